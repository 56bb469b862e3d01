"""The rotation diagnostics (updateRotation, src/core/model.ts:1609-1648; refreshRotation :1607), on the CPU: the
oracle on planted vortices whose helicity, core vorticity and positions have closed forms; on the GPU: the device's
k_rotation on the same planted fields, and after a step of a mature storm against the oracle's step.

Planted fields (field index x + nx y + nz-level z * nx ny, as the reference's): a solid-body vortex u = -om (y - yc) dy,
v = om (x - xc) dx over a window one column wider than a patch has central-difference vorticity dv/dx - du/dy = 2 om
exactly on the patch; with w = W on every level above the ground the 2-5 km helicity of a patch column is
W 2 om S, S the sum of hz over the levels from the first at or above 2 km to the last at or below 5 km, and the
core's w-hz-weighted mean vorticity is 2 om. A patch across the periodic x edge with a vorticity that differs between
its two columns tells one core (the mean) from two (the larger)."""
import numpy as np
import pytest

from meteolab_amd import config as C
from oracle.binding import OracleRun

REL = 2e-6  # the planted winds are float32: the vorticity of a node is 2 om to float32 rounding


def planted(g, scenario):
    """u, v, w (float32, nx*ny*nz) of a scenario and its expected rotation record's closed forms."""
    nx, ny, nz = g["nx"], g["ny"], g["nz"]
    dx, dy = g["width"] / nx, g["depth"] / ny
    u, v, w = (np.zeros((nz, ny, nx), np.float64) for _ in range(3))

    def vortex(xs, ys, om, weights):
        """Solid-body rotation om about the patch centre on the patch plus one column around it; w by column."""
        xc, yc = (xs[0] + xs[-1]) / 2, (ys[0] + ys[-1]) / 2
        for y in range(ys[0] - 1, ys[-1] + 2):
            for x in range(xs[0] - 1, xs[-1] + 2):
                u[:, y % ny, x % nx] = -om * (y - yc) * dy
                v[:, y % ny, x % nx] = om * (x - xc) * dx
        for (x, y), wv in weights.items():
            w[1:, y, x] = wv

    if scenario == "pair":
        # Cyclonic patch A (3x3 at x 2-4, y 2-4; its centre column the strongest), anticyclonic patch B (x 9-11).
        vortex([2, 3, 4], [2, 3, 4], 0.01, {(x, y): 25.0 if (x, y) == (3, 3) else 20.0 for x in (2, 3, 4) for y in (2, 3, 4)})
        vortex([9, 10, 11], [2, 3, 4], -0.008, {(x, y): 30.0 if (x, y) == (10, 3) else 20.0 for x in (9, 10, 11) for y in (2, 3, 4)})
    else:
        # One patch of two columns across the x edge (x = nx-1 and 0, y 7-9): u = -om (y - 8) dy gives om; v = c off^2 dx,
        # off the signed distance from x = 0, gives 2 c off: vorticity om - 2c at x = nx-1, om at x = 0.
        om, c = 0.005, 0.001
        for y in range(6, 11):
            for x in (nx - 1, 0):
                u[:, y, x] = -om * (y - 8) * dy
        for y in range(7, 10):
            for off in (-2, -1, 0, 1):
                v[:, y, off % nx] = c * off * off * dx
            for x in (nx - 1, 0):
                w[1:, y, x] = 15.0
    return [a.astype(np.float32).reshape(-1) for a in (u, v, w)]


def expected(g, zs, hz, scenario):
    nx = g["nx"]
    z0 = int(np.argmax(zs >= 2000))
    z1 = int(np.nonzero(zs <= 5000)[0][-1])
    top1 = int(np.nonzero(zs <= 1000)[0][-1])
    top3 = int(np.nonzero(zs <= 3000)[0][-1])
    S, S1, S3 = hz[z0:z1 + 1].sum(), hz[1:top1 + 1].sum(), hz[1:top3 + 1].sum()
    if scenario == "pair":
        return {"uh": 25 * 0.02 * S, "x": 3, "y": 3, "anticyclonic": 30 * 0.016 * S, "uh01": 25 * 0.02 * S1,
                "uh03": 25 * 0.02 * S3, "low": 3 + nx * 3, "coreZeta": 0.02, "coreUH": 25 * 0.02 * S, "coreAnti": 0.016}
    # One core across the edge: the mean of 0.003 and 0.005 over equal weights; its peak helicity the x = 0 column's.
    return {"uh": 15 * 0.005 * S, "x": 0, "y": 7, "anticyclonic": 0.0, "uh01": 15 * 0.005 * S1,
            "uh03": 15 * 0.005 * S3, "low": nx * 7, "coreZeta": 0.004, "coreUH": 15 * 0.005 * S, "coreAnti": 0.0}


def check(got, want):
    for k, e in want.items():
        if k in ("x", "y", "low"):
            assert got[k] == e, (k, got[k], e)
        else:
            assert got[k] == pytest.approx(e, rel=REL, abs=1e-12), (k, got[k], e)


def oracle_with(g, scenario):
    o = OracleRun(g, C.profile(), C.run(g, seed=1))
    for name, a in zip(("u", "v", "w"), planted(g, scenario)):
        o.set_field(name, a)
    return o


@pytest.mark.parametrize("scenario", ["pair", "edge"])
def test_oracle_rotation_of_planted_vortices(scenario):
    g = C.small_grid()
    o = oracle_with(g, scenario)
    zs, hz = o.levels()
    o.update_rotation(3.0, keep=False)
    check(o.rotation(), expected(g, zs, hz, scenario))


DEFAULT = {"nx": 40, "ny": 32, "nz": 50, "width": 48000.0, "depth": 36000.0, "height": 18900.0, "bottom": 100.0}


def rankine(g, zs, vortices):
    """model.test.ts:36-49: a broad updraught (25 m/s peak, 5 km e-folding) at 1.5-6 km about column (20, 16) and
    Rankine vortices (vorticity zeta inside 2 km) at the given column offsets."""
    nx, ny, nz = g["nx"], g["ny"], g["nz"]
    dx, dy, cx, cy = g["width"] / nx, g["depth"] / ny, 20, 16
    u, v, w = (np.zeros((nz, ny, nx)) for _ in range(3))
    ex = (np.arange(nx) - cx) * dx
    ey = (np.arange(ny) - cy) * dy
    EX, EY = np.meshgrid(ex, ey)
    for z in range(nz):
        if zs[z] < 1500 or zs[z] > 6000:
            continue
        w[z] = 25 * np.exp(-(EX * EX + EY * EY) / 5000 ** 2)
        for ox, zeta in vortices:
            rx = EX - ox * dx
            r = np.hypot(rx, EY)
            with np.errstate(divide="ignore", invalid="ignore"):
                speed = zeta / 2 * np.where(r < 2000, r, 2000 ** 2 / r)
                u[z] += np.where(r > 0, -speed * EY / r, 0)
                v[z] += np.where(r > 0, speed * rx / r, 0)
    return [a.astype(np.float32).reshape(-1) for a in (u, v, w)]


def rotation_of(o, g, vortices):
    zs, _ = o.levels()
    for name, a in zip(("u", "v", "w"), rankine(g, zs, vortices)):
        o.set_field(name, a)
    o.update_rotation(1.0, keep=False)
    return o.rotation()


def test_oracle_tells_a_mesocyclone_from_a_vortex_pair():
    # model.test.ts:33-56: a rotating updraught has core vorticity above 0.004 s-1; an updraught with a cyclonic and an
    # anticyclonic vortex on its flanks has strong helicity of both signs but core vorticity under 0.002 s-1.
    g = DEFAULT
    o = OracleRun(g, C.profile(), C.run(g, seed=1))
    meso = rotation_of(o, g, [(0, .02)])
    pair = rotation_of(o, g, [(-3, .02), (3, -.02)])
    assert meso["coreZeta"] > .004
    assert pair["uh"] > 200 and pair["anticyclonic"] > 200
    assert abs(pair["coreZeta"]) < .002


def test_oracle_rotation_holds_while_the_core_spins():
    # persisted adds dt while the core's vorticity is at least mesoCoreZeta(dx) (3e-3 at 1200 m) and resets otherwise;
    # refreshRotation keeps it.
    g = C.small_grid()
    o = oracle_with(g, "pair")
    o.update_rotation(3.0, keep=False)
    o.update_rotation(3.0, keep=False)
    assert o.rotation()["persisted"] == 6.0
    o.update_rotation(0.0, keep=True)
    assert o.rotation()["persisted"] == 6.0
    for name in ("u", "v"):
        o.set_field(name, np.zeros(g["nx"] * g["ny"] * g["nz"], np.float32))
    o.update_rotation(0.0, keep=True)
    r = o.rotation()
    assert r["persisted"] == 6.0 and r["coreZeta"] == 0 and r["uh"] == 0
    o.update_rotation(3.0, keep=False)
    assert o.rotation()["persisted"] == 0.0


def test_oracle_rotation_is_the_last_step_before_the_lightning():
    # The record after a step is updateRotation of the step's winds (the lightning changes no wind).
    g = C.small_grid()
    o = OracleRun(g, C.profile(), C.run(g, seed=1))
    o.step(460)
    after_step = o.rotation()
    o.update_rotation(0.0, keep=True)
    assert o.rotation() == after_step
    assert after_step["uh"] > 1 and after_step["uh03"] > 1


@pytest.mark.gpu
@pytest.mark.parametrize("scenario", ["pair", "edge"])
def test_device_rotation_of_planted_vortices(scenario):
    import meteolab_amd as M
    g = C.small_grid()
    o = oracle_with(g, scenario)
    zs, hz = o.levels()
    o.update_rotation(0.0, keep=True)
    with M.Context(0, 32) as ctx:
        r = ctx.pool(g, 1).run(C.profile(), C.run(g, seed=1))
        for name, a in zip(("u", "v", "w"), planted(g, scenario)):
            r.set_field(name, a)
        r.set_totals(0.0, 0, 0, 9.0)
        print("device rotation of planted vortices:", scenario, flush=True)
        r.refresh_rotation()
        got = r.rotation()
    check(got, expected(g, zs, hz, scenario))
    want = o.rotation()
    assert got["persisted"] == 9.0
    for k in want:
        if k != "persisted":
            assert got[k] == pytest.approx(want[k], rel=1e-12, abs=1e-15), k


@pytest.mark.gpu
def test_device_tells_a_mesocyclone_from_a_vortex_pair():
    import meteolab_amd as M
    g = DEFAULT
    o = OracleRun(g, C.profile(), C.run(g, seed=1))
    zs, _ = o.levels()
    with M.Context(0, 32) as ctx:
        r = ctx.pool(g, 1).run(C.profile(), C.run(g, seed=1))
        for vortices in ([(0, .02)], [(-3, .02), (3, -.02)]):
            for name, a in zip(("u", "v", "w"), rankine(g, zs, vortices)):
                r.set_field(name, a)
            print("device rotation of the vortices", vortices, flush=True)
            r.refresh_rotation()
            got, want = r.rotation(), rotation_of(o, g, vortices)
            for k in want:
                if k != "persisted":
                    assert got[k] == pytest.approx(want[k], rel=1e-12, abs=1e-15), (k, got[k], want[k])
            if len(vortices) == 1:
                assert got["coreZeta"] > .004
            else:
                assert got["uh"] > 200 and got["anticyclonic"] > 200 and abs(got["coreZeta"]) < .002


@pytest.mark.gpu
@pytest.mark.parametrize("steps", [449, 650])
def test_device_rotation_after_a_step_of_a_mature_storm(steps):
    # At 449 steps (22.5 min) the next step has a cyclonic updraught core; at 650 an anticyclonic one.
    import meteolab_amd as M
    g = C.small_grid()
    o = OracleRun(g, C.profile(), C.run(g, seed=1))
    o.step(steps)
    with M.Context(0, 32) as ctx:
        r = ctx.pool(g, 1).run(C.profile(), C.run(g, seed=1))
        for name in M.CARRIED + ["fieldZ"]:
            r.set_field(name, o.field(name))
        for name in M.COLUMNS:
            r.set_column(name, o.column(name))
        r.set_time(o.diag()["time"])
        print("device rotation after a step from", steps, flush=True)
        o.step(1)
        r.step(1)
        got, want = r.rotation(), o.rotation()
    assert want["coreZeta"] > 0 or want["coreAnti"] > 0
    for k in ("x", "y", "low", "persisted"):
        assert got[k] == want[k], (k, got[k], want[k])
    for k in ("uh", "anticyclonic", "uh01", "uh03", "coreZeta", "coreUH", "coreAnti"):
        assert got[k] == pytest.approx(want[k], rel=1e-4, abs=1e-9), (k, got[k], want[k])
