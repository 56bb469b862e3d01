"""Interventions in a run going on (intervene.ts:38-68): the oracle held to the reference's own checks
(intervene.test.ts:9-68), and the device's ml_run_intervene to the oracle.

The reference's tests run on SUMMER_DAY; these run on the Weisman-Klemp sounding of the step's other tests. Their checks
do not depend on the sounding: the patch formula, what it adds at its centre, the saturation cap and the vortex's shape.
"""
import numpy as np
import pytest

from meteolab_amd import config as C
from oracle.binding import INTERVENTIONS, OracleRun

G = C.STANDARD


def fresh(g=G, seed=7):
    return OracleRun(g, C.profile(), C.run(g, seed=seed))


def node(g, x, y, z=0):
    """intervene.test.ts:7: the node nearest (x, y) m at level z (Math.round: halves up)."""
    dx, dy = g["width"] / g["nx"], g["depth"] / g["ny"]
    return int(np.floor(x / dx + .5)) + g["nx"] * (int(np.floor(y / dy + .5)) + g["ny"] * z)


def ulps32(a, b):
    """Distance in float32 units in the last place."""
    a, b = np.asarray(a, np.float32), np.asarray(b, np.float32)
    return np.abs(a.view(np.int32).astype(np.int64) - b.view(np.int32).astype(np.int64))


def test_the_thermal_is_the_formula_injectbubble_always_used():
    # intervene.test.ts:10-23: +3.2 K and +1.1 m/s times exp(-(r/4.2 km)^2 - (z/1.8 km)^2) below 2.7 km, periodic.
    m = fresh()
    theta, w = m.field("theta"), m.field("w")
    zs, _ = m.levels()
    nx, ny, nz = G["nx"], G["ny"], G["nz"]
    dx, dy, W, D = G["width"] / nx, G["depth"] / ny, G["width"], G["depth"]
    m.intervene("thermal", 20_000, 15_000, 1.25)
    want_t, want_w = theta.astype(np.float64).reshape(nz, ny, nx), w.astype(np.float64).reshape(nz, ny, nx)
    x, y = np.meshgrid(np.arange(nx) * dx - 20_000, np.arange(ny) * dy - 15_000)
    x = np.where(x > W / 2, x - W, x); x = np.where(x < -W / 2, x + W, x)
    y = np.where(y > D / 2, y - D, y); y = np.where(y < -D / 2, y + D, y)
    for z in range(nz):
        if zs[z] >= 2700:
            break
        a = np.exp(-((x / 4200) ** 2 + (y / 4200) ** 2 + (zs[z] / 1800) ** 2)) * 1.25
        want_t[z] += 3.2 * a
        want_w[z] += 1.1 * a
    # The reference asserts equality; numpy's exp may differ from the C library's in the last bit of a double, which
    # moves a float32 result by at most one unit in the last place.
    assert ulps32(m.field("theta"), want_t.ravel()).max() <= 1
    assert ulps32(m.field("w"), want_w.ravel()).max() <= 1
    assert np.mean(ulps32(m.field("theta"), want_t.ravel()) == 0) > .999
    assert not np.array_equal(m.field("theta"), theta)


def test_what_is_put_in_is_counted():
    # intervene.test.ts:25-38: warming puts heat in, +2 K at its centre; cooling takes the same out; a cold pool is
    # -6 K on its centre node.
    m, before = fresh(), fresh()
    assert m.added() == {"heat": 0.0, "vapour": 0.0, "count": 0}
    i = node(G, 20_000, 15_000)
    warm = m.intervene("warm", 20_000, 15_000, 1)
    assert warm["heat"] > 0
    assert abs(m.field("theta")[i] - before.field("theta")[i] - 2) < .05
    cool = m.intervene("cool", 20_000, 15_000, 1)
    assert abs(cool["heat"] + warm["heat"]) < 5e5
    assert m.added()["count"] == 2 and abs(m.added()["heat"]) < warm["heat"] * 1e-3
    dx, dy = G["width"] / G["nx"], G["depth"] / G["ny"]
    cx, cy = 8 * dx, 9 * dy
    j = node(G, cx, cy)
    t0 = m.field("theta")[j]
    m.intervene("coldPool", cx, cy, 1)
    assert abs(m.field("theta")[j] - t0 + 6) < .05


def test_moistening_stops_short_of_saturation_and_drying_at_nothing():
    # intervene.test.ts:40-48.
    m = fresh()
    i = node(G, 20_000, 15_000)
    for _ in range(20):
        m.intervene("moisten", 20_000, 15_000, 2)
    bs = m.base_state()
    t = m.field("theta")[i] * bs["exner"][0] - 273.15
    es = 611.2 * np.exp(17.67 * t / (t + 243.5))
    qs = es * .622 / (bs["p"][0] - es)
    assert abs(m.field("q")[i] / qs - .95) < .005
    assert m.added()["vapour"] > 0
    for _ in range(20):
        m.intervene("dry", 20_000, 15_000, 2)
    assert m.field("q")[i] == 0 and m.added()["vapour"] < 0


def test_the_vortex_spins_cyclonically_without_converging_the_air():
    # intervene.test.ts:50-67: > 7 m/s round its radius (3 km), the added wind's divergence under a tenth of its
    # vorticity.
    m = fresh()
    u, v = m.field("u"), m.field("v")
    m.intervene("vortex", 24_000, 18_000, 1)
    east, north = node(G, 27_000, 18_000), node(G, 24_000, 21_000)
    du, dv = m.field("u").astype(np.float64) - u, m.field("v").astype(np.float64) - v
    assert dv[east] > 7 and du[north] < -7
    nx, ny = G["nx"], G["ny"]
    dx, dy = G["width"] / nx, G["depth"] / ny
    a, b = du[:nx * ny].reshape(ny, nx), dv[:nx * ny].reshape(ny, nx)
    div = (a[1:-1, 2:] - a[1:-1, :-2]) / (2 * dx) + (b[2:, 1:-1] - b[:-2, 1:-1]) / (2 * dy)
    vort = (b[1:-1, 2:] - b[1:-1, :-2]) / (2 * dx) - (a[2:, 1:-1] - a[:-2, 1:-1]) / (2 * dy)
    assert np.abs(div).max() < .1 * np.abs(vort).max()


def test_the_heat_is_the_energy_of_the_warming():
    # What intervene counts is c_p exner dtheta times each node's air mass (rho dx dy hz), summed: recomputed here from
    # the fields before and after, which fixes the mass and the exner factor of the count.
    m, before = fresh(), fresh()
    out = m.intervene("warm", 10_000, 30_000, 1.7)
    bs = m.base_state()
    zs, hz = m.levels()
    nx, ny, nz = G["nx"], G["ny"], G["nz"]
    dth = (m.field("theta").astype(np.float64) - before.field("theta")).reshape(nz, ny * nx)
    mass = bs["rho"] * (G["width"] / nx) * (G["depth"] / ny) * hz
    heat = float(np.sum(1004 * bs["exner"][:, None] * dth * mass[:, None]))
    assert heat == pytest.approx(out["heat"], rel=1e-5)
    assert m.added() == {"heat": out["heat"], "vapour": 0.0, "count": 1}


def test_a_patch_at_the_corner_wraps_round_the_periodic_domain():
    # Centred on the corner node: the nodes one column east and one west (across the edge), one row north and one
    # south, are the same distance away the shortest way round, so they warm alike, as much as at 1200 m inside.
    m, before = fresh(), fresh()
    m.intervene("warm", 0.0, 0.0, 1)
    nx, ny = G["nx"], G["ny"]
    d = m.field("theta").astype(np.float64)[:nx * ny].reshape(ny, nx) - before.field("theta")[:nx * ny].reshape(ny, nx)
    assert d[0, 0] > 1.9
    assert d[0, 1] == d[0, nx - 1] and d[1, 0] == d[ny - 1, 0] and d[1, 1] == d[ny - 1, nx - 1] == d[1, nx - 1]
    assert d[0, nx - 1] > 1.8 and d[0, nx // 2] < 1e-6


# ---- the device's interventions -----------------------------------------------------------------------------

@pytest.mark.gpu
@pytest.mark.parametrize("precision", [32, 64])
def test_every_intervention_on_the_device_is_the_oracles(precision):
    # Each kind at a place near an edge (so that the patch wraps round the periodic domain), in turn on the same run,
    # with the thermal's radius of a config that sets bubbleRadius; then a step of the run.
    import meteolab_amd as M
    from test_gpu_step import compare, margin
    g = C.STANDARD
    o = fresh(g)
    with M.Context(0, precision) as ctx:
        r = ctx.pool(g, 1).run(C.profile(), C.run(g, seed=7))
        print("interventions, float", precision, flush=True)
        places = {"thermal": (2_000, 34_000, 1.3, 5.0), "warm": (47_000, 1_000, .8, 0), "cool": (24_000, 18_000, 1.1, 0),
                  "moisten": (40_000, 30_000, 2, 0), "dry": (5_000, 5_000, 1.5, 0), "coldPool": (46_500, 17_000, 1, 0),
                  "vortex": (1_000, 35_500, 1.2, 0)}
        for kind in INTERVENTIONS:
            x, y, s, radius = places[kind]
            got, want = r.intervene(kind, x, y, s, radius), o.intervene(kind, x, y, s, radius)
            for k in ("heat", "vapour"):
                assert got[k] == pytest.approx(want[k], rel=1e-12, abs=1e-9), (kind, k, got[k], want[k])
            for name in ("u", "v", "w", "theta", "q"):
                d = ulps32(r.field(name), o.field(name))
                assert d.max() <= 1, (kind, name, int(d.max()))
        a, b = r.added(), o.added()
        assert a["count"] == b["count"] == 7
        assert a["heat"] == pytest.approx(b["heat"], rel=1e-12) and a["vapour"] == pytest.approx(b["vapour"], rel=1e-12)
        r.step(1); o.step(1)
        print("a step after the interventions:", margin(compare(r, o)))


@pytest.mark.gpu
def test_the_counts_survive_a_saved_state():
    import meteolab_amd as M
    from meteolab_amd.state import apply_state, read_state, save_state
    g = C.small_grid()
    with M.Context(0, 32) as ctx:
        r = ctx.pool(g, 1).run(C.profile(), C.run(g))
        r.intervene("warm", 5_000, 6_000, 1)
        r.intervene("moisten", 9_000, 4_000, 1)
        state = read_state(save_state(r))
        assert state[0]["added"] == r.added() and state[0]["added"]["count"] == 2
        q = ctx.pool(g, 1).run(C.profile(), C.run(g))
        assert q.added() == {"heat": 0.0, "vapour": 0.0, "count": 0}
        apply_state(q, state)
        assert q.added() == r.added()
        for name in M.CARRIED:
            assert np.array_equal(q.field(name), r.field(name)), name


@pytest.mark.gpu
def test_intervene_refuses_an_unknown_kind_and_a_bad_strength():
    import meteolab_amd as M
    g = C.small_grid()
    with M.Context(0, 32) as ctx:
        r = ctx.pool(g, 1).run(C.profile(), C.run(g))
        with pytest.raises(ValueError):
            r.intervene("hurricane", 0, 0, 1)
        with pytest.raises(M.MeteolabError):
            r.intervene("warm", 0, 0, float("nan"))
        assert r.added()["count"] == 0
