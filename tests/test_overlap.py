"""A step's lightning beside the next step's dynamics (ml_ctx_set_overlap): runs stepped with the overlap, in one call
or in several, end bit for bit as the same runs stepped with each step's parts in line: fields, ground columns, run
metrics, flashes, rotation and model diagnostics, through steps that flash."""
import numpy as np
import pytest

import meteolab_amd as M
from meteolab_amd import config as C
from oracle.binding import OracleRun

pytestmark = pytest.mark.gpu


def charged(g, run, dry, spec):
    """A charged cloud above the breakdown threshold (the state test's, state.test.ts:25-52): it flashes at once."""
    nx, ny, nz = g["nx"], g["ny"], g["nz"]
    dx, dy = g["width"] / nx, g["depth"] / ny
    rho = run.levels()["rho"]
    zlev, _ = OracleRun(g, dry, spec).levels()
    x, y, z = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="xy")
    r2 = ((x - nx / 2) * dx) ** 2 + ((y - ny / 2) * dy) ** 2
    ci = 3e-9 * np.exp(-r2 / 2500 ** 2 - ((zlev[z] - 9000) / 700) ** 2) / rho[z]
    cg = -3e-9 * np.exp(-r2 / 2500 ** 2 - ((zlev[z] - 6500) / 700) ** 2) / rho[z]
    run.set_field("chargeIce", np.transpose(ci, (2, 0, 1)).reshape(-1))
    run.set_field("chargeGraupel", np.transpose(cg, (2, 0, 1)).reshape(-1))


def setup(ctx, storm):
    g = C.small_grid()
    pool = ctx.pool(g, 3)
    dry = C.profile(qv_max=0.0)
    spec = C.run(g, seed=7, amp=0.0, noise_amp=0.0)
    a = pool.run(dry, spec)
    charged(g, a, dry, spec)
    b = pool.run(C.profile(), C.run(g, seed=1))
    for name in M.CARRIED + ["fieldZ"]:
        b.set_field(name, storm.field(name))
    for name in M.COLUMNS:
        b.set_column(name, storm.column(name))
    b.set_time(storm.diag()["time"])
    c = pool.run(dry, dict(spec, seed=11))
    charged(g, c, dry, spec)
    return [a, b, c]


def state(r):
    return ({k: r.field(k) for k in M.CARRIED + ["fieldZ"]}, {k: r.column(k) for k in M.COLUMNS}, r.diagnostics(),
            r.flashes(), r.rotation(), r.model_diagnostics())


def same(s, t):
    for name in s[0]:
        assert np.array_equal(s[0][name], t[0][name]), name
    for name in s[1]:
        assert np.array_equal(s[1][name], t[1][name]), name
    assert s[2:] == t[2:]


@pytest.fixture(scope="module")
def storm():
    g = C.small_grid()
    o = OracleRun(g, C.profile(), C.run(g, seed=1))
    o.step(650)
    return o


def test_the_overlapped_steps_end_as_the_steps_in_line(storm):
    # One context, a pool of the same three runs for each way of stepping them.
    results = []
    with M.Context(0, 32) as ctx:
        # In line in one call and in one-step calls, overlapped in one call, in calls of 1, 2, 3 and in one-step calls
        # (each a step, then its discharge after the step's end).
        for overlap, calls in ((False, [6]), (False, [1] * 6), (True, [6]), (True, [1, 2, 3]), (True, [1] * 6)):
            ctx.set_overlap(overlap)
            runs = setup(ctx, storm)
            print("overlap", overlap, "calls", calls, flush=True)
            for n in calls:
                M.step(runs, n)
            ctx.sync()
            results.append([state(r) for r in runs])
            assert results[-1][0][2]["flashesIC"] + results[-1][0][2]["flashesCG"] > 0
            runs[0].pool.destroy()
    for k in range(1, len(results)):
        for a, b in zip(results[0], results[k]):
            same(a, b)


def test_the_charged_cloud_flashes_as_the_oracle():
    # The charged cloud these tests step, in clear air (no cloud anywhere): its first two steps flash as the oracle's
    # (9, then 17 flashes in all) and leave the charges within the parity bar. The cloud is symmetric about the domain's
    # centre, so from the third step on the choice between mirror-image nodes of equal field is left to rounding and the
    # flash sequences part; the counts of the first steps are the check.
    g = C.small_grid()
    dry = C.profile(qv_max=0.0)
    spec = C.run(g, seed=7, amp=0.0, noise_amp=0.0)
    o = OracleRun(g, dry, spec)
    with M.Context(0, 32) as ctx:
        r = ctx.pool(g, 1).run(dry, spec)
        charged(g, r, dry, spec)
        for name in ("chargeIce", "chargeGraupel"):
            o.set_field(name, r.field(name))
        for step in range(2):
            r.step(1)
            o.step(1)
            d, od = r.diagnostics(), o.diag()
            print("step", step + 1, "flashes", d["flashesIC"], d["flashesCG"], "oracle", od["flashesIC"], od["flashesCG"], flush=True)
            assert (d["flashesIC"], d["flashesCG"]) == (od["flashesIC"], od["flashesCG"])
        for name in ("chargeIce", "chargeGraupel", "chargeRain", "chargeAir"):
            ref = o.field(name).astype(np.float64)
            diff = np.linalg.norm(r.field(name).astype(np.float64) - ref) / max(np.linalg.norm(ref), 1e-300)
            assert diff <= 1e-5, (name, diff)


def test_a_flashing_step_repeats_exactly():
    """The two charged clear-air clouds made anew eight times, each time in a new pool, and stepped once (9 flashes
    each): every repetition ends as the first, bit for bit. The discharge's result must not depend on how its waves
    are scheduled: its failure marks, front and potential pass between waves through LDS and global memory."""
    g = C.small_grid()
    dry = C.profile(qv_max=0.0)
    spec = C.run(g, seed=7, amp=0.0, noise_amp=0.0)
    outs = []
    with M.Context(0, 32) as ctx:
        for rep in range(8):
            print("repetition", rep, flush=True)
            pool = ctx.pool(g, 2)
            runs = [pool.run(dry, spec), pool.run(dry, dict(spec, seed=11))]
            for r in runs:
                charged(g, r, dry, spec)
            M.step(runs, 1)
            ctx.sync()
            outs.append([state(r) for r in runs])
            pool.destroy()
    assert outs[0][0][2]["flashesIC"] + outs[0][0][2]["flashesCG"] > 1
    for k in range(1, len(outs)):
        for a, b in zip(outs[0], outs[k]):
            same(a, b)
