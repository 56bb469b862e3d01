"""Parity of the HIP step (through the C ABI) with the oracle, the CPU restatement of the reference's step
(AtmosphereModel.step, src/core/model.ts:873-877). Per-field relative L2 ||gpu - oracle|| / ||oracle|| <= 1e-5 after one
and after ten steps; fields the oracle holds at exactly zero must stay within 1e-12 of zero."""
import numpy as np
import pytest

import meteolab_amd as M
from meteolab_amd import config as C
from oracle.binding import OracleRun

pytestmark = pytest.mark.gpu

TOL = 1e-5


def margin(worst, tol=TOL):
    """The worst per-field relative L2, its field and the margin to the bound, printed by each parity test."""
    k = max(worst, key=worst.get)
    return f"worst {worst[k]:.2e} ({k}), bound {tol:g}, margin {tol / worst[k]:.1f}x"


@pytest.fixture(scope="module")
def ctx():
    c = M.Context(0, 32)
    print("device:", c.adapter, flush=True)
    yield c
    c.close()


def rel_l2(a, b):
    a, b = a.astype(np.float64), b.astype(np.float64)
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / nb if nb > 0 else np.linalg.norm(a)


def compare(run, oracle, tol=TOL, fields=M.CARRIED, zero_tol=1e-12):
    worst = {}
    bad_nodes = {name: int((~np.isfinite(run.field(name))).sum()) for name in fields}
    assert not any(bad_nodes.values()), f"non-finite values: {bad_nodes}"
    for name in fields:
        a, b = run.field(name), oracle.field(name)
        if np.linalg.norm(b.astype(np.float64)) == 0:
            assert np.linalg.norm(a.astype(np.float64)) <= zero_tol, (name, np.linalg.norm(a))
            continue
        worst[name] = rel_l2(a, b)
    bad = {k: v for k, v in worst.items() if v > tol}
    assert not bad, f"fields over {tol}: {bad} (all: {worst})"
    return worst


def test_base_state_and_initial_state_match(ctx):
    g = C.STANDARD
    prof, run = C.profile(), C.run(g)
    o = OracleRun(g, prof, run)
    r = ctx.pool(g, 1).run(prof, run)
    lv, bs = r.levels(), o.base_state()
    for k in ("p", "exner", "theta", "q", "u", "v", "rho"):
        scale = max(np.abs(bs[k]).max(), 1e-300)
        assert np.abs(lv[k] - bs[k]).max() / scale < 1e-13, k
    for name in M.CARRIED:
        assert np.array_equal(r.field(name), o.field(name)), name


@pytest.mark.parametrize("grid", ["small", "standard", "large"])
def test_parity_after_one_and_ten_steps(ctx, grid):
    # The three sizes of the workloads: the small test grid, the standard (40x32x50) domain the 1e-5 bar is set on, and
    # the large (80x64x50) one, each held to 1e-5 per field. With the fluxes taken relative to their node, w's relative
    # difference after ten steps is 1.0e-6 on the small grid, 2.8e-6 on the standard and 5.9e-6 on the large.
    g = C.small_grid() if grid == "small" else C.STANDARD if grid == "standard" else C.LARGE
    prof, run = C.profile(), C.run(g)
    o = OracleRun(g, prof, run)
    r = ctx.pool(g, 1).run(prof, run)
    o.step(1); r.step(1)
    w1 = compare(r, o)
    o.step(9); r.step(9)
    w10 = compare(r, o)
    print(grid, "after 1 step:", margin(w1), "| after 10 steps:", margin(w10))
    d, od = r.diagnostics(), o.diag()
    assert d["time"] == od["time"] == 30.0
    assert d["maxW"] == pytest.approx(od["maxW"], rel=1e-5)


def test_large_domain_parity_of_the_float64_kernels():
    # The same steps of the large domain with the float64 kernels (Context precision 64): w stays within 1e-5 after ten
    # steps too (3.5e-6), the float32 rounding of the production kernels aside.
    g = C.LARGE
    prof, run = C.profile(), C.run(g)
    o = OracleRun(g, prof, run)
    with M.Context(0, 64) as c64:
        print("large domain, float64 kernels", flush=True)
        r = c64.pool(g, 1).run(prof, run)
        o.step(1); r.step(1)
        w1 = compare(r, o)
        o.step(9); r.step(9)
        w10 = compare(r, o)
    print("large float64: 1 step", max(w1.values()), "10 steps", max(w10.values()))


def mature(g, steps, seed=1):
    """The oracle run `steps` steps into the storm (cloud, rain and ice present), for a restart on the GPU."""
    o = OracleRun(g, C.profile(), C.run(g, seed=seed))
    o.step(steps)
    return o


def restart(ctx, o, g, seed=1):
    r = ctx.pool(g, 1).run(C.profile(), C.run(g, seed=seed))
    for name in M.CARRIED + ["fieldZ"]:
        r.set_field(name, o.field(name))
    for name in M.COLUMNS:
        r.set_column(name, o.column(name))
    r.set_time(o.diag()["time"])
    return r


def test_parity_in_a_mature_storm(ctx):
    # 32.5 min into the storm on the small grid: cloud, rain, snow, graupel, hail, ice and charge are all present.
    g = C.small_grid()
    o = mature(g, 650)
    st = o.state()
    for name in ("cloud", "rain", "ice", "snow", "graupel", "hail", "chargeIce"):
        assert np.abs(st[name]).max() > 0, name
    r = restart(ctx, o, g)
    o.step(1); r.step(1)
    w1 = compare(r, o)
    o.step(9); r.step(9)
    w10 = compare(r, o)
    print("mature storm after 1 step:", margin(w1), "| after 10 steps:", margin(w10))
    assert r.column("precipitation").max() == pytest.approx(o.column("precipitation").max(), rel=1e-4)


def test_batch_alone_equals_in_the_mix(ctx):
    # Three different runs stepped together give what each gives alone, bit for bit (the serving loop's batching
    # invariant: a run's arithmetic and its sums' order do not depend on the batch).
    g = C.small_grid()
    specs = [C.run(g, seed=s, amp=a, offset=off) for s, a, off in ((1, 2.0, (0, 0)), (2, 2.6, (3000, -2000)), (3, 1.6, (-4000, 1000)))]
    pool = ctx.pool(g, 4)
    mixed = [pool.run(C.profile(), s) for s in specs]
    M.step(mixed, 20)
    ctx.sync()
    for s, rm in zip(specs, mixed):
        alone = ctx.pool(g, 1).run(C.profile(), s)
        alone.step(20)
        for name in M.CARRIED:
            assert np.array_equal(rm.field(name), alone.field(name)), name


def test_float64_kernels_match_the_oracle_closely():
    # The same kernels in float64: what is left is the float32 storage of the fields, whose last-bit rounding depends on
    # the order of operations (the oracle's and the kernels' differ), fed back through the projections: ~1e-6 in w.
    g = C.small_grid()
    prof, run = C.profile(), C.run(g)
    o = OracleRun(g, prof, run)
    with M.Context(0, 64) as c64:
        r = c64.pool(g, 1).run(prof, run)
        o.step(10); r.step(10)
        w = compare(r, o, tol=3e-6)
        print("float64 kernels, 10 steps", max(w.values()))


def test_graph_replay_equals_eager_launches(ctx):
    g = C.small_grid()
    pool = ctx.pool(g, 2)
    a = pool.run(C.profile(), C.run(g))
    b = pool.run(C.profile(), C.run(g))
    a.step(4)                      # captured graphs, replayed (the lightning beside the next step)
    M.profile_step([b], 4)         # the same kernels launched one by one, in line
    for name in M.CARRIED:
        assert np.array_equal(a.field(name), b.field(name)), name


@pytest.fixture(scope="module")
def mature_small():
    return mature(C.small_grid(), 650)


MARCH_OFF = pytest.mark.skip(reason="the column-march transport layout was taken out after a whole-step run with it faulted "
                                   "the GPU, cause not found on the CPU; set_batch_tuning rejects it")
TRANSPORT_SHAPES = [("level", 1, False, 1),
                    ("level", 1, True, 1), ("level", 1, True, 5), ("level", 1, False, 10), ("plane", 1, False, 1),
                    ("plane", 1, True, 2), ("plane", 1, True, 5), ("plane", 1, False, 10),
                    pytest.param("march", 1, True, 5, marks=MARCH_OFF), pytest.param("march", 1, False, 24, marks=MARCH_OFF),
                    pytest.param("march", 1, True, 10, marks=MARCH_OFF)]


@pytest.mark.parametrize("layout,groups,xcd,levels", TRANSPORT_SHAPES)
def test_every_transport_layout_keeps_parity(ctx, mature_small, layout, groups, xcd, levels):
    # Each launch shape the autotuner may pick: one step of the mature storm against the oracle, eight different runs
    # batched in that shape (the XCD block order needs a multiple of eight) against each run stepped alone in the same
    # layout (the same kernel: batching changes nothing but the order of the positivity fix's float64 sums), and against
    # each run stepped alone in the level layout (another kernel: float32 rounding, fused differently, is all).
    g = C.small_grid()
    ref = oracle_restart(mature_small, g)
    r = restart(ctx, mature_small, g)
    M.set_batch_tuning([r], layout, groups, xcd, levels)
    ref.step(1); r.step(1)
    compare(r, ref)
    specs = [C.run(g, seed=s, amp=1.5 + .2 * s, offset=(500 * s - 2000, 300 * s - 1000)) for s in range(8)]
    pool = ctx.pool(g, 8)
    mixed = [pool.run(C.profile(), s) for s in specs]
    M.set_batch_tuning(mixed, layout, groups, xcd, levels)
    assert M.batch_tuning(mixed)["transport_layout"] == layout
    M.step(mixed, 10)
    ctx.sync()
    for s, rm in zip(specs, mixed):
        alone, other = ctx.pool(g, 1).run(C.profile(), s), ctx.pool(g, 1).run(C.profile(), s)
        M.set_batch_tuning([alone], layout, groups, xcd, levels)
        M.set_batch_tuning([other], "level", 1, False, 1)
        alone.step(10)
        other.step(10)
        for name in M.CARRIED:
            assert np.array_equal(rm.field(name), alone.field(name)), (name, s["seed"])
            d = rel_l2(rm.field(name), other.field(name))
            assert d <= 1e-5, (name, s["seed"], "against the level layout", d)


def oracle_restart(o, g, seed=1):
    """A fresh oracle run holding o's state (fields, fieldZ, columns, time)."""
    ref = OracleRun(g, C.profile(), C.run(g, seed=seed))
    for name in M.CARRIED + ["fieldZ"]:
        ref.set_field(name, o.field(name))
    for name in M.COLUMNS:
        ref.set_column(name, o.column(name))
    ref.set_time(o.diag()["time"])
    return ref


@pytest.mark.parametrize("front_lds", [None, 16], ids=["front-in-lds", "front-spilled"])
def test_lightning_matches_the_oracle_flash_for_flash(ctx, mature_small, front_lds, monkeypatch):
    # From 32.5 min into the small-grid storm the first flash comes at step 656 and 24 more follow by step 700 (the
    # oracle's counts): the same flashes, step by step, from the same seeded streams, and the charge fields after them.
    # With front_lds the pool keeps only that many entries of a flash's front in LDS and the rest in its global arrays
    # (METEOLAB_FRONT_LDS_ENTRIES, read at pool creation): the same flashes.
    g = C.small_grid()
    ref = oracle_restart(mature_small, g)
    if front_lds is not None:
        monkeypatch.setenv("METEOLAB_FRONT_LDS_ENTRIES", str(front_lds))
    r = restart(ctx, mature_small, g)
    monkeypatch.delenv("METEOLAB_FRONT_LDS_ENTRIES", raising=False)
    counts = []
    for k in range(50):
        ref.step(1); r.step(1)
        d, od = r.diagnostics(), ref.diag()
        counts.append((int(od["flashesIC"]), int(od["flashesCG"])))
        assert (d["flashesIC"], d["flashesCG"]) == counts[-1], (k, counts[-1], d["flashesIC"], d["flashesCG"])
    assert sum(counts[-1]) >= 20, counts[-1]
    assert r.diagnostics()["firstFlash"] == ref.diag()["firstFlash"] == 1968.0
    log = r.flashes()
    assert len(log) == sum(counts[-1])
    w = compare(r, ref)
    print("after 50 steps and", counts[-1], "flashes:", margin(w))
