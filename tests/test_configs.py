"""The batched workloads at their own sizes, through the C ABI: the ensemble (64 standard-domain members in one launch),
the regime map (cells of both size classes, batched by class) and a mixed-class serve on the real grids. Each run of a
batch matches the oracle (the CPU restatement of AtmosphereModel.step, src/core/model.ts:873-877) as a run alone does,
and equals the same run stepped alone bit for bit."""
import numpy as np
import pytest

import meteolab_amd as M
from meteolab_amd import config as C
from meteolab_amd import serve as S
from oracle.binding import OracleRun
from test_gpu_step import TOL, compare, margin

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = M.Context(0, 32)
    yield c
    c.close()


def equal_alone(ctx, grid, run, profile, spec, steps):
    alone = ctx.pool(grid, 1).run(profile, spec)
    alone.step(steps)
    for name in M.CARRIED:
        assert np.array_equal(run.field(name), alone.field(name)), name
    alone.pool.destroy()


def test_ensemble_of_64_standard_members_in_one_launch(ctx):
    # The ensemble configuration: 64 standard-domain members (seeded amplitude, offset and noise) stepped together.
    g = C.STANDARD
    print("case: 64 standard members, 10 + 10 steps", flush=True)
    pool = ctx.pool(g, 64)
    runs = [pool.run(C.profile(), C.ensemble_member(g, k)) for k in range(64)]
    M.step(runs, 10)
    ctx.sync()
    for k in (0, 37):
        o = OracleRun(g, C.profile(), C.ensemble_member(g, k))
        o.step(10)
        print(f"member {k} after 10 steps:", margin(compare(runs[k], o)))
    M.step(runs, 10)
    ctx.sync()
    for k in (5, 63):
        equal_alone(ctx, g, runs[k], C.profile(), C.ensemble_member(g, k), 20)
    pool.destroy()


REGIME_CELLS = [(0, 0, 0), (7, 7, 1), (3, 5, 0), (5, 2, 1)]  # (moisture row, shear column, seed)


@pytest.mark.parametrize("size", ["standard", "large"])
def test_regime_cells_batched_by_size_class(ctx, size):
    # The regime map's cells (surface vapour cap 11-16 g/kg x shear 0-35 m/s, two seeds), four of a size class in one
    # batch: each against the oracle after 10 steps.
    g = C.GRIDS[size]
    print(f"case: 4 {size} regime cells, 10 steps", flush=True)
    pool = ctx.pool(g, len(REGIME_CELLS))
    starts = [(C.regime_cell(i, j, s), C.run(g, seed=1 + s, amp=2.0)) for i, j, s in REGIME_CELLS]
    runs = [pool.run(p, spec) for p, spec in starts]
    M.step(runs, 10)
    ctx.sync()
    for (i, j, s), (p, spec), r in zip(REGIME_CELLS, starts, runs):
        o = OracleRun(g, p, spec)
        o.step(10)
        print(f"cell ({i}, {j}, seed {s}) after 10 steps:", margin(compare(r, o, tol=TOL)))
    pool.destroy()


def test_mixed_class_serve_on_the_real_grids(ctx):
    # The serving configuration's mix on the standard and large domains: 8 standard and 4 large requests of 1-2 model
    # minutes arriving together into fewer slots than requests, so runs join and leave between rounds; every request
    # ends as its run alone.
    reqs = []
    for k in range(12):
        size = "large" if k % 3 == 2 else "standard"
        g = C.GRIDS[size]
        reqs.append(S.Request(k, size, 1 + k % 2, 0.0, C.regime_cell(k % 8, (3 * k) % 8, k),
                              C.run(g, seed=k + 1, amp=1.8 + 0.1 * k, offset=(400.0 * k - 2000, -300.0 * k + 1500))))
    print("case: serve 8 standard + 4 large requests with memory for 5 + 2 runs", flush=True)
    finals = {}
    memory = {size: n * ctx.slot_bytes(C.GRIDS[size]) for size, n in (("standard", 5), ("large", 2))}
    with S.Server(memory=memory, chunk=2,
                  on_finish=lambda r: finals.setdefault(r.id, {name: r.handle.field(name) for name in M.CARRIED})) as server:
        done, clock = server.serve(reqs)
        assert server.pools["standard"].capacity == 5 and server.pools["large"].capacity == 2
    assert sorted(r.id for r in done) == list(range(12))
    for r in reqs:
        g = C.GRIDS[r.size]
        alone = ctx.pool(g, 1).run(r.profile, r.run)
        alone.step(r.minutes * S.MINUTE_STEPS)
        for name in M.CARRIED:
            assert np.array_equal(finals[r.id][name], alone.field(name)), (r.id, r.size, name)
        alone.pool.destroy()
