"""The continuous-batching serving loop (meteolab_amd/serve.py): admission by device memory into pools that grow,
rounds of one model minute, release; every request ends where it would alone."""
import numpy as np
import pytest

import meteolab_amd as M
from meteolab_amd import config as C
from meteolab_amd import serve as S


def test_workload_is_seeded_and_shaped():
    a, b = S.workload(200, rate=5.0, seed=3), S.workload(200, rate=5.0, seed=3)
    assert [(r.size, r.minutes, r.arrival) for r in a] == [(r.size, r.minutes, r.arrival) for r in b]
    arrivals = np.array([r.arrival for r in a])
    assert np.all(np.diff(arrivals) > 0)
    assert 0.15 < np.mean([r.size == "large" for r in a]) < 0.45
    assert min(r.minutes for r in a) >= 10 and max(r.minutes for r in a) <= 120
    gaps = np.diff(np.concatenate([[0], arrivals]))
    assert gaps.mean() == pytest.approx(1 / 5.0, rel=0.25)


@pytest.mark.gpu
def test_requests_served_together_end_as_alone():
    # Six small-grid requests of 1-3 model minutes arriving together into 3 slots of one class: runs join and leave
    # between rounds, so the captured step of each batch size is replayed over changing runs.
    g = C.small_grid()
    grids = {"small": g}
    reqs = [S.Request(k, "small", 1 + k % 3, 0.0, C.profile(qv_max=0.012 + 0.0005 * k), C.run(g, seed=k + 1, amp=2.0 + 0.1 * k))
            for k in range(6)]
    with M.Context(0, 32) as ctx:
        finals = {}
        # Memory for three runs, a slot at a time: the pool grows to three, and the other requests wait.
        with S.Server(grids=grids, memory={"small": 3 * ctx.slot_bytes(g)}, chunk=1,
                      on_finish=lambda r: finals.setdefault(r.id, {name: r.handle.field(name) for name in M.CARRIED})) as server:
            done, clock = server.serve(reqs)
            assert server.pools["small"].capacity == 3
        assert sorted(r.id for r in done) == list(range(6))
        assert all(len(r.minute_ms) == r.minutes for r in done)
        assert max(r.admitted for r in done) > 0  # some waited for a slot
        out = S.summary(done, clock, grids)
        assert out["latency_per_model_minute_ms"]["p50"] > 0
        for r in reqs:
            alone = ctx.pool(g, 1).run(r.profile, r.run)
            alone.step(r.minutes * S.MINUTE_STEPS)
            for name in M.CARRIED:
                assert np.array_equal(finals[r.id][name], alone.field(name)), (r.id, name)


@pytest.mark.gpu
def test_an_evicted_request_readmitted_from_its_state_file_ends_as_alone():
    # Checkpoint and restore of a resident run: after a round a request leaves the device as a .meteolab file (and its
    # discharge's vertical field), a round goes on without it, and it comes back and finishes as it would alone.
    g = C.small_grid()
    grids = {"small": g}
    a, b = [S.Request(k, "small", 3, 0.0, C.profile(qv_max=0.013), C.run(g, seed=k + 3, amp=2.5)) for k in range(2)]
    with M.Context(0, 32) as ctx, S.Server(grids=grids, memory={"small": 2 * ctx.slot_bytes(g)}, chunk=2,
                                           on_finish=lambda r: finals.setdefault(r.id, {name: r.handle.field(name) for name in M.CARRIED})) as server:
        finals = {}
        server.admit(a)
        server.admit(b)
        server.round()
        server.evict(a)
        assert a.handle is None and a.saved[0][:4] == b"MLAB"
        server.round()           # b alone
        server.readmit(a)
        server.round()           # b finishes
        server.round()           # a finishes
        assert a.done_minutes == 3 and b.done_minutes == 3 and set(finals) == {0, 1}
        for r in (a, b):
            alone = ctx.pool(g, 1).run(r.profile, r.run)
            alone.step(r.minutes * S.MINUTE_STEPS)
            for name in M.CARRIED:
                assert np.array_equal(finals[r.id][name], alone.field(name)), (r.id, name)


@pytest.mark.gpu
def test_a_grown_pool_keeps_its_runs_and_batched_creation_equals_one_by_one():
    # ml_pool_grow: runs stepped before the growth go on as they would have; ml_runs_create: runs made in one batch
    # start (and step) as the same runs made one at a time.
    g = C.small_grid()
    starts = [(C.profile(qv_max=0.012 + 0.0005 * k), C.run(g, seed=k + 1, amp=2.0 + 0.2 * k)) for k in range(5)]
    with M.Context(0, 32) as ctx:
        pool = ctx.pool(g, 2)
        first = pool.runs(starts[:2])
        M.step(first, 5)
        ctx.sync()
        pool.grow(5)
        assert pool.capacity == 5 and pool.free() == 3
        later = pool.runs(starts[2:])
        for r, (p, spec) in zip(later, starts[2:]):
            one = ctx.pool(g, 1).run(p, spec)
            for name in M.CARRIED:
                assert np.array_equal(r.field(name), one.field(name)), name
            one.pool.destroy()
        M.step(first + later, 5)
        ctx.sync()
        for k, r in enumerate(first + later):
            alone = ctx.pool(g, 1).run(*starts[k])
            alone.step(10 if k < 2 else 5)
            for name in M.CARRIED:
                assert np.array_equal(r.field(name), alone.field(name)), (k, name)
            alone.pool.destroy()
