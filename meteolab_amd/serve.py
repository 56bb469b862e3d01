"""Continuous batching of StormLab runs on one device: requests of two size classes arrive over time, are admitted while
the device's memory holds their slots, advance together one model minute per round (one launch per kernel per step for
all active runs of a class), and leave when done; a run's state stays resident from admission to release.

Each size class has a context of its own (its own HIP streams) and one pool that grows, in chunks of slots, as requests
are admitted. A request is admitted when its class's pool has a free slot, or when the pool can grow: the slots' bytes
(ml_slot_bytes) fit both the server's memory budget and the device's free memory (ml_mem_info) less a reserve. One host
thread keeps both classes' rounds in flight and polls them (ml_ctx_query): a class whose round ended hands back its
finished requests, admits the requests of its class that have arrived, and starts its next round at once, without
waiting for the other class, so the two classes' kernels run side by side on the device.

The clock is the wall clock (time.perf_counter) from the start of the stream, and requests arrive at their Poisson
times on it. A request's latency of a model minute is the wall time from the end of its previous minute (for the first,
from the start of its admission) to the end of this one: its class's round and the host's work between rounds.

    python -m meteolab_amd.serve [--requests 1000] [--rate R] [--seed 1]

The workload restates the serving configuration: Poisson arrivals at a fixed rate, 70 % standard / 30 % large domains,
lengths U(10, 120) model minutes, the sounding and thermal drawn per request as for the regime map and ensemble.
"""
import argparse
import collections
import json
import time

import numpy as np

from . import config as C
from .state import apply_state, read_state, save_state

MINUTE_STEPS = 20  # model minute at dt = 3 s
DEFAULT_RATE = 4.0  # requests per second of wall time


class Request:
    def __init__(self, rid, size, minutes, arrival, profile, run):
        self.id, self.size, self.minutes, self.arrival = rid, size, minutes, arrival
        self.profile, self.run = profile, run
        self.done_minutes = 0
        self.admitted = None   # ms on the server's clock
        self.finished = None
        self.last = None       # end of the request's last minute (or its admission), ms
        self.minute_ms = []
        self.handle = None
        self.saved = None  # (.meteolab bytes, fieldZ) while evicted


def workload(requests=1000, rate=DEFAULT_RATE, seed=1, min_minutes=10, max_minutes=120, large_share=0.3):
    """Poisson arrivals (`rate` requests per second), size class, length and the run's sounding and thermal, all from
    one seeded generator."""
    rng = np.random.default_rng(seed)
    t, out = 0.0, []
    for k in range(requests):
        t += rng.exponential(1 / rate)
        size = "large" if rng.random() < large_share else "standard"
        grid = C.GRIDS[size]
        i, j = int(rng.integers(0, 8)), int(rng.integers(0, 8))
        offset = (float(rng.uniform(-5000, 5000)), float(rng.uniform(-5000, 5000)))
        out.append(Request(k, size, int(rng.integers(min_minutes, max_minutes + 1)), t, C.regime_cell(i, j, k),
                           C.run(grid, seed=k + 1, amp=float(rng.uniform(1.5, 3.0)), offset=offset)))
    return out


class Server:
    """The runs of each size class in a growing pool on a context of the class's own; admission by device memory.

    grids: size class -> grid. memory: the bytes the server's slots may take: a dict by class, or in all (None: the
    device's free memory at start less `reserve`), split so that each class may hold as many runs as the others (a
    class's pool never shrinks, so a class keeps room of its own). chunk: slots a pool grows by. max_active: an optional
    bound on a class's active runs (the latency a caller is willing to trade for throughput), None for memory alone."""

    def __init__(self, grids=None, device=0, memory=None, reserve=4 << 30, chunk=16, max_active=None, on_finish=None,
                 dt=3.0, precision=32):
        import meteolab_amd as M
        self.M, self.dt, self.on_finish, self.chunk, self.reserve = M, dt, on_finish, int(chunk), int(reserve)
        self.grids = dict(grids or C.GRIDS)
        self.ctxs = {size: M.Context(device, precision) for size in self.grids}
        self.slot = {size: self.ctxs[size].slot_bytes(g) for size, g in self.grids.items()}
        free, self.total_memory = next(iter(self.ctxs.values())).mem_info()
        if isinstance(memory, dict):
            self.memory = {size: int(memory[size]) for size in self.grids}
        else:
            total = int(memory) if memory is not None else free - self.reserve
            self.memory = {size: total * b // sum(self.slot.values()) for size, b in self.slot.items()}
        short = [size for size in self.grids if self.memory[size] < self.slot[size]]
        if short:
            raise ValueError(f"serving memory {self.memory} holds no run of {short} (slots {self.slot})")
        self.max_active = max_active
        self.pools = {size: None for size in self.grids}
        self.active = {size: [] for size in self.grids}
        self.rounds = []  # (class, runs, wall ms) of every round of serve()

    def close(self):
        for size in self.grids:
            if self.pools[size] is not None:
                self.pools[size].destroy()
                self.pools[size] = None
        for ctx in self.ctxs.values():
            ctx.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def held(self, size=None):
        """Device bytes the pools' slots take (of one class, or of all)."""
        return sum(p.capacity * self.slot[s] for s, p in self.pools.items() if p is not None and size in (None, s))

    def room(self, size, wanted):
        """Makes room for up to `wanted` more runs of a class, growing its pool within the memory budget and the
        device's free memory; returns how many slots are free for it."""
        pool = self.pools[size]
        free = pool.free() if pool is not None else 0
        if self.max_active is not None:
            wanted = min(wanted, self.max_active - len(self.active[size]))
        if wanted <= free:
            return max(0, wanted)
        grow = max(self.chunk, wanted - free)
        device_free, _ = self.ctxs[size].mem_info()
        afford = min(self.memory[size] - self.held(size), device_free - self.reserve) // self.slot[size]
        grow = int(min(grow, afford))
        if grow <= 0:
            return free
        if pool is None:
            self.pools[size] = self.ctxs[size].pool(self.grids[size], grow)
        else:
            pool.grow(pool.capacity + grow)
        return min(wanted, free + grow)

    def admit_many(self, reqs, now_ms=0.0):
        """Admits requests of one class into free slots (ml_runs_create: one launch for all); returns those admitted."""
        if not reqs:
            return []
        size = reqs[0].size
        n = self.room(size, len(reqs))
        take = reqs[:n]
        if take:
            for r, h in zip(take, self.pools[size].runs([(r.profile, r.run) for r in take])):
                r.handle, r.admitted, r.last = h, now_ms, now_ms
            self.active[size].extend(take)
        return take

    def admit(self, req):
        if not self.admit_many([req]):
            raise RuntimeError(f"no device memory for a {req.size} run")

    def evict(self, req):
        """Takes a running request off the device to free its slot: its state as a .meteolab file (save_state) and
        the vertical field of its last discharge (fieldZ, which the file does not hold and the next step's charging
        reads before it is recomputed)."""
        req.saved = (save_state(req.handle, config=req.handle.spec), req.handle.field("fieldZ"))
        self.active[req.size].remove(req)
        req.handle.release()
        req.handle = None

    def readmit(self, req):
        """Puts an evicted request back into a free slot of its class from what evict kept; it goes on as if it had
        never left."""
        if self.room(req.size, 1) < 1:
            raise RuntimeError(f"no device memory for a {req.size} run")
        req.handle = self.pools[req.size].run(req.profile, req.run)
        apply_state(req.handle, read_state(req.saved[0]))
        req.handle.set_field("fieldZ", req.saved[1])
        req.saved = None
        self.active[req.size].append(req)

    def start(self, size):
        """Starts a class's round: every active run one model minute (asynchronous)."""
        self.M.step([r.handle for r in self.active[size]], MINUTE_STEPS, self.dt)

    def end(self, size, now_ms):
        """Ends a class's round (its context idle): each run's minute counted, finished requests released; returns
        them."""
        keep, finished = [], []
        for r in self.active[size]:
            r.done_minutes += 1
            r.minute_ms.append(now_ms - r.last)
            r.last = now_ms
            if r.done_minutes >= r.minutes:
                r.finished = now_ms
                finished.append(r)
                if self.on_finish:
                    self.on_finish(r)  # the run's final state is still resident here
                r.handle.release()
                r.handle = None
            else:
                keep.append(r)
        self.active[size] = keep
        return finished

    def round(self):
        """One round of every class with active runs, side by side, waited for; returns the requests it finished."""
        t0 = time.perf_counter()
        sizes = [s for s in self.grids if self.active[s]]
        for s in sizes:
            self.start(s)
        finished = []
        for s in sizes:
            self.ctxs[s].sync()
            finished += self.end(s, (time.perf_counter() - t0) * 1e3)
        return finished

    def serve(self, requests, log=None):
        """Serves the requests on the wall clock; returns (finished requests, wall ms to the last finish)."""
        pending = sorted(requests, key=lambda r: r.arrival)
        queues = {size: collections.deque() for size in self.grids}
        inflight = {size: None for size in self.grids}  # start of the class's round in flight (ms)
        done, k = [], 0
        t0 = time.perf_counter()
        clock = lambda: (time.perf_counter() - t0) * 1e3  # noqa: E731
        last_log = 0.0
        while k < len(pending) or any(queues.values()) or any(self.active.values()):
            now = clock()
            while k < len(pending) and pending[k].arrival * 1e3 <= now:
                queues[pending[k].size].append(pending[k])
                k += 1
            busy = False
            for size in self.grids:
                if inflight[size] is not None:
                    if not self.ctxs[size].idle():
                        busy = True
                        continue
                    end = clock()
                    self.rounds.append((size, len(self.active[size]), end - inflight[size]))
                    done += self.end(size, end)
                    inflight[size] = None
                q = queues[size]
                if q:
                    took = self.admit_many(list(q), clock())
                    for _ in took:
                        q.popleft()
                if self.active[size]:
                    inflight[size] = clock()
                    self.start(size)
                    busy = True
            if log and clock() - last_log > 10e3:
                last_log = clock()
                log(f"t {last_log / 1e3:.1f} s: active {[len(v) for v in self.active.values()]}, "
                    f"queued {[len(q) for q in queues.values()]}, done {len(done)}, held {self.held() / 2**30:.1f} GiB")
            if not busy:
                # Nothing on the device: wait for the next arrival.
                if k < len(pending):
                    time.sleep(max(0.0, pending[k].arrival - clock() / 1e3))
            else:
                time.sleep(2e-4)
        return done, clock()


def summary(done, clock_ms, grids=None, rounds=None):
    grids = grids or C.GRIDS
    minute = np.concatenate([np.asarray(r.minute_ms) for r in done]) if done else np.zeros(0)
    cells = sum(C.nodes(grids[r.size]) * r.minutes * MINUTE_STEPS for r in done)
    wait = np.asarray([r.admitted - r.arrival * 1e3 for r in done])
    out = {"requests": len(done), "wall_s": clock_ms / 1e3,
           "cell_steps_per_s": cells / (clock_ms / 1e3) if clock_ms > 0 else 0.0,
           "latency_per_model_minute_ms": {"p50": float(np.percentile(minute, 50)), "p99": float(np.percentile(minute, 99)),
                                           "minutes": int(minute.size)},
           "queue_wait_ms": {"p50": float(np.percentile(wait, 50)), "p99": float(np.percentile(wait, 99))}}
    by = {}
    for size in grids:
        m = [x for r in done if r.size == size for x in r.minute_ms]
        if m:
            by[size] = {"requests": sum(r.size == size for r in done), "p50": float(np.percentile(m, 50)),
                        "p99": float(np.percentile(m, 99))}
    out["latency_by_class_ms"] = by
    if rounds:
        for size in grids:
            rs = [(n, ms) for s, n, ms in rounds if s == size]
            if rs:
                out.setdefault("rounds", {})[size] = {"count": len(rs), "mean_runs": float(np.mean([n for n, _ in rs])),
                                                      "max_runs": int(max(n for n, _ in rs)),
                                                      "mean_ms": float(np.mean([ms for _, ms in rs]))}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--requests", type=int, default=1000)
    ap.add_argument("--rate", type=float, default=DEFAULT_RATE)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--min-minutes", type=int, default=10)
    ap.add_argument("--max-minutes", type=int, default=120)
    ap.add_argument("--chunk", type=int, default=16)
    ap.add_argument("--max-active", type=int, default=None)
    a = ap.parse_args()
    reqs = workload(a.requests, a.rate, a.seed, a.min_minutes, a.max_minutes)
    with Server(chunk=a.chunk, max_active=a.max_active) as server:
        done, clock = server.serve(reqs, log=lambda s: print(s, flush=True))
        out = summary(done, clock, rounds=server.rounds)
        out["slot_bytes"] = server.slot
        out["held_bytes"] = server.held()
        out["memory_budget_bytes"] = server.memory
        out["device_memory_bytes"] = server.total_memory
        out["workload"] = {"requests": a.requests, "rate_per_s": a.rate, "minutes": [a.min_minutes, a.max_minutes],
                           "large_share": 0.3, "seed": a.seed, "chunk": a.chunk, "max_active": a.max_active}
        out["command"] = (f"python -m meteolab_amd.serve --requests {a.requests} --rate {a.rate} --seed {a.seed} "
                          f"--min-minutes {a.min_minutes} --max-minutes {a.max_minutes} --chunk {a.chunk}"
                          + (f" --max-active {a.max_active}" if a.max_active else ""))
        print(json.dumps(out))


if __name__ == "__main__":
    main()
