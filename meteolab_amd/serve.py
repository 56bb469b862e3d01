"""Continuous batching of StormLab runs on one device: requests of two size classes arrive over time, are admitted into
the free slots of their class's pool, advance together one model minute per round (one launch per kernel per step for
all active runs of a class), and leave when done; a run's state stays resident from admission to release.

The clock is the device's: a round costs the measured device time of its steps (HIP events, ml_step_timed) plus the
wall time of the admissions made before it. Each request's model minutes are timed by the rounds that advanced them,
so the p50 / p99 latency per model-minute is over all (request, minute) pairs.

    python -m meteolab_amd.serve [--requests 1000] [--rate R] [--seed 1]

The workload restates the serving configuration: Poisson arrivals at a fixed rate, 70 % standard / 30 % large domains,
lengths U(10, 120) model minutes, the sounding and thermal drawn per request as for the regime map and ensemble.
"""
import argparse
import heapq
import json
import time

import numpy as np

from . import config as C
from .state import apply_state, read_state, save_state

MINUTE_STEPS = 20  # model minute at dt = 3 s


class Request:
    def __init__(self, rid, size, minutes, arrival, profile, run):
        self.id, self.size, self.minutes, self.arrival = rid, size, minutes, arrival
        self.profile, self.run = profile, run
        self.done_minutes = 0
        self.admitted = None
        self.finished = None
        self.minute_ms = []
        self.handle = None
        self.saved = None  # (.meteolab bytes, fieldZ) while evicted


def workload(requests=1000, rate=None, seed=1, min_minutes=10, max_minutes=120, large_share=0.3):
    """Poisson arrivals (rate per second of device time; by default the rate that keeps about 48 runs active),
    size class, length and the run's sounding and thermal, all from one seeded generator."""
    rng = np.random.default_rng(seed)
    if rate is None:
        rate = 48 / ((min_minutes + max_minutes) / 2 * 0.1)
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
    """Pools of `capacity` slots per size class on one context; a round steps every active run of a class one model
    minute."""

    def __init__(self, ctx, capacity=None, dt=3.0, grids=None, on_finish=None):
        import meteolab_amd as M
        self.M, self.ctx, self.dt, self.on_finish = M, ctx, dt, on_finish
        self.capacity = dict(capacity or {"standard": 64, "large": 16})
        self.grids = dict(grids or C.GRIDS)
        self.pools = {size: ctx.pool(self.grids[size], cap) for size, cap in self.capacity.items()}
        self.active = {size: [] for size in self.capacity}

    def slot_bytes(self):
        return {size: self.ctx.slot_bytes(self.grids[size]) for size in self.capacity}

    def admit(self, req):
        req.handle = self.pools[req.size].run(req.profile, req.run)
        self.active[req.size].append(req)

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
        req.handle = self.pools[req.size].run(req.profile, req.run)
        apply_state(req.handle, read_state(req.saved[0]))
        req.handle.set_field("fieldZ", req.saved[1])
        req.saved = None
        self.active[req.size].append(req)

    def round(self):
        """Steps every active run one model minute; returns the device ms of the round and the requests it
        finished."""
        total, finished = 0.0, []
        for size, reqs in self.active.items():
            if not reqs:
                continue
            ms = self.M.step_timed([r.handle for r in reqs], MINUTE_STEPS, self.dt)
            total += float(ms.sum())
        for size, reqs in self.active.items():
            keep = []
            for r in reqs:
                r.done_minutes += 1
                r.minute_ms.append(total)
                if r.done_minutes >= r.minutes:
                    finished.append(r)
                    if self.on_finish:
                        self.on_finish(r)  # the run's final state is still resident here
                    r.handle.release()
                    r.handle = None
                else:
                    keep.append(r)
            self.active[size] = keep
        return total, finished

    def serve(self, requests, log=None):
        pending = sorted(requests, key=lambda r: r.arrival)
        queue = []  # arrived, waiting for a slot: (arrival, id)
        clock_ms, k, done = 0.0, 0, []
        rounds = 0
        while k < len(pending) or queue or any(self.active.values()):
            while k < len(pending) and pending[k].arrival * 1e3 <= clock_ms:
                heapq.heappush(queue, (pending[k].arrival, pending[k].id, pending[k]))
                k += 1
            # Admit in arrival order while the request's class has a free slot.
            t0 = time.perf_counter()
            waiting = []
            while queue:
                item = heapq.heappop(queue)
                r = item[2]
                if len(self.active[r.size]) < self.capacity[r.size]:
                    self.admit(r)
                    r.admitted = clock_ms
                else:
                    waiting.append(item)
            for item in waiting:
                heapq.heappush(queue, item)
            self.ctx.sync()
            clock_ms += (time.perf_counter() - t0) * 1e3
            if not any(self.active.values()):
                if k < len(pending):
                    clock_ms = max(clock_ms, pending[k].arrival * 1e3)
                continue
            ms, finished = self.round()
            clock_ms += ms
            rounds += 1
            for r in finished:
                r.finished = clock_ms
                done.append(r)
            if log and rounds % 50 == 0:
                log(f"round {rounds}: clock {clock_ms / 1e3:.1f} s, active {[len(v) for v in self.active.values()]}, "
                    f"queued {len(queue)}, done {len(done)}")
        return done, clock_ms


def summary(done, clock_ms, grids=None):
    grids = grids or C.GRIDS
    minute = np.concatenate([np.asarray(r.minute_ms) for r in done]) if done else np.zeros(0)
    cells = sum(C.nodes(grids[r.size]) * r.minutes * MINUTE_STEPS for r in done)
    wait = np.asarray([r.admitted - r.arrival * 1e3 for r in done])
    return {"requests": len(done), "device_and_admission_s": clock_ms / 1e3,
            "cell_steps_per_s": cells / (clock_ms / 1e3) if clock_ms > 0 else 0.0,
            "latency_per_model_minute_ms": {"p50": float(np.percentile(minute, 50)), "p99": float(np.percentile(minute, 99)),
                                            "minutes": int(minute.size)},
            "queue_wait_ms": {"p50": float(np.percentile(wait, 50)), "p99": float(np.percentile(wait, 99))}}


def main():
    import meteolab_amd as M
    ap = argparse.ArgumentParser()
    ap.add_argument("--requests", type=int, default=1000)
    ap.add_argument("--rate", type=float, default=None)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--min-minutes", type=int, default=10)
    ap.add_argument("--max-minutes", type=int, default=120)
    ap.add_argument("--standard-slots", type=int, default=64)
    ap.add_argument("--large-slots", type=int, default=16)
    a = ap.parse_args()
    reqs = workload(a.requests, a.rate, a.seed, a.min_minutes, a.max_minutes)
    with M.Context(0, 32) as ctx:
        server = Server(ctx, {"standard": a.standard_slots, "large": a.large_slots})
        done, clock = server.serve(reqs, log=lambda s: print(s, flush=True))
        out = summary(done, clock)
        out["slot_bytes"] = server.slot_bytes()
        out["workload"] = {"requests": a.requests, "rate_per_s": a.rate, "minutes": [a.min_minutes, a.max_minutes],
                           "large_share": 0.3, "slots": server.capacity, "seed": a.seed}
        out["command"] = (f"python -m meteolab_amd.serve --requests {a.requests} --seed {a.seed} --min-minutes {a.min_minutes} "
                          f"--max-minutes {a.max_minutes} --standard-slots {a.standard_slots} --large-slots {a.large_slots}"
                          + (f" --rate {a.rate}" if a.rate else ""))
        print(json.dumps(out))


if __name__ == "__main__":
    main()
