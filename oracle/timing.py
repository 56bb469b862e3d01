"""Timing of the oracle on the host's cores: the CPU baseline reported beside the GPU step.

TEST INFRASTRUCTURE ONLY: part of the oracle package (the CPU restatement of the reference's step the HIP step is
checked against). bench.py's cpu_baseline leg runs it; the product never does.

Each job is one oracle run (a grid, a sounding, a run configuration and optionally a state to restart from) advanced a
number of steps. Jobs run in separate processes, one thread each, all released together after their set-up, so the wall
time of a batch of jobs is that of the model steps alone.
"""
import multiprocessing as mp
import os
import queue
import time


def cpu_model():
    """The host CPU's model name (/proc/cpuinfo), or the platform's processor string."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    import platform
    return platform.processor() or "unknown"


def host_cores():
    """Cores this process may run on."""
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def _restore(o, state):
    for name, values in state["fields"].items():
        o.set_field(name, values)
    for name, values in state["columns"].items():
        o.set_column(name, values)
    o.set_time(state["time"])


def _worker(job, barrier, results, index):
    from oracle.binding import OracleRun
    grid, profile, run, state, steps = job
    o = OracleRun(grid, profile, run)
    if state is not None:
        _restore(o, state)
    barrier.wait()
    t0 = time.perf_counter()
    o.step(steps)
    results.put((index, t0, time.perf_counter()))


def time_jobs(jobs):
    """Runs `jobs` [(grid, profile, run, state or None, steps)] in as many single-threaded processes at once; returns
    the wall seconds from the common start to the last job's end, and each job's own seconds."""
    ctx = mp.get_context("spawn")
    barrier, results = ctx.Barrier(len(jobs)), ctx.Queue()
    procs = [ctx.Process(target=_worker, args=(job, barrier, results, k)) for k, job in enumerate(jobs)]
    for p in procs:
        p.start()
    got, waited = [], 0.0
    while len(got) < len(jobs):
        try:
            got.append(results.get(timeout=1.0))
        except queue.Empty:
            waited += 1.0
            if any(p.exitcode not in (None, 0) for p in procs) or waited > 900:
                for p in procs:
                    p.kill()
                raise RuntimeError("an oracle timing process failed or did not report back")
    for p in procs:
        p.join()
        if p.exitcode != 0:
            raise RuntimeError(f"oracle timing process exited with {p.exitcode}")
    start = min(t0 for _, t0, _ in got)
    end = max(t1 for _, _, t1 in got)
    own = [t1 - t0 for _, t0, t1 in sorted(got)]
    return end - start, own
