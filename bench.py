"""Benchmark of the StormLab step on one MI355X: cell-steps/s of batched runs against the plain one-run step, p50/p99
latency per model-minute, the transport kernel's roofline and the oracle's CPU rate on a bounded sample.

Workload (the ensemble configuration): 64 standard-domain runs (40x32x50, dt 3 s), each with its own seeded thermal
(amplitude 1.5-3 K, centre offset +-5 km) and theta noise, stepped together with one launch per kernel for all runs and
one captured hipGraph per step. A cell-step is one grid node advanced by one model step.

* The timed window: W untimed warm-up steps, then exactly K timed steps (the line's `value`, `ms_per_step`, `roofline`).
  The window's model time is named in `config.workload`.
* `full_run`: the same 64 members from their initial state through the whole 30-minute run (600 steps) in the same
  process: ms/step, p50/p99 per model-minute, the per-kernel table and the transport roofline over the whole run.
* `plain`: members run one at a time, each on its own through the same steps.
* `cpu_baseline`: the oracle (the C++ restatement of the reference's step) on the host's cores, one thread and all cores
  (one process per core), on a bounded sample: the first steps from the initial state and steps of the mature storm
  (restarted from the GPU state at 22.5 model minutes).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--runs R] [--grid standard|large] [--full-steps S]

Prints one JSON line. `--table-only` prints the per-kernel table of the window (eager launches only: the mode the
profiler runs are taken in); `--summarize` turns a profiler run of that mode into the summaries under profiles/.
"""
import argparse
import csv
import glob
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import meteolab_amd as M  # noqa: E402
from meteolab_amd import config as C  # noqa: E402

PEAK_HBM = 8.0e12  # MI355X HBM3E, bytes/s
DT = 3.0
METRIC = "cell-steps/s on 1 MI355X, batched vs plain; p50/p99 latency per model-minute"
STAGE_KERNELS = ("k_stage_plane", "k_stage_level", "k_stage")

# Algorithmic bytes of the RK3 transport per cell-step: each of the 17 carried fields read as the stage input and as the
# stage start and written (float32; stage 0 reads one array for both: 8 + 12 + 12 B over the three stages), and the
# face mass fluxes with their divergence read once a stage (4 x float32 x 3 stages).
TRANSPORT_BYTES_PER_CELL = 17 * (8 + 12 + 12) + 3 * 16  # 592
# Compulsory float32 traffic of a whole step per cell (the step's fields, each phase fused, no reuse across the step's
# global syncs): fall speeds 44, sedimentation 84, face projection 96, RK3 transport 592, stage projections 160,
# positivity fix 92, column physics 148, turbulence and mixing 104, final projection 80, lightning 88.
STEP_BYTES_PER_CELL = 1488
SNAPSHOT_STEP = 450  # 22.5 model minutes: the mature storms the CPU baseline's storm sample restarts from


def tuning_args(t):
    """set_batch_tuning's arguments for a shape reported by batch_tuning."""
    return {"layout": t["transport_layout"], "groups": t["field_groups"], "xcd": t["xcd_per_run"], "levels": t["levels_per_block"]}


def shape_key(t):
    return {k: t[k] for k in ("transport_layout", "field_groups", "levels_per_block", "xcd_per_run")}


def workload_tag(runs, grid):
    return f"ensemble{runs}_{grid['nx']}x{grid['ny']}x{grid['nz']}"


def window_tag(steps, warmup):
    return f"steps{warmup + 1}-{warmup + steps}"


def model_minutes(steps, warmup):
    return f"model minutes {warmup * DT / 60:g}-{(warmup + steps) * DT / 60:g} (steps {warmup + 1}-{warmup + steps})"


def percentile(a, q):
    return float(np.percentile(np.asarray(a), q)) if len(a) else float("nan")


def per_minute(ms):
    """Device time of each consecutive 20 steps (one model minute at dt 3 s)."""
    k = int(round(60 / DT))
    return [float(ms[i:i + k].sum()) for i in range(0, len(ms) - k + 1, k)]


def transport_roofline(kt, runs, n):
    """The transport kernel's roofline from a per-kernel table: algorithmic bytes per launch (592 B a cell-step over
    the step's launches of the kernel) over the table's mean launch time."""
    kernel = next(k for k in STAGE_KERNELS if k in kt)
    launches, ms = kt[kernel]
    per_launch = TRANSPORT_BYTES_PER_CELL * n * runs / launches
    achieved = per_launch / (ms / launches / 1e3)
    return kernel, {"bound": "hbm", "kernel": kernel, "achieved": achieved / 1e9, "peak": PEAK_HBM / 1e9, "unit": "GB/s",
                    "frac": achieved / PEAK_HBM, "bytes_per_launch": per_launch, "ms_per_launch": ms / launches,
                    "launches_per_step": launches, "bytes_per_cell_step": TRANSPORT_BYTES_PER_CELL}


def measured_traffic(kernel, runs, grid, tuning, steps, warmup):
    """HBM bytes per launch of the transport kernel over the same window and launch shape, from the committed counter
    summary of that window (rocprofv3 --pmc FETCH_SIZE, then WRITE_SIZE, scaled by a copy kernel of known size), else
    None."""
    path = os.path.join(ROOT, "profiles", f"{kernel}_{workload_tag(runs, grid)}_{window_tag(steps, warmup)}_hbm.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        rec = json.load(f)
    return rec["bytes_per_launch"] if rec.get("shape") == shape_key(tuning) else None


def members(pool, grid, count):
    """The ensemble's first `count` members, made in one batch (ml_runs_create)."""
    prof = C.profile()
    return pool.runs([(prof, C.ensemble_member(grid, k)) for k in range(count)])


def snapshot(run):
    """A run's state as the oracle restarts from it: the carried fields, fieldZ, the ground columns and the time."""
    return {"fields": {name: run.field(name) for name in M.CARRIED + ["fieldZ"]},
            "columns": {name: run.column(name) for name in M.COLUMNS}, "time": float(run.diagnostics()["time"])}


def cpu_baseline(grid, states, pre_steps, storm_steps, cores):
    """The oracle on the host: one thread, then `cores` single-threaded processes at once, each on a sample of the
    workload: `pre_steps` steps of ensemble members from their initial state and `storm_steps` steps of the same
    members restarted from the GPU's state at the snapshot step (the mature storms)."""
    from oracle.timing import cpu_model, host_cores, time_jobs
    prof, n = C.profile(), C.nodes(grid)
    cores = max(1, min(cores, len(states)))

    def sample(count):
        pre = [(grid, prof, C.ensemble_member(grid, k), None, pre_steps) for k in range(count)]
        storm = [(grid, prof, C.ensemble_member(grid, k), states[k], storm_steps) for k in range(count)]
        t_pre, _ = time_jobs(pre)
        t_storm, _ = time_jobs(storm)
        cells_pre, cells_storm = count * n * pre_steps, count * n * storm_steps
        return {"value": (cells_pre + cells_storm) / (t_pre + t_storm), "unit": "cell-steps/s", "cores": count,
                "pre_storm": {"value": cells_pre / t_pre, "steps": pre_steps, "wall_s": round(t_pre, 2)},
                "storm": {"value": cells_storm / t_storm, "steps": storm_steps, "wall_s": round(t_storm, 2)}}

    one = sample(1)
    every = sample(cores) if cores > 1 else one
    snap_min = SNAPSHOT_STEP * DT / 60
    return {"value": every["value"], "unit": "cell-steps/s", "cores": every["cores"], "kind": "port",
            "sample": f"the oracle (C++ restatement, float64 arithmetic), one single-threaded process per core: "
                      f"{every['cores']} ensemble members of {grid['nx']}x{grid['ny']}x{grid['nz']} for steps 1-{pre_steps} "
                      f"from their initial state, then the same members restarted from the GPU state at {snap_min:g} model "
                      f"minutes (mature storms) for {storm_steps} steps; value = all cell-steps of both over both wall times",
            "one_thread": one, "all_cores": every, "cpu_model": cpu_model(), "host_cores": host_cores()}


def table_only(a, grid, n):
    """The per-kernel table of the window (eager launches, warm-up included), as the profiler runs take it."""
    ctx = M.Context(0, 32)
    pool = ctx.pool(grid, a.runs)
    runs = members(pool, grid, a.runs)
    apply_transport(a, runs)
    if a.warmup:
        M.profile_step(runs, a.warmup, DT)
    kt = M.kernel_times(M.profile_step(runs, a.steps, DT))
    kernel, roof = transport_roofline(kt, a.runs, n)
    print(json.dumps({"steps": a.steps, "warmup": a.warmup, "runs": a.runs, "grid": f"{grid['nx']}x{grid['ny']}x{grid['nz']}",
                      "window": model_minutes(a.steps, a.warmup), "tuning": M.batch_tuning(runs), "roofline": roof,
                      "kernels_ms_per_step": {k: round(v[1], 4) for k, v in sorted(kt.items(), key=lambda kv: -kv[1][1])},
                      # a copy of known size, whose counter readings calibrate the transport's
                      "copy_GBps": ctx.copy_bandwidth()}))
    ctx.close()


def apply_transport(a, runs):
    if a.transport != "auto":
        layout, count, xcd = a.transport.split(":")
        M.set_batch_tuning(runs, layout, xcd=xcd == "1", levels=int(count))


# ---- summaries of profiler runs of --table-only (rocprofv3 -f csv) ---------------------------------------------------

def _csv_rows(directory, suffix):
    """The rows of the newest *suffix file under a profiler output directory (one profiled process)."""
    paths = glob.glob(os.path.join(directory, "**", f"*{suffix}"), recursive=True)
    if not paths:
        raise SystemExit(f"no *{suffix} under {directory}")
    with open(max(paths, key=os.path.getmtime), newline="") as f:
        return list(csv.DictReader(f))


def _grid(row):
    for key in ("Grid_Size_X", "Grid_Size"):
        if key in row and row[key] not in (None, ""):
            return int(row[key])
    return 0


def _short(name):
    """A kernel's name without its return type, namespaces and arguments (k_stage_plane<float>)."""
    name = name.replace("(anonymous namespace)::", "").replace("void ", "").replace("ml::", "")
    return name.split("(")[0].strip()


def _window_dispatches(rows, kernel, steps):
    """The transport's dispatches of the measured window: the last 3 * steps at each grid size of the step's last launches
    (the stage launches of the fields before the charges and of the charges), in dispatch order."""
    unique = {int(r["Dispatch_Id"]): r for r in rows if _short(r["Kernel_Name"]).startswith(kernel + "<")}  # a counter file has a row a counter
    mine = [unique[d] for d in sorted(unique)]
    sizes = sorted({_grid(r) for r in mine[-6:]})
    chosen = []
    for s in sizes:
        chosen += [r for r in mine if _grid(r) == s][-3 * steps:]
    return sorted(chosen, key=lambda r: int(r["Dispatch_Id"])), sizes


def _counter(rows, name, dispatches):
    by = {}
    for r in rows:
        if r["Counter_Name"] == name:
            by[int(r["Dispatch_Id"])] = by.get(int(r["Dispatch_Id"]), 0.0) + float(r["Counter_Value"])
    return [by[d] for d in dispatches if d in by]


def summarize(a):
    """Writes the kernel-trace summary, the HBM traffic record and the VALU record of a --table-only window from the
    rocprofv3 CSV directories given, beside the bench JSON of the traced command."""
    bench = json.load(open(a.bench_json))
    grid = C.GRIDS[a.grid]
    n, steps, warmup, runs = C.nodes(grid), bench["steps"], bench["warmup"], bench["runs"]
    kernel = bench["roofline"]["kernel"]
    base = f"{kernel}_{workload_tag(runs, grid)}_{window_tag(steps, warmup)}"
    cmd = f"python3 bench.py --table-only --no-cpu --steps {steps} --warmup {warmup} --runs {runs} --grid {a.grid}"
    out = {}
    if a.trace:
        rows = _csv_rows(a.trace, "kernel_trace.csv")
        chosen, sizes = _window_dispatches(rows, kernel, steps)
        us = [(int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3 for r in chosen]
        groups = {}
        for r in rows:
            key = (_short(r["Kernel_Name"]), _grid(r))
            d = (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3
            c, t = groups.get(key, (0, 0.0))
            groups[key] = (c + 1, t + d)
        total = sum(t for _, t in groups.values())
        lines = [f"# {kernel}: {runs} {a.grid}-domain runs, {model_minutes(steps, warmup)} — kernel trace", "",
                 f"Command (repository root, one MI355X): `rocprofv3 --kernel-trace --stats -f csv -d gpurun_out/trace -- {cmd}`", "",
                 "Every dispatch of the command, grouped by kernel and grid: the autotuner's stage-0 timing launches, the "
                 f"{warmup} warm-up steps and the {steps} measured steps, all launched eagerly, and the copy probe.", "",
                 "| kernel | grid x | dispatches | total ms | average us | % |", "|---|---|---|---|---|---|"]
        for (name, gx), (c, t) in sorted(groups.items(), key=lambda kv: -kv[1][1]):
            if t / total < 0.002:
                continue
            lines.append(f"| {name} | {gx} | {c} | {t / 1e3:.1f} | {t / c:.1f} | {100 * t / total:.1f} |")
        lines += ["", f"The measured window's transport dispatches: the last {3 * steps} at each of grids {sizes} "
                      f"({len(chosen)} in all) average **{np.mean(us):.1f} us** a launch; the bench JSON of the same "
                      f"command reports {bench['roofline']['ms_per_launch'] * 1e3:.1f} us (HIP events on the library stream), "
                      f"achieved {bench['roofline']['achieved']:.0f} GB/s at {TRANSPORT_BYTES_PER_CELL} B a cell-step "
                      f"(frac {bench['roofline']['frac']:.3f}).", f"Launch shape: {bench['tuning']}.", ""]
        with open(os.path.join(ROOT, "profiles", base + "_trace.md"), "w") as f:
            f.write("\n".join(lines))
        out["trace_us"] = float(np.mean(us))
        ids = [int(r["Dispatch_Id"]) for r in chosen]
    if a.fetch and a.write:
        ids_f = [int(r["Dispatch_Id"]) for r in _window_dispatches(_csv_rows(a.fetch, "counter_collection.csv"), kernel, steps)[0]]
        ids_w = [int(r["Dispatch_Id"]) for r in _window_dispatches(_csv_rows(a.write, "counter_collection.csv"), kernel, steps)[0]]
        frows, wrows = _csv_rows(a.fetch, "counter_collection.csv"), _csv_rows(a.write, "counter_collection.csv")
        copy_f = [float(r["Counter_Value"]) for r in frows if _short(r["Kernel_Name"]).startswith("k_copy") and r["Counter_Name"] == "FETCH_SIZE"]
        copy_w = [float(r["Counter_Value"]) for r in wrows if _short(r["Kernel_Name"]).startswith("k_copy") and r["Counter_Name"] == "WRITE_SIZE"]
        # The copy kernel reads and writes 1 GiB a dispatch: the scale from each counter's KiB to bytes.
        f_scale, w_scale = (1 << 30) / (np.mean(copy_f) * 1024), (1 << 30) / (np.mean(copy_w) * 1024)
        fetch = np.mean(_counter(frows, "FETCH_SIZE", ids_f)) * 1024 * f_scale
        write = np.mean(_counter(wrows, "WRITE_SIZE", ids_w)) * 1024 * w_scale
        algorithmic = bench["roofline"]["bytes_per_launch"]
        rec = {"kernel": kernel, "workload": f"{runs} {a.grid}-domain runs, ensemble members 0-{runs - 1}",
               "window": model_minutes(steps, warmup), "shape": shape_key(bench["tuning"]),
               "bytes_per_launch": fetch + write, "read_bytes_per_launch": fetch, "written_bytes_per_launch": write,
               "algorithmic_bytes_per_launch": algorithmic, "traffic_over_algorithmic": (fetch + write) / algorithmic,
               "fetch_scale": f_scale, "write_scale": w_scale, "dispatches": len(ids_f),
               "calibration": "k_copy of the copy probe (1 GiB read + 1 GiB written a dispatch): scale = 1 GiB / counter bytes",
               "commands": [f"rocprofv3 --pmc FETCH_SIZE -f csv -d gpurun_out/pmc_fetch -- {cmd}",
                            f"rocprofv3 --pmc WRITE_SIZE -f csv -d gpurun_out/pmc_write -- {cmd}"]}
        with open(os.path.join(ROOT, "profiles", base + "_hbm.json"), "w") as f:
            json.dump(rec, f, indent=1)
        out["traffic"] = rec["bytes_per_launch"]
    if a.valu:
        vrows = _csv_rows(a.valu, "counter_collection.csv")
        chosen, sizes = _window_dispatches(vrows, kernel, steps)
        names = sorted({r["Counter_Name"] for r in vrows})
        per = {}
        for s in sizes:
            ids_s = [int(r["Dispatch_Id"]) for r in chosen if _grid(r) == s]
            per[s] = {c: float(np.mean(_counter(vrows, c, ids_s))) for c in names}
        # Lane instructions per node and field of a stage launch: wave instructions x 64 lanes / (runs x n x fields).
        fields = {max(sizes): 13, min(sizes): 4} if len(sizes) == 2 else {sizes[0]: 17}
        rec = {"kernel": kernel, "window": model_minutes(steps, warmup), "shape": shape_key(bench["tuning"]),
               "counters_per_launch": {str(s): v for s, v in per.items()},
               "valu_lane_instructions_per_node_field": {str(s): per[s]["SQ_INSTS_VALU"] * 64 / (runs * n * fields[s]) for s in sizes
                                                         if "SQ_INSTS_VALU" in per[s]},
               "command": f"rocprofv3 --pmc {' '.join(names)} -f csv -d gpurun_out/pmc_valu -- {cmd}"}
        with open(os.path.join(ROOT, "profiles", base + "_valu.json"), "w") as f:
            json.dump(rec, f, indent=1)
        out["valu"] = rec["valu_lane_instructions_per_node_field"]
    print(json.dumps(out))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--runs", type=int, default=64)
    ap.add_argument("--grid", default="standard", choices=["standard", "large"])
    ap.add_argument("--full-steps", type=int, default=600, help="steps of the full run (0: none)")
    ap.add_argument("--plain-members", type=int, default=4)
    ap.add_argument("--cpu-pre-steps", type=int, default=60)
    ap.add_argument("--cpu-storm-steps", type=int, default=20)
    ap.add_argument("--cpu-cores", type=int, default=16, help="most processes of the all-core CPU sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--table-only", action="store_true",
                    help="only the per-kernel table of the window and the transport roofline (eager launches, no graph "
                         "replays: for the profiler's runs, which cannot follow graph replays)")
    ap.add_argument("--compare-in-line", action="store_true",
                    help="also time the window with each step's lightning in line (ml_ctx_set_overlap off)")
    ap.add_argument("--transport", default="auto",
                    help="the transport kernel's launch shape: auto (the grid's layout, tuned levels and block order), or "
                         "layout:n:xcd with layout level or plane, n levels a block and xcd 0 or 1")
    ap.add_argument("--summarize", action="store_true", help="write profiles/ summaries of --table-only profiler runs")
    ap.add_argument("--bench-json", help="--summarize: the --table-only JSON line of the profiled command")
    ap.add_argument("--trace", help="--summarize: the rocprofv3 --kernel-trace -f csv directory")
    ap.add_argument("--fetch", help="--summarize: the rocprofv3 --pmc FETCH_SIZE -f csv directory")
    ap.add_argument("--write", help="--summarize: the rocprofv3 --pmc WRITE_SIZE -f csv directory")
    ap.add_argument("--valu", help="--summarize: the rocprofv3 --pmc SQ_... -f csv directory")
    a = ap.parse_args()
    if a.summarize:
        summarize(a)
        return
    if a.gpus != 1:
        print(f"bench.py runs on one GPU only (--gpus {a.gpus} was asked): the batched runs are independent and a "
              "multi-GPU run would shard them by run with no exchange; no multi-GPU figure is measured.", file=sys.stderr)
        sys.exit(2)

    grid = C.GRIDS[a.grid]
    n = C.nodes(grid)
    if a.table_only:
        table_only(a, grid, n)
        return
    ctx = M.Context(0, 32)

    # The timed window: R runs in one pool, one launch per kernel for all of them, a graph replay per step.
    pool = ctx.pool(grid, a.runs)
    runs = members(pool, grid, a.runs)
    apply_transport(a, runs)
    if a.warmup:
        M.step(runs, a.warmup, DT)
        ctx.sync()
    ms = M.step_timed(runs, a.steps, DT)
    batched = a.runs * n * a.steps / (ms.sum() / 1e3)
    minutes = per_minute(ms)
    tuning = M.batch_tuning(runs)
    pool.destroy()

    # The window's kernel times: the same members from their initial state through the same warm-up and the same steps,
    # launched eagerly with HIP events between the kernels on the library's stream (a graph replay cannot be split).
    epool = ctx.pool(grid, a.runs)
    eruns = members(epool, grid, a.runs)
    M.set_batch_tuning(eruns, **tuning_args(tuning))
    if a.warmup:
        M.step(eruns, a.warmup, DT)
    kt = M.kernel_times(M.profile_step(eruns, a.steps, DT))
    epool.destroy()
    kernel, roofline = transport_roofline(kt, a.runs, n)
    roofline["traffic"] = measured_traffic(kernel, a.runs, grid, tuning, a.steps, a.warmup)
    roofline["window"] = model_minutes(a.steps, a.warmup)

    in_line = None
    if a.compare_in_line:
        ctx.set_overlap(False)
        lpool = ctx.pool(grid, a.runs)
        lruns = members(lpool, grid, a.runs)
        M.set_batch_tuning(lruns, **tuning_args(tuning))
        if a.warmup:
            M.step(lruns, a.warmup, DT)
        lms = M.step_timed(lruns, a.steps, DT)
        lpool.destroy()
        ctx.set_overlap(True)
        in_line = {"ms_per_step": float(lms.mean()), "value": a.runs * n * a.steps / (lms.sum() / 1e3),
                   "overlap_speedup": float(lms.sum() / ms.sum())}

    # Plain: members run one at a time, each alone through the window and through the full run.
    picks = [k * a.runs // a.plain_members for k in range(a.plain_members)]
    ppool = ctx.pool(grid, 1)
    plain_ms, plain_full = [], []
    for k in picks:
        r = ppool.run(C.profile(), C.ensemble_member(grid, k))
        if a.warmup:
            M.step([r], a.warmup, DT)
        plain_ms.append(M.step_timed([r], a.steps, DT))
        r.release()
        if a.full_steps:
            r = ppool.run(C.profile(), C.ensemble_member(grid, k))
            plain_full.append(M.step_timed([r], a.full_steps, DT))
            r.release()
    plain_tuning = M.batch_tuning([ppool.run(C.profile(), C.ensemble_member(grid, 0))])
    pms = np.concatenate(plain_ms)
    plain = n * pms.size / (pms.sum() / 1e3)

    # The full run: the same members from their initial state through the 30 model minutes, a graph replay per step;
    # at the snapshot step the states the CPU baseline's storm sample restarts from are read (outside the timed steps).
    full, states = None, []
    if a.full_steps:
        fpool = ctx.pool(grid, a.runs)
        fruns = members(fpool, grid, a.runs)
        M.set_batch_tuning(fruns, **tuning_args(tuning))
        first = min(SNAPSHOT_STEP, a.full_steps)
        fms = [M.step_timed(fruns, first, DT)]
        if not a.no_cpu:
            states = [snapshot(r) for r in fruns[:min(a.cpu_cores, a.runs)]]
        if a.full_steps > first:
            fms.append(M.step_timed(fruns, a.full_steps - first, DT))
        fms = np.concatenate(fms)
        fpool.destroy()
        tpool = ctx.pool(grid, a.runs)
        truns = members(tpool, grid, a.runs)
        M.set_batch_tuning(truns, **tuning_args(tuning))
        fkt = M.kernel_times(M.profile_step(truns, a.full_steps, DT))
        tpool.destroy()
        _, froof = transport_roofline(fkt, a.runs, n)
        froof["traffic"] = measured_traffic(kernel, a.runs, grid, tuning, a.full_steps, 0)
        fvalue = a.runs * n * a.full_steps / (fms.sum() / 1e3)
        pf = np.concatenate(plain_full)
        fmin = per_minute(fms)
        full = {"steps": a.full_steps, "window": model_minutes(a.full_steps, 0), "value": fvalue,
                "ms_per_step": float(fms.mean()),
                "latency_per_model_minute_ms": {"p50": percentile(fmin, 50), "p99": percentile(fmin, 99),
                                                "max": max(fmin), "minutes": len(fmin)},
                "plain": {"value": n * pf.size / (pf.sum() / 1e3), "ms_per_step": float(pf.mean()), "members": picks},
                "batched_over_plain": fvalue / (n * pf.size / (pf.sum() / 1e3)),
                "roofline": froof,
                "step_roofline": {"bytes_per_cell_step": STEP_BYTES_PER_CELL, "achieved_GBps": fvalue * STEP_BYTES_PER_CELL / 1e9,
                                  "frac": fvalue * STEP_BYTES_PER_CELL / PEAK_HBM},
                "kernels_ms_per_step": {k: round(v[1], 4) for k, v in sorted(fkt.items(), key=lambda kv: -kv[1][1])},
                "dominant_kernel": max(fkt.items(), key=lambda kv: kv[1][1])[0]}

    out = {
        "metric": METRIC,
        "value": batched,
        "unit": "cell-steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": float(ms.mean()),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"ensemble of {a.runs} {a.grid}-domain runs ({grid['nx']}x{grid['ny']}x{grid['nz']}, dt 3 s) "
                               f"batched in one launch per kernel: {model_minutes(a.steps, a.warmup)} of the 30-minute run "
                               f"after {a.warmup} warm-up steps (full_run: all 600 steps)",
                   "runs": a.runs, "grid": f"{grid['nx']}x{grid['ny']}x{grid['nz']}", "dt": DT, "parallelism": "1 GPU"},
        "plain": {"value": plain, "unit": "cell-steps/s", "ms_per_step": float(pms.mean()), "steps": a.steps, "members": picks},
        "batched_over_plain": batched / plain,
        "latency_per_model_minute_ms": {"p50": percentile(minutes, 50), "p99": percentile(minutes, 99),
                                        "runs_per_batch": a.runs, "minutes": len(minutes)},
        "roofline": roofline,
        "kernels_ms_per_step": {k: round(v[1], 4) for k, v in sorted(kt.items(), key=lambda kv: -kv[1][1])},
        "dominant_kernel": max(kt.items(), key=lambda kv: kv[1][1])[0],
        "tuning": {"batched": tuning, "plain": plain_tuning},
        "copy_GBps": ctx.copy_bandwidth(),
        "device": ctx.adapter,
    }
    if full:
        out["full_run"] = full
    if in_line:
        out["lightning_in_line"] = in_line
    ctx.close()
    if not a.no_cpu:
        from oracle.timing import host_cores
        if not states:
            raise SystemExit("the CPU baseline's storm sample needs --full-steps of at least the snapshot step")
        out["cpu_baseline"] = cpu_baseline(grid, states, a.cpu_pre_steps, a.cpu_storm_steps, min(a.cpu_cores, host_cores()))
    print(json.dumps(out))


if __name__ == "__main__":
    main()
