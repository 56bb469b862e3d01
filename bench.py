"""Benchmark of the StormLab step on one MI355X: cell-steps/s of batched runs against the plain one-run step, p50/p99
latency per model-minute, the dominant kernel's roofline and the oracle's CPU rate on a bounded sample.

Default workload (the ensemble configuration): 64 standard-domain runs (40x32x50, dt 3 s), each with its own seeded
thermal (amplitude 1.5-3 K, centre offset +-5 km) and theta noise, advanced together for 600 steps (30 model minutes),
one launch per kernel for all runs, one captured hipGraph per step. The plain rate is that of members run one at a time,
each on its own through the same steps. A cell-step is one grid node advanced by one model step.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--runs R] [--grid standard|large]

Prints one JSON line.
"""
import argparse
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


def stage_bytes(n, stage, carried=17, buoyancy=True):
    """Algorithmic HBM bytes of one k_stage launch (one RK3 stage of `carried` fields) per run of n nodes: each field's
    stage input read, its start read (stage 0: the same array), its output written (float32), the face mass fluxes and
    their divergence read once (4 x float32) and, with w among the fields, the 8 fields of the buoyancy read once."""
    per_field = 8 if stage == 0 else 12
    return n * (carried * per_field + 16 + (8 * 4 if buoyancy else 0))


def stage_bytes_per_step(n, split):
    """The transport's algorithmic bytes of a step per run: three launches of the 17 carried fields, or, with the step
    split around its lightning, three of the 13 fields before the charges and three of the 4 charges."""
    if split:
        return sum(stage_bytes(n, s, 13) + stage_bytes(n, s, 4, buoyancy=False) for s in range(3))
    return sum(stage_bytes(n, s) for s in range(3))


# Compulsory float32 traffic of a whole step per cell (the step's fields, each phase fused, no reuse across the step's
# global syncs): fall speeds 44, sedimentation 84, face projection 96, RK3 transport 592, stage projections 160,
# positivity fix 92, column physics 148, turbulence and mixing 104, final projection 80, lightning 88.
STEP_BYTES_PER_CELL = 1488


def tuning_args(t):
    """set_batch_tuning's arguments for a shape reported by batch_tuning."""
    return {"layout": t["transport_layout"], "groups": t["field_groups"], "xcd": t["xcd_per_run"], "levels": t["levels_per_block"]}


def measured_traffic(kernel, runs, grid, tuning, split):
    """HBM bytes per launch of the transport kernel (the mean over a step's launches) from the committed counter summary
    (rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE, scaled by a copy kernel of known size) when it was taken on this
    workload, launch shape and step split, else None."""
    path = os.path.join(ROOT, "profiles", f"{kernel}_ensemble{runs}_{grid['nx']}x{grid['ny']}x{grid['nz']}_hbm.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        rec = json.load(f)
    shape = {k: tuning[k] for k in ("transport_layout", "field_groups", "levels_per_block", "xcd_per_run")}
    return rec["bytes_per_launch"] if rec.get("shape") == shape and rec.get("split", False) == split else None


def percentile(a, q):
    return float(np.percentile(np.asarray(a), q)) if len(a) else float("nan")


def cpu_baseline(grid, steps):
    """The oracle (single-threaded C++ restatement of the reference's step) on a bounded sample of the workload."""
    from oracle.binding import OracleRun
    o = OracleRun(grid, C.profile(), C.ensemble_member(grid, 0))
    t0 = time.perf_counter()
    o.step(steps)
    dt = time.perf_counter() - t0
    n = C.nodes(grid)
    return {"value": n * steps / dt, "unit": "cell-steps/s", "cores": 1, "kind": "port",
            "sample": f"1 run of {grid['nx']}x{grid['ny']}x{grid['nz']} (ensemble member 0), steps 1-{steps} from the "
                      f"initial state, 1 thread, {dt:.1f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=600)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--runs", type=int, default=64)
    ap.add_argument("--grid", default="standard", choices=["standard", "large"])
    ap.add_argument("--plain-members", type=int, default=4)
    ap.add_argument("--cpu-steps", type=int, default=300)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--table-only", action="store_true",
                    help="only the per-kernel table and the transport roofline (eager launches, no graph replays: "
                         "for hardware-counter runs, whose profiler cannot follow graph replays)")
    ap.add_argument("--compare-in-line", action="store_true",
                    help="also time the same runs with each step's lightning in line (ml_ctx_set_overlap off)")
    ap.add_argument("--transport", default="auto",
                    help="the transport kernel's launch shape: auto (the autotuner's), or layout:n:xcd with layout strip "
                         "(n field groups) or level / plane (n levels a block) and xcd 0 or 1")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py runs on one GPU only (--gpus {a.gpus} was asked): the batched runs are independent and a "
              "multi-GPU run would shard them by run with no exchange; no multi-GPU figure is measured.", file=sys.stderr)
        sys.exit(2)

    grid = C.GRIDS[a.grid]
    n = C.nodes(grid)
    prof = C.profile()
    ctx = M.Context(0, 32)

    # Batched: R runs in one pool, one launch per kernel for all of them.
    pool = ctx.pool(grid, a.runs)
    runs = [pool.run(prof, C.ensemble_member(grid, k)) for k in range(a.runs)]
    if a.transport != "auto":
        layout, count, xcd = a.transport.split(":")
        M.set_batch_tuning(runs, layout, groups=int(count) if layout == "strip" else 1, xcd=xcd == "1",
                           levels=int(count) if layout != "strip" else 1)
    if a.table_only:
        # The per-kernel table of the members from their initial state, in the shape asked (or the autotuner's).
        M.set_batch_tuning(runs, **tuning_args(M.batch_tuning(runs)))
        if a.warmup:
            M.step(runs, a.warmup, DT)
        kt = M.kernel_times(M.profile_step(runs, a.steps, DT))
        stage_kernel = next(k for k in ("k_stage_plane", "k_stage_level", "k_stage") if k in kt)
        split = "k_charge_courant" in kt
        launches, stage_ms = kt[stage_kernel]
        stage_b = stage_bytes_per_step(n, split) * a.runs
        print(json.dumps({"steps": a.steps, "tuning": {"batched": M.batch_tuning(runs)},
                          "roofline": {"kernel": stage_kernel, "bytes_per_launch": stage_b / launches,
                                       "ms_per_launch": stage_ms / launches, "launches_per_step": launches,
                                       "achieved": stage_b / (stage_ms / 1e3) / 1e9},
                          "kernels_ms_per_step": {k: round(v[1], 4) for k, v in sorted(kt.items(), key=lambda kv: -kv[1][1])},
                          # a copy of known size, whose counter readings calibrate the transport's
                          "copy_GBps": ctx.copy_bandwidth()}))
        ctx.close()
        return
    if a.warmup:
        M.step(runs, a.warmup, DT)
        ctx.sync()
    ms = M.step_timed(runs, a.steps, DT)
    total_s = ms.sum() / 1e3
    batched = a.runs * n * a.steps / total_s
    # Latency per model-minute: the device time of each consecutive 20 steps (60 s of model time) of the batch.
    per_minute = int(round(60 / DT))
    minutes = [ms[k:k + per_minute].sum() for k in range(0, len(ms) - per_minute + 1, per_minute)]

    # Optionally the same members through the same steps with each step's lightning in line (ml_ctx_set_overlap off):
    # what running a step's discharge beside the next step's dynamics gains. (Off by default: its whole-field transport
    # launches would mix into a kernel trace of this command.)
    in_line = None
    if a.compare_in_line:
        ctx.set_overlap(False)
        lpool = ctx.pool(grid, a.runs)
        lruns = [lpool.run(prof, C.ensemble_member(grid, k)) for k in range(a.runs)]
        M.set_batch_tuning(lruns, **tuning_args(M.batch_tuning(runs)))
        if a.warmup:
            M.step(lruns, a.warmup, DT)
        lms = M.step_timed(lruns, a.steps, DT)
        lpool.destroy()
        ctx.set_overlap(True)
        in_line = {"ms_per_step": float(lms.mean()), "value": a.runs * n * a.steps / (lms.sum() / 1e3),
                   "overlap_speedup": float(lms.sum() / ms.sum())}

    # Kernel times over the same workload: the same members from their initial state through the same steps, launched
    # eagerly with HIP events between the kernels on the library's stream (the graph replays above cannot be split by
    # kernel). The lightning kernels cost nothing until the storms electrify, so the table is over the whole run.
    epool = ctx.pool(grid, a.runs)
    eruns = [epool.run(prof, C.ensemble_member(grid, k)) for k in range(a.runs)]
    M.set_batch_tuning(eruns, **tuning_args(M.batch_tuning(runs)))
    if a.warmup:
        M.step(eruns, a.warmup, DT)
    kt = M.kernel_times(M.profile_step(eruns, a.steps, DT))
    epool.destroy()
    stage_kernel = next(k for k in ("k_stage_plane", "k_stage_level", "k_stage") if k in kt)
    # Per step: the transport's launches (3, or 6 with the step split around its lightning), their time and bytes.
    split = "k_charge_courant" in kt
    stage_launches, stage_ms = kt[stage_kernel]
    stage_b = stage_bytes_per_step(n, split) * a.runs
    achieved = stage_b / (stage_ms / 1e3)
    traffic = measured_traffic(stage_kernel, a.runs, grid, M.batch_tuning(runs), split)
    dominant = max(kt.items(), key=lambda kv: kv[1][1])[0]

    # Plain: members run one at a time, each through the same steps on its own (one-run launches).
    members = [k * a.runs // a.plain_members for k in range(a.plain_members)]
    ppool = ctx.pool(grid, len(members))
    plain_ms = []
    for k in members:
        r = ppool.run(prof, C.ensemble_member(grid, k))
        if a.warmup:
            M.step([r], a.warmup, DT)
        plain_ms.append(M.step_timed([r], a.steps, DT))
    pms = np.concatenate(plain_ms)
    plain = n * pms.size / (pms.sum() / 1e3)
    plain_tuning = M.batch_tuning([r])

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
        "config": {"workload": f"ensemble of {a.runs} {a.grid}-domain runs ({grid['nx']}x{grid['ny']}x{grid['nz']}, "
                               f"dt 3 s) batched in one launch per kernel, {a.steps} steps after {a.warmup} warm-up steps",
                   "runs": a.runs, "grid": f"{grid['nx']}x{grid['ny']}x{grid['nz']}", "dt": DT, "parallelism": "1 GPU"},
        "plain": {"value": plain, "unit": "cell-steps/s", "ms_per_step": float(pms.mean()), "steps": a.steps,
                  "members": members},
        "batched_over_plain": batched / plain,

        "latency_per_model_minute_ms": {"p50": percentile(minutes, 50), "p99": percentile(minutes, 99),
                                        "runs_per_batch": a.runs, "minutes": len(minutes)},
        "roofline": {"bound": "hbm", "kernel": stage_kernel, "achieved": achieved / 1e9, "peak": PEAK_HBM / 1e9,
                     "unit": "GB/s", "frac": achieved / PEAK_HBM, "traffic": traffic,
                     "bytes_per_launch": stage_b / stage_launches, "ms_per_launch": stage_ms / stage_launches,
                     "launches_per_step": stage_launches,
                     # The transport skips the blocks whose field is zero within the stage's reach, so `achieved` (the
                     # full-field algorithmic bytes over the time) is an effective rate; the HBM rate the kernel draws
                     # is the measured traffic over the same time.
                     "measured_GBps": traffic / (stage_ms / stage_launches / 1e3) / 1e9 if traffic else None},
        "step_roofline": {"bytes_per_cell_step": STEP_BYTES_PER_CELL, "achieved_GBps": batched * STEP_BYTES_PER_CELL / 1e9,
                          "frac": batched * STEP_BYTES_PER_CELL / PEAK_HBM},
        "kernels_ms_per_step": {k: round(v[1], 4) for k, v in sorted(kt.items(), key=lambda kv: -kv[1][1])},
        # The longest kernel a step in the per-kernel table (in line). The lightning's discharge, serial by the
        # reference's algorithm (a block per run growing the channel link by link, with a solve per stage), has no
        # bandwidth roofline to speak of; `roofline` is the transport's, the largest data-parallel kernel.
        "dominant_kernel": dominant,
        "tuning": {"batched": M.batch_tuning(runs), "plain": plain_tuning},
        "copy_GBps": ctx.copy_bandwidth(),
        "device": ctx.adapter,
    }
    if in_line:
        out["lightning_in_line"] = in_line
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_baseline(grid, a.cpu_steps)
    ctx.close()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
