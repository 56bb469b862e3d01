"""The precision study and the regime map on the batched step.

    python -m meteolab_amd.studies precision [--steps 1200] [--grid large]
    python -m meteolab_amd.studies regime [--steps 600]

precision: one run on the large domain stepped with the float32 kernels and with the same kernels in float64
(ml_ctx_create precision 64), the storm's metrics every 10 model minutes from both, the per-field relative L2
difference at the end, and each variant's cell-steps/s.

regime: the regime map of the reference's regime front-end (src/regime/grid.ts:32-109: moisture x shear cells), here
8 surface vapour caps (11..16 g/kg) x 8 shears (0..35 m/s) x 2 seeds = 128 runs, batched by size class (the 32 runs of
the two strongest shears on the large domain, the other 96 on the standard one), 30 model minutes; each cell's max w,
max theta', rain total, first flash and flash count, and the batch's cell-steps/s.
"""
import argparse
import json
import time

import numpy as np

from . import config as C

METRICS = ["maxW", "maxThetaP", "rainTotal", "firstFlash", "flashesIC", "flashesCG"]


def _metrics(d):
    return {k: float(d[k]) for k in METRICS}


def precision(steps=1200, grid="large", every=200):
    import meteolab_amd as M
    g = C.GRIDS[grid]
    out = {"grid": grid, "steps": steps, "variants": {}}
    finals = {}
    for prec in (32, 64):
        with M.Context(0, prec) as ctx:
            r = ctx.pool(g, 1).run(C.profile(), C.run(g))
            samples, ms_total = [], 0.0
            for k in range(steps // every):
                ms = M.step_timed([r], every, 3.0)
                ms_total += float(ms.sum())
                samples.append({"time": (k + 1) * every * 3.0, **_metrics(r.diagnostics())})
            finals[prec] = {name: r.field(name).astype(np.float64) for name in M.CARRIED}
            out["variants"][f"f{prec}"] = {"samples": samples, "ms_per_step": ms_total / steps,
                                            "cell_steps_per_s": C.nodes(g) * steps / (ms_total / 1e3)}
    rel = {}
    for name in finals[32]:
        a, b = finals[32][name], finals[64][name]
        nb = np.linalg.norm(b)
        rel[name] = float(np.linalg.norm(a - b) / nb) if nb > 0 else float(np.linalg.norm(a))
    out["final_relative_l2_f32_vs_f64"] = rel
    return out


def regime(steps=600, seeds=2):
    """The regime map: 8 moisture caps x 8 shears x `seeds` seeds, the standard-domain cells (shear index < 6) and the
    large ones batched by size class. Each class is timed alone (graph-replayed steps, device events), then both
    classes are stepped again from their initial state at the same time, each on a context of its own, on the wall
    clock: a class's discharge leaves most CUs idle, which the other class's steps then fill."""
    import meteolab_amd as M
    cells = [(i, j, s) for s in range(seeds) for i in range(8) for j in range(8)]
    classes = {"standard": [c for c in cells if c[1] < 6], "large": [c for c in cells if c[1] >= 6]}
    out = {"steps": steps, "cells": {}, "classes": {}}
    starts = {size: [(C.regime_cell(i, j, s), C.run(C.GRIDS[size], seed=1 + s, amp=2.0)) for (i, j, s) in members]
              for size, members in classes.items()}
    ctxs = {size: M.Context(0, 32) for size in classes}
    try:
        for size, members in classes.items():
            g = C.GRIDS[size]
            pool = ctxs[size].pool(g, len(members))
            t0 = time.perf_counter()
            runs = pool.runs(starts[size])
            ctxs[size].sync()
            create_s = time.perf_counter() - t0
            ms = M.step_timed(runs, steps, 3.0)
            for (i, j, s), r in zip(members, runs):
                out["cells"][f"{i},{j},{s}"] = {"size": size, "qv_max_gkg": 11 + 5 * i / 7, "shear": 5.0 * j, "seed": s,
                                               **_metrics(r.diagnostics())}
            out["classes"][size] = {"runs": len(members), "ms_per_step": float(ms.mean()), "create_s": create_s,
                                    "cell_steps_per_s": len(members) * C.nodes(g) * steps / (ms.sum() / 1e3)}
            pool.destroy()
        total_cells = sum(v["runs"] * C.nodes(C.GRIDS[k]) * steps for k, v in out["classes"].items())
        total_ms = sum(v["ms_per_step"] * steps for v in out["classes"].values())
        out["one_class_at_a_time"] = {"cell_steps_per_s": total_cells / (total_ms / 1e3), "device_s": total_ms / 1e3}
        # Both classes at once from their initial state: the same runs, so the same metrics.
        batches = {}
        for size, members in classes.items():
            pool = ctxs[size].pool(C.GRIDS[size], len(members))
            batches[size] = (pool, pool.runs(starts[size]))
            M.step(batches[size][1], 1, 3.0)   # graphs captured and the transport tuned before the clock starts
        for size in classes:
            ctxs[size].sync()
        t0 = time.perf_counter()
        for size in classes:
            M.step(batches[size][1], steps - 1, 3.0)
        for size in classes:
            ctxs[size].sync()
        wall = time.perf_counter() - t0
        cells_conc = sum(len(m) * C.nodes(C.GRIDS[k]) * (steps - 1) for k, m in classes.items())
        out["both_classes_at_once"] = {"cell_steps_per_s": cells_conc / wall, "wall_s": wall, "steps": steps - 1}
        out["cell_steps_per_s"] = out["both_classes_at_once"]["cell_steps_per_s"]
        out["metrics_equal_one_class_at_a_time"] = all(
            {k: v for k, v in _metrics(r.diagnostics()).items()} == {k: v for k, v in out["cells"][f"{i},{j},{s}"].items()
                                                                    if k not in ("size", "qv_max_gkg", "shear", "seed")}
            for size, members in classes.items() for (i, j, s), r in zip(members, batches[size][1]))
        for pool, _ in batches.values():
            pool.destroy()
    finally:
        for c in ctxs.values():
            c.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("study", choices=["precision", "regime"])
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--grid", default="large")
    a = ap.parse_args()
    if a.study == "precision":
        out = precision(a.steps or 1200, a.grid)
        cmd = f"python -m meteolab_amd.studies precision --steps {out['steps']} --grid {a.grid}"
    else:
        out = regime(a.steps or 600)
        cmd = f"python -m meteolab_amd.studies regime --steps {out['steps']}"
    print(json.dumps({"command": cmd, **out}))


if __name__ == "__main__":
    main()
