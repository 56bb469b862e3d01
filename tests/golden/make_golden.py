"""Writes the golden fixtures of the oracle (the CPU restatement of the reference step): the small grid's carried
fields after 10 steps, and the 30-minute metrics (max w, max theta', rain total, first flash, flash counts) of the
small and standard grids at 10, 20 and 30 model minutes. Run from the repository root:

    python tests/golden/make_golden.py
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from meteolab_amd import config as C  # noqa: E402
from oracle.binding import CARRIED, OracleRun  # noqa: E402

METRICS = ["time", "maxW", "maxThetaP", "rainTotal", "firstFlash", "flashesIC", "flashesCG"]


def metrics_run(grid, steps=600, every=200):
    o = OracleRun(grid, C.profile(), C.run(grid))
    out = []
    for k in range(steps // every):
        o.step(every)
        d = o.diag()
        out.append({m: float(d[m]) for m in METRICS})
    return out


def main():
    g = C.small_grid()
    o = OracleRun(g, C.profile(), C.run(g))
    o.step(10)
    np.savez_compressed(os.path.join(HERE, "small_10steps.npz"), **{name: o.field(name) for name in CARRIED})
    metrics = {"small": {"grid": g, "samples": metrics_run(g)}, "standard": {"grid": C.STANDARD, "samples": metrics_run(C.STANDARD)}}
    with open(os.path.join(HERE, "metrics_30min.json"), "w") as f:
        json.dump(metrics, f, indent=1)
    print(json.dumps({k: v["samples"][-1] for k, v in metrics.items()}))


if __name__ == "__main__":
    main()
