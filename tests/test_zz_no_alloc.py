"""ml_model_diagnostics_noalloc, the entry point that allocates nothing: through the C ABI it gives what
ml_model_diagnostics gives (AtmosphereModel.diagnostics(), src/core/model.ts:1672-1696) for the same runs, bit for bit,
call after call, its scratch being the library's static device array; and it refuses more runs than that array holds."""
import ctypes

import numpy as np
import pytest

import meteolab_amd as M
from meteolab_amd import config as C

pytestmark = pytest.mark.gpu


def rows(diags):
    return np.array([[d[k] for k in M.MODEL_DIAGNOSTICS] for d in diags], np.float64)


def test_noalloc_diagnostics_equal_the_allocating_entry_point():
    g = C.small_grid()
    with M.Context(0, 32) as ctx:
        pool = ctx.pool(g, 6)
        runs = [pool.run(C.profile(), C.run(g, seed=s + 1, amp=1.5 + .3 * s, offset=(600.0 * s - 1500, 300.0 * s - 800)))
                for s in range(6)]
        M.step(runs, 200)  # 10 model minutes: updraughts, cloud and rain
        ctx.sync()
        print("case: 6 runs of the 16x12x24 grid after 200 steps, then 3 of them", flush=True)
        # Two calls in a row, the second with fewer runs than the first (the array keeps the first call's rows beyond
        # the second's count; none of them may leak into the output).
        for batch in (runs, runs[1:4]):
            want = rows(M.model_diagnostics(batch))
            got = rows(M.model_diagnostics_noalloc(batch))
            assert got.shape == (len(batch), len(M.MODEL_DIAGNOSTICS))
            assert np.array_equal(got, want, equal_nan=True), (got, want)
            assert want[:, M.MODEL_DIAGNOSTICS.index("updraft")].max() > 1.0
        # Over the limit: refused with the limit named, nothing launched.
        n = M.NOALLOC_RUNS + 1
        ids = (ctypes.c_int * n)(*([runs[0].h] * n))
        out = np.zeros((n, len(M.MODEL_DIAGNOSTICS)), np.float64)
        rc = M.lib().ml_model_diagnostics_noalloc(ctx.h, ids, n, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        assert rc == -1 and b"ML_NOALLOC_RUNS" in M.lib().ml_last_error(ctx.h)
        assert not out.any()
        # After the refusal the entry point still serves a call.
        assert np.array_equal(rows(M.model_diagnostics_noalloc(runs[:2])), rows(M.model_diagnostics(runs[:2])), equal_nan=True)
