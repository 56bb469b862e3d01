"""The transport kernels' WENO5 face reconstruction (the device's own function, through ml_check_weno) against the
oracle's weno (advection.ts:388-399, pinned on the reference's known answers in test_oracle.py), over stencils of every
scale a run meets and more: 1e-30 to 1e8, signed, one-signed, near-constant, with zeros, all zero.

The bound: the face value minus c within 4e-7 of the stencil's variation max |q - c| for the float32 kernels, and
never a non-finite value. The float64 kernels are held to 1e-9: the oracle's own rounding of values 1e5 times their
variation (the smoothest stencils here) is ~1e-11 of it."""
import numpy as np
import pytest

from oracle.binding import lib as oracle_lib

pytestmark = pytest.mark.gpu


def stencils():
    rng = np.random.default_rng(5)
    out = []
    for scale in range(-30, 9):
        for mode in range(5):
            u = rng.uniform(-1, 1, (400, 5))
            if mode == 0:
                v = u
            elif mode == 1:
                v = np.abs(u)
            elif mode == 2:
                v = 1 + 0.01 * u            # near-constant
            elif mode == 3:
                v = u.copy(); v[:, :2] = 0  # zeros upwind
            else:
                v = 1 + 1e-5 * u            # smooth to float32's last digits
            out.append((v * 10.0 ** scale).astype(np.float32))
    out.append(np.zeros((4, 5), np.float32))
    return np.concatenate(out)


@pytest.mark.parametrize("precision,tol", [(32, 4e-7), (64, 1e-9)])
def test_face_values_match_the_oracle_at_every_scale(precision, tol):
    import meteolab_amd as M
    st = stencils()
    L = oracle_lib()
    want = np.array([L.orc_weno(*map(float, s)) - float(s[2]) for s in st])
    with M.Context(0, precision) as ctx:
        print("WENO check,", len(st), "stencils, float", precision, flush=True)
        got = ctx.check_weno(st)
    assert np.all(np.isfinite(got))
    variation = np.max(np.abs(st.astype(np.float64) - st[:, 2:3]), axis=1)
    err = np.abs(got - want)
    rel = np.where(variation > 0, err / np.where(variation > 0, variation, 1), err)
    k = int(np.argmax(rel))
    print(f"worst {rel[k]:.2e} of the variation at {st[k].tolist()}, bound {tol:g}")
    assert rel[k] <= tol
    assert np.all(got[variation == 0] == 0)
