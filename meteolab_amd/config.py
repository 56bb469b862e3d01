"""Run configurations of the StormLab step: grids, soundings and the initial thermal, as plain dicts.

The values restate the reference's: the standard and large grids (src/core/grid.ts:31-34), the Weisman-Klemp
sounding family (src/core/profiles.ts:349-361) with a unidirectional tanh shear, and the SimConfig values of
SUMMER_DAY in single-storm mode (src/core/fixtures.ts:6-12, src/core/config.ts:152) that matter for the step.
"""

# src/core/grid.ts:31 DEFAULT_GRID and :34 LARGE_GRID (50 levels stretched from 100 m to an 18.9 km top).
STANDARD = dict(nx=40, ny=32, nz=50, width=48_000.0, depth=36_000.0, height=18_900.0, bottom=100.0)
LARGE = dict(nx=80, ny=64, nz=50, width=96_000.0, depth=72_000.0, height=18_900.0, bottom=100.0)
GRIDS = {"standard": STANDARD, "large": LARGE}


def small_grid(nx=16, ny=12, nz=24, height=15_000.0, bottom=150.0):
    """A small grid with the standard spacing (for tests that run the oracle in seconds)."""
    return dict(nx=nx, ny=ny, nz=nz, width=nx * 1200.0, depth=ny * 1125.0, height=height, bottom=bottom)


def nodes(grid):
    return grid["nx"] * grid["ny"] * grid["nz"]


def profile(qv_max=0.014, shear=25.0, shear_depth=3000.0):
    """Weisman-Klemp thermodynamics (theta 300 + 43 (z/12 km)^1.25, 213 K above; RH 1 - .75 (z/12 km)^1.25) and
    the wind u = shear tanh(z / shear_depth)."""
    return dict(theta0=300.0, dtheta=43.0, ztr=12_000.0, ttr=213.0, rh_drop=0.75, rh_top=0.25, qv_max=float(qv_max),
                shear=float(shear), shear_depth=float(shear_depth))


def run(grid, seed=1, amp=2.0, offset=(0.0, 0.0), noise_seed=None, noise_amp=0.1, tropopause_km=12.0):
    """A run: the thermal of Weisman (1993) (src/core/model.ts:94 ELLIPSOID: 10 km / 1.4 km radii, centred at 1.4 km,
    a third of the way in from the west, mid-domain in y) of amplitude `amp` K shifted by `offset` m, plus uniform
    +-noise_amp K noise in theta below 1 km from mulberry32(noise_seed)."""
    return dict(seed=int(seed), latitude=45.0, turbulence=0.55, tropopause_km=float(tropopause_km), roughness=0.03,
                frame_u=0.0, frame_v=0.0, bubble_amp=float(amp),
                bubble_x=grid["width"] / 3 + offset[0], bubble_y=grid["depth"] / 2 + offset[1], bubble_z=1400.0,
                bubble_rh=10_000.0, bubble_rz=1400.0,
                noise_seed=int(seed if noise_seed is None else noise_seed), noise_amp=float(noise_amp),
                noise_top=1000.0)


def mulberry32(seed):
    """src/core/math.ts:7-14, as a Python generator of floats in [0, 1)."""
    s = seed & 0xFFFFFFFF
    while True:
        s = (s + 0x6D2B79F5) & 0xFFFFFFFF
        t = ((s ^ (s >> 15)) * (1 | s)) & 0xFFFFFFFF
        t = (((t + (((t ^ (t >> 7)) * (61 | t)) & 0xFFFFFFFF)) & 0xFFFFFFFF) ^ t) & 0xFFFFFFFF
        yield ((t ^ (t >> 14)) & 0xFFFFFFFF) / 4294967296.0


def ensemble_member(grid, k):
    """Ensemble member k (seed k): amplitude U(1.5, 3) K, centre offset U(-5, 5) km per axis, its own noise."""
    r = mulberry32(1000 + k)
    amp = 1.5 + 1.5 * next(r)
    ox = -5000.0 + 10000.0 * next(r)
    oy = -5000.0 + 10000.0 * next(r)
    return run(grid, seed=k, amp=amp, offset=(ox, oy), noise_seed=k)


def regime_cell(i, j, seed):
    """Regime-map cell (i, j): surface vapour cap 11..16 g/kg in 8 steps and shear 0..35 m/s in 8 steps."""
    return profile(qv_max=(11.0 + 5.0 * i / 7) / 1000.0, shear=5.0 * j)
