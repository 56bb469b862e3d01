"""The ensemble and regime front-ends on the batched step (src/ensemble/member.ts, combine.ts; src/regime/grid.ts).

run_members steps a batch of runs as runMember (member.ts:103-117) steps one: the swaths folded in after every step on
the device (ml_run_set_swaths, addSwaths member.ts:85-100), and every SAMPLE_EVERY model seconds a sample of each run's
strongest values from one launch for the batch (ml_model_diagnostics, diagnostics model.ts:1672-1696) and its rotation
record. The results are MemberResult records (member.ts:17-22) that the reference's combine.ts reads as they are:
member_peaks, EVENTS, probabilities, spread and regime_of restate it here, and probability_map runs on the device
(ml_probability_map, probabilityMap combine.ts:70-75).

The step's sounding is the analytic one (u = shear tanh(z / shear_depth), vapour capped at qv_max), so withMoisture and
withShear act as on the reference's Weisman-Klemp profile (grid.ts:35-36, 57-62): the vapour cap scaled, the hodograph
scaled as a whole to the 0-6 km difference asked. perturb_run restates perturbConfig (member.ts:45-65) with the same
draws in the same order: the ground temperature, the low-level humidity, the wind (errors at 0, 3, 6 and 10 km, as for a
sounding) and the thermal's strength and place. The reference's slider and sounding branches and the CAPE of
environmentOf (a parcel computation) have no counterpart on this sounding.
"""
import math

import numpy as np

from . import config as C

# member.ts:26 and model.ts:65, :136 (UH_ROTATING, MESO_PERSISTENCE, uhScale with MESO_GRID = 1200 m).
SAMPLE_EVERY = 60
UH_ROTATING, UH_MESOCYCLONE, MESO_PERSISTENCE = 200, 400, 600
MESO_GRID = 1200


def uh_scale(dx):
    return MESO_GRID / dx


# combine.ts:12-17: what each probability map counts and from what threshold.
MAPS = {"rain": ("rainTotal", 10), "gust": ("gustMax", 20), "uh": ("uhMax", UH_ROTATING), "hail": ("hailSize", 20)}
# The device's ground column of each member field.
COLUMN_OF = {"rainTotal": "precipitation", "hailSize": "hailSize", "gustMax": "gustMax", "uhMax": "uhMax"}


def run_members(runs, minutes, dt=3.0, progress=None):
    """runMember (member.ts:103-117) for every run of a batch (one pool) at once: MemberResult dicts in run order."""
    import meteolab_amd as M
    for r in runs:
        r.set_swaths(True)
    samples = [[] for _ in runs]
    flashes = [0] * len(runs)

    def sample():
        ds = M.model_diagnostics(runs)
        for k, (r, d) in enumerate(zip(runs, ds)):
            rot, m = r.rotation(), r.diagnostics()
            total = int(m["flashesIC"] + m["flashesCG"])
            samples[k].append({"t": m["time"], "updraft": d["updraft"], "uh": rot["uh"], "rain": d["rainRate"],
                               "gust": d["gust"], "hail": d["hail"], "flashes": (total - flashes[k]) * 60 / SAMPLE_EVERY,
                               "rotationHeld": rot["persisted"], "cores": d["cores"], "outflow": d["outflow"],
                               "outflowLow": d["outflowLow"]})
            flashes[k] = total
        if progress:
            progress(samples[0][-1]["t"])

    sample()
    per = int(round(SAMPLE_EVERY / dt))
    end = minutes * 60
    t = 0.0
    while t < end - 1e-6:
        n = min(per, int(math.ceil((end - t) / dt - 1e-9)))
        M.step(runs, n, dt)
        t += n * dt
        runs[0].ctx.sync()
        sample()
    g = runs[0].grid
    return [{"seed": int(r.spec["run"]["seed"]) if hasattr(r, "spec") else None, "samples": samples[k],
             "nx": g["nx"], "ny": g["ny"], "width": g["width"], "depth": g["depth"],
             "rainTotal": r.column("precipitation"), "hailSize": r.column("hailSize"),
             "gustMax": r.column("gustMax"), "uhMax": r.column("uhMax")} for k, r in enumerate(runs)]


def member_peaks(r):
    """memberPeaks (combine.ts:23-28): a member's highlights over its run."""
    def peak(k):
        return max([0.0] + [s[k] for s in r["samples"]])
    meso = next((s for s in r["samples"] if s["rotationHeld"] >= MESO_PERSISTENCE), None)
    return {"updraft": peak("updraft"), "uh": peak("uh"), "rain": peak("rain"),
            "rainTotal": float(max([0.0] + [float(x) for x in r["rainTotal"]])), "gust": peak("gust"), "hail": peak("hail"),
            "flashes": peak("flashes"), "mesocyclone": meso["t"] - meso["rotationHeld"] if meso else None,
            "outflow": max([0.0] + [s.get("outflow") or 0.0 for s in r["samples"]]),
            "outflowLow": max([0.0] + [s.get("outflowLow") or 0.0 for s in r["samples"]])}


# combine.ts:31-38: the events counted, each with its test on a member's highlights.
EVENTS = {
    "storm": lambda p: p["updraft"] >= 10 and p["rain"] >= 1,
    "mesocyclone": lambda p: p["mesocyclone"] is not None,
    "hail": lambda p: p["hail"] >= MAPS["hail"][1],
    "gust": lambda p: p["gust"] >= MAPS["gust"][1],
    "downpour": lambda p: p["rain"] >= 50,
    "total": lambda p: p["rainTotal"] >= 20,
}


def probabilities(peaks):
    """probabilities (combine.ts:41-45): the share of members in which each event happened."""
    return {k: (sum(1 for p in peaks if test(p)) / len(peaks) if peaks else 0.0) for k, test in EVENTS.items()}


def spread(values):
    """spread (combine.ts:48-52): least, median and most (NaN when there are none)."""
    if not values:
        return {"min": math.nan, "median": math.nan, "max": math.nan}
    v = sorted(values)
    mid = len(v) >> 1
    return {"min": v[0], "median": v[mid] if len(v) % 2 else (v[mid - 1] + v[mid]) / 2, "max": v[-1]}


def map_threshold(kind, dx):
    """The threshold of a probability map on a grid of spacing dx (helicity's scaled, combine.ts:72)."""
    field, threshold = MAPS[kind]
    return threshold * uh_scale(dx) if kind == "uh" else threshold


def probability_map(runs, kind):
    """probabilityMap (combine.ts:70-75) of the runs' swaths on the device (ml_probability_map)."""
    import meteolab_amd as M
    g = runs[0].grid
    return M.probability_map(runs, COLUMN_OF[MAPS[kind][0]], map_threshold(kind, g["width"] / g["nx"]))


# grid.ts:97: an ordinary cell's life, s.
CELL_LIFE = 40 * 60


def regime_of(r, p=None):
    """regimeOf (grid.ts:98-104): none, supercell, multicell or ordinary from a run's course."""
    p = p or member_peaks(r)
    if not EVENTS["storm"](p):
        return "none"
    if p["mesocyclone"] is not None:
        return "supercell"
    end = r["samples"][-1]["t"]
    late = max([0.0] + [s["updraft"] for s in r["samples"] if s["t"] >= max(CELL_LIFE, end * 2 / 3)])
    return "multicell" if late >= max(10, p["updraft"] / 2) else "ordinary"


# grid.ts:14-17 and :21: the grids (vapour factors by row, 0-6 km shears by column) and the calm limit.
GRIDS = {"small": {"moisture": [.9, 1, 1.1], "shear": [0, 20, 40]},
         "full": {"moisture": [.8, .9, 1, 1.1], "shear": [0, 10, 20, 30, 40]}}
CALM = 2


WIND_NODES = (0.0, 3000.0, 6000.0, 10000.0)


def wind_error(d, z):
    """The wind error at height z from errors d at WIND_NODES (member.ts:56): linear between, held beyond."""
    k = 0
    while k < 2 and z >= WIND_NODES[k + 1]:
        k += 1
    return d[k] + (d[k + 1] - d[k]) * min(max((z - WIND_NODES[k]) / (WIND_NODES[k + 1] - WIND_NODES[k]), 0.0), 1.0)


def wind(profile, z):
    """The sounding's wind (u, v) at height z: the tanh profile plus its wind errors."""
    zero = (0.0,) * 4
    return (profile["shear"] * math.tanh(z / profile["shear_depth"]) + wind_error(profile.get("wind_u", zero), z),
            wind_error(profile.get("wind_v", zero), z))


def shear06(profile):
    """The sounding's 0-6 km wind difference, m/s (the shear of environmentOf, grid.ts:24-29; model.ts:1682)."""
    (u0, v0), (u6, v6) = wind(profile, 0.0), wind(profile, 6000.0)
    return math.hypot(u6 - u0, v6 - v0)


def with_moisture(profile, factor):
    """withMoisture (grid.ts:32-36) on the analytic sounding: its vapour cap times `factor`."""
    return dict(profile, qv_max=profile["qv_max"] * factor)


def with_shear(profile, shear):
    """withShear (grid.ts:49-80) on the analytic sounding, to a 0-6 km difference of `shear` m/s with the hodograph's
    shape kept. Without wind errors it is the Weisman-Klemp hodograph, calm at the ground: scaled as a whole (a calm
    one stays calm). With wind errors it is a sounding's: the difference from the ground wind scaled, or, if calm, a
    straight westerly shear over 0-6 km added; both stay a tanh profile plus errors at the wind nodes."""
    now = shear06(profile)
    if "wind_u" not in profile and "wind_v" not in profile:
        return dict(profile, shear=profile["shear"] * (shear / now if now > CALM else 0.0))
    f, added = (shear / now, 0.0) if now > CALM else (1.0, float(shear))
    zero = (0.0,) * 4
    eu, ev = profile.get("wind_u", zero), profile.get("wind_v", zero)
    ramp = [min(z, 6000.0) / 6000 for z in WIND_NODES]
    return dict(profile, shear=profile["shear"] * f, wind_u=[eu[0] + (e - eu[0]) * f + added * r for e, r in zip(eu, ramp)],
                wind_v=[ev[0] + (e - ev[0]) * f for e in ev])


def regime_cells(profile, size):
    """regimeCells (grid.ts:69-75): the grid's cells row by row (driest first), each with its sounding and shear."""
    g = GRIDS[size]
    return [{"row": i, "column": j, "profile": with_shear(with_moisture(profile, m), s), "shear": shear06(with_shear(with_moisture(profile, m), s))}
            for i, m in enumerate(g["moisture"]) for j, s in enumerate(g["shear"])]


# member.ts:33: the starting data's uncertainty (standard deviations): K, % relative humidity, m/s, degrees, the
# thermal's relative strength, m.
UNCERTAINTY = {"temperature": .5, "humidity": 4, "wind": 1.5, "direction": 8, "thermal": .1, "offset": 1500}


def normal(rng):
    """member.ts:36: normal deviates by Box-Muller from a generator of uniforms in [0, 1), two uniforms a deviate (the
    first under the logarithm)."""
    def draw():
        a = next(rng)
        return math.sqrt(-2 * math.log(1 - a)) * math.cos(2 * math.pi * next(rng))
    return draw


# The wind slider nodes of perturbConfig's loop (member.ts:50) in its order, and those SUMMER_DAY-based configurations
# carry (src/core/fixtures.ts:9: wind0, 3, 6, 10 with their directions).
SLIDER_KEYS = ("0", "05", "1", "3", "6", "10")
SLIDER_NODES = ("0", "3", "6", "10")


def perturb_run(grid, profile, spec, sliders=SLIDER_NODES):
    """perturbConfig (member.ts:45-65) on the analytic sounding and a run: new (profile, spec) dicts, the inputs
    untouched. From the run's own seed (mulberry32(seed ^ 0x5bd1e995)) in the reference's order: the ground temperature
    error dT (the sounding's theta0 moves by dT, as surfaceTemp does), the humidity error dRh in % (the vapour cap,
    which sets the low-level vapour of this sounding, times 1 + dRh / 100, as a sounding's vapour at the ground), the
    wind errors at 0, 3, 6 and 10 km (u then v, added to the sounding's wind errors, as a sounding's are), a speed and a
    direction error for each wind slider node the emulated configuration carries (`sliders`; drawn and not applied, as
    the analytic sounding has no sliders), the thermal's strength (times max(.5, 1 + .1 n)) and its place (startOffset:
    from the thermal's default place, replacing any offset the run had)."""
    n, U = normal(C.mulberry32(int(spec["seed"]) ^ 0x5bd1e995)), UNCERTAINTY
    dT, dRh = n() * U["temperature"], n() * U["humidity"]
    du = [n() * U["wind"] for _ in WIND_NODES]
    dv = [n() * U["wind"] for _ in WIND_NODES]
    for key in SLIDER_KEYS:
        if key in sliders:
            n(), n()
    zero = (0.0,) * 4
    prof = dict(profile, theta0=profile["theta0"] + dT, qv_max=profile["qv_max"] * (1 + dRh / 100),
                wind_u=[a + b for a, b in zip(profile.get("wind_u", zero), du)],
                wind_v=[a + b for a, b in zip(profile.get("wind_v", zero), dv)])
    amp = spec["bubble_amp"] * max(.5, 1 + n() * U["thermal"])
    ox = n() * U["offset"]
    oy = n() * U["offset"]
    return prof, dict(spec, bubble_amp=amp, bubble_x=grid["width"] / 3 + ox, bubble_y=grid["depth"] / 2 + oy)


def ensemble(grid, members, minutes, ctx, profile=None, data=False):
    """An ensemble of `members` runs (the ensemble configuration's members, config.ensemble_member) stepped together:
    their MemberResults, the event probabilities and the four probability maps. With `data`, every member but the
    first (the control) has its starting data perturbed by perturb_run, as the reference's "seed and data" ensembles
    (panel.ts:110)."""
    pool = ctx.pool(grid, members)
    prof = profile or C.profile()
    starts = [(prof, C.ensemble_member(grid, k)) for k in range(members)]
    if data:
        starts = [s if k == 0 else perturb_run(grid, *s) for k, s in enumerate(starts)]
    runs = pool.runs(starts)
    results = run_members(runs, minutes)
    peaks = [member_peaks(r) for r in results]
    maps = {kind: probability_map(runs, kind) for kind in MAPS}
    return {"results": results, "peaks": peaks, "probabilities": probabilities(peaks), "maps": maps, "runs": runs}
