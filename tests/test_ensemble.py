"""The ensemble and regime front-ends (src/ensemble, src/regime): on the CPU the reference's own tests of addSwaths,
the probabilities, spread and probability maps (ensemble.test.ts:21-31, 58-74) and regimeOf (grid.test.ts:39-52), on
the oracle and the host logic; the analytic sounding's withMoisture / withShear; on the GPU a batch of members run
together (swaths folded in on the device, samples from one launch) against the oracle's runMember, and the device's
probability maps against the restated probabilityMap."""
import math

import numpy as np
import pytest

from meteolab_amd import config as C
from meteolab_amd import ensemble as E
from oracle.binding import OracleRun


def sample(t, **o):
    s = {"t": t, "updraft": 0, "uh": 0, "rain": 0, "gust": 0, "hail": 0, "flashes": 0, "rotationHeld": 0, "cores": 0}
    s.update(o)
    return s


def member(samples, **maps):
    """A 2 x 2 member with the given highlights and swaths (ensemble.test.ts:8-12)."""
    f = {k: np.asarray(maps.get(k, [0, 0, 0, 0]), np.float32) for k in ("rainTotal", "hailSize", "gustMax", "uhMax")}
    return {"seed": 1, "samples": samples, "nx": 2, "ny": 2, "width": 2000, "depth": 2000, **f}


def host_probability_map(results, kind):
    """probabilityMap (combine.ts:70-75) restated on the host, as the check of the device's."""
    field, threshold = E.MAPS[kind]
    out = np.zeros(len(results[0]["rainTotal"]), np.float32)
    for r in results:
        at = threshold * E.uh_scale(r["width"] / r["nx"]) if kind == "uh" else threshold
        out += (np.asarray(r[field], np.float64) >= at).astype(np.float32)
    return (out.astype(np.float64) / len(results)).astype(np.float32)


def test_probabilities_spread_and_maps():
    # ensemble.test.ts:58-74.
    results = [
        member([sample(0), sample(600, updraft=30, rain=60, rotationHeld=E.MESO_PERSISTENCE)], rainTotal=[20, 0, 0, 0], gustMax=[25, 0, 0, 0]),
        member([sample(0), sample(600, updraft=20, rain=5, gust=22)], rainTotal=[12, 3, 0, 0]),
        member([sample(0), sample(600, updraft=4)]),
        member([sample(0), sample(600, updraft=15, rain=2, hail=25)], rainTotal=[0, 11, 0, 0]),
    ]
    peaks = [E.member_peaks(r) for r in results]
    p = E.probabilities(peaks)
    assert peaks[0]["mesocyclone"] == 0 and peaks[1]["mesocyclone"] is None
    assert [p["storm"], p["mesocyclone"], p["downpour"], p["hail"], p["gust"], p["total"]] == [.75, .25, .25, .25, .25, .25]
    assert E.spread([x["updraft"] for x in peaks]) == {"min": 4, "median": 17.5, "max": 30}
    assert list(host_probability_map(results, "rain")) == [.5, .25, 0, 0]
    assert list(host_probability_map(results, "gust")) == [.25, 0, 0, 0]


def test_regime_of_a_runs_course():
    # grid.test.ts:39-52.
    def run(updrafts, every=600, **o):
        return {"seed": 1, "nx": 1, "ny": 1, "width": 1, "depth": 1, "rainTotal": np.zeros(1, np.float32),
                "samples": [sample(k * every, updraft=w, rain=20 if w > 5 else 0, cores=1, **o) for k, w in enumerate(updrafts)]}
    assert E.regime_of(run([0, 4, 6, 3])) == "none"
    assert E.regime_of(run([0, 30, 12, 2, 1])) == "ordinary"
    assert E.regime_of(run([0, 30, 12, 20, 18])) == "multicell"
    assert E.regime_of(run([0, 30, 12, 20, 18], rotationHeld=E.MESO_PERSISTENCE)) == "supercell"
    calm = "2 1 3 4 6 8 10 13 16 21 25 30 37 44 52 60 69 68 62 57 54 55 57 54 56 57 53 52 54 51 46 45 47 41 36 30 27 28 30 30 27 21 15 10 7 6"
    sheared = "2 1 3 4 5 7 9 11 14 17 21 24 28 33 38 42 49 54 57 57 52 52 51 49 51 48 43 40 42 45 51 50 45 39 33 35 37 41 42 41 38 37 37 39 36 32"
    assert E.regime_of(run([float(x) for x in calm.split()], every=60)) == "ordinary"
    assert E.regime_of(run([float(x) for x in sheared.split()], every=60)) == "multicell"


def test_moisture_and_shear_of_the_analytic_sounding():
    # grid.test.ts:11-37 on this sounding: the 0-6 km shear asked is made and the shape (shear depth) kept; a calm
    # sounding stays calm (the Weisman-Klemp branch); the vapour cap scales and the shear stays.
    prof = C.profile()
    for s in (0, 15, 40):
        assert E.shear06(E.with_shear(prof, s)) == pytest.approx(s, abs=1e-9)
    assert E.with_shear(prof, 15)["shear_depth"] == prof["shear_depth"]
    assert E.shear06(E.with_shear(C.profile(shear=1.0), 20)) == 0
    assert E.with_moisture(prof, .75)["qv_max"] == pytest.approx(.0105, abs=1e-12)
    assert E.shear06(E.with_moisture(prof, 1.1)) == E.shear06(prof)
    cells = E.regime_cells(prof, "full")
    g = E.GRIDS["full"]
    assert len(cells) == len(g["moisture"]) * len(g["shear"])
    for c in cells:
        assert c["shear"] == pytest.approx(g["shear"][c["column"]], abs=1e-9)
        assert c["profile"]["qv_max"] == pytest.approx(prof["qv_max"] * g["moisture"][c["row"]])


def test_oracle_folds_columns_into_the_ground_shifted_by_the_domain_motion():
    # ensemble.test.ts:21-31 on the small grid: the domain moving 5 m/s east for 480 s has moved 2 columns (2.4 km at
    # dx = 1200 m): its column 1, with u = 10 m/s over the domain, lies over ground column 3 (15 m/s over the ground);
    # a column at rest over the domain has the domain's 5 m/s.
    # Aloft, the rotation tests' cyclonic and anticyclonic updraught patches give helicity columns of both signs.
    from test_rotation import planted
    g = C.small_grid()
    nx, layer = g["nx"], g["nx"] * g["ny"]
    o = OracleRun(g, C.profile(), dict(C.run(g), frame_u=5.0, frame_v=0.0))
    u, v, w = planted(g, "pair")
    u[:layer] = 0
    v[:layer] = 0
    u[1] = 10
    o.set_field("u", u); o.set_field("v", v); o.set_field("w", w)
    o.update_rotation(0.0, keep=True)
    o.set_time(480.0)
    o.add_swaths()
    gust, uh, col = o.column("gustMax"), o.column("uhMax"), o.uh_column()
    assert gust[3] == 15 and gust[0] == 5
    assert np.all(gust[np.arange(g["nx"] * g["ny"]) != 3] == 5)
    shift = (np.arange(g["nx"] * g["ny"]) % nx + 2) % nx + nx * (np.arange(g["nx"] * g["ny"]) // nx)
    expect = np.zeros_like(uh)
    expect[shift] = np.maximum(col, 0)
    assert col.max() > 0 and col.min() < 0
    assert np.array_equal(uh, expect)


def test_normal_deviates_of_the_members_generator():
    # member.ts:36 (Box-Muller): mean 0, variance 1, and the normal law's 68.27 % within one deviation and 95.45 %
    # within two, over 40,000 deviates of one stream.
    n = E.normal(C.mulberry32(0x5bd1e995 ^ 7))
    x = np.array([n() for _ in range(40_000)])
    assert abs(x.mean()) < .02 and abs(x.std() - 1) < .02
    assert abs(np.mean(np.abs(x) < 1) - .6827) < .01 and abs(np.mean(np.abs(x) < 2) - .9545) < .005


def test_perturbed_starting_data_from_the_members_seed():
    # ensemble.test.ts:32-40: one seed gives one perturbation, the settings given untouched; another seed another; over
    # 400 seeds the deviations have about the standard deviations asked (the reference's tolerances).
    g, prof = C.small_grid(), C.profile()
    base = C.run(g, seed=5)
    kept = (dict(prof), dict(base))
    a = E.perturb_run(g, prof, base)
    assert a == E.perturb_run(g, prof, dict(base)) and (prof, base) == kept
    assert E.perturb_run(g, prof, C.run(g, seed=6))[0]["theta0"] != a[0]["theta0"]
    many = [E.perturb_run(g, prof, C.run(g, seed=k + 1)) for k in range(400)]

    def sd(v):
        return math.sqrt(sum(x * x for x in v) / len(v))
    U = E.UNCERTAINTY
    assert abs(sd([p["wind_u"][2] for p, _ in many]) / E.UNCERTAINTY["wind"] - 1) < .5
    assert abs(sd([p["wind_v"][0] for p, _ in many]) / E.UNCERTAINTY["wind"] - 1) < .5
    assert abs(sd([p["theta0"] - prof["theta0"] for p, _ in many]) - U["temperature"]) < .05
    assert abs(sd([100 * (p["qv_max"] / prof["qv_max"] - 1) for p, _ in many]) / U["humidity"] - 1) < .5
    assert abs(sd([s["bubble_x"] - g["width"] / 3 for _, s in many]) / U["offset"] - 1) < .5
    assert abs(sd([s["bubble_y"] - g["depth"] / 2 for _, s in many]) / U["offset"] - 1) < .5
    strength = [s["bubble_amp"] / 2.0 for _, s in many]
    assert min(strength) >= .5 and abs(sd([x - 1 for x in strength]) / U["thermal"] - 1) < .5
    assert all(p["shear"] == prof["shear"] and p["dtheta"] == prof["dtheta"] for p, _ in many)
    assert "wind_u" not in prof and E.perturb_run(g, dict(prof, wind_u=[1, 1, 1, 1]), base)[0]["wind_u"] == [1 + x for x in a[0]["wind_u"]]


def test_perturbed_thermal_takes_its_draws_after_the_slider_winds():
    # perturbConfig (member.ts:45-65) draws, from mulberry32(seed ^ 0x5bd1e995), dT and dRh, the 4 + 4 wind errors, then
    # a speed and a direction for each wind slider the configuration has (SUMMER_DAY: wind0, 3, 6, 10 and windDir0, 3, 6,
    # 10; fixtures.ts:9), then the thermal's strength and the two offsets: deviates 18, 19 and 20 of the stream. Without
    # sliders they are deviates 10, 11 and 12.
    g, prof = C.small_grid(), C.profile()
    for seed in (1, 7, 123456):
        spec = C.run(g, seed=seed)
        draw = E.normal(C.mulberry32(seed ^ 0x5bd1e995))
        z = [draw() for _ in range(21)]
        U = E.UNCERTAINTY
        for sliders, at in ((E.SLIDER_NODES, 18), ((), 10), (("0", "05", "1", "3", "6", "10"), 22)):
            if at + 2 >= len(z):
                z += [draw() for _ in range(at + 3 - len(z))]
            p, s = E.perturb_run(g, prof, spec, sliders=sliders)
            assert p["theta0"] == prof["theta0"] + z[0] * U["temperature"]
            assert p["wind_v"][3] == z[9] * U["wind"]
            assert s["bubble_amp"] == spec["bubble_amp"] * max(.5, 1 + z[at] * U["thermal"])
            assert s["bubble_x"] == g["width"] / 3 + z[at + 1] * U["offset"]
            assert s["bubble_y"] == g["depth"] / 2 + z[at + 2] * U["offset"]
    assert E.perturb_run(g, prof, C.run(g, seed=3)) == E.perturb_run(g, prof, C.run(g, seed=3), sliders=("0", "3", "6", "10"))


WINDY ={"wind_u": [1.0, 2.0, -1.0, 3.0], "wind_v": [-2.0, 0.0, 1.0, .5]}


def test_wind_errors_reach_the_sounding():
    # The errors at 0, 3, 6 and 10 km: the environment's wind less the tanh profile lies on the straight line through
    # them in each interval (fitted on the levels there), is held above 10 km, and the 0-6 km shear is the closed form.
    g = dict(nx=8, ny=8, nz=50, width=9600.0, depth=9000.0, height=18900.0, bottom=100.0)
    prof = dict(C.profile(), **WINDY)
    o = OracleRun(g, prof, C.run(g))
    bs = o.base_state()
    zs, _ = o.levels()
    err = {"wind_u": bs["u"] - 25 * np.tanh(zs / 3000), "wind_v": bs["v"]}
    for key, d in WINDY.items():
        for k, (lo, hi) in enumerate(((0, 3000), (3000, 6000), (6000, 10000))):
            at = (zs >= lo) & (zs <= hi)
            slope, icpt = np.polyfit(zs[at], err[key][at], 1)
            assert icpt + slope * lo == pytest.approx(d[k], abs=1e-9) and icpt + slope * hi == pytest.approx(d[k + 1], abs=1e-9)
        assert np.allclose(err[key][zs >= 10000], d[3], atol=1e-12)
    want = math.hypot(25 * math.tanh(2.0) - 1.0 - 1.0, 1.0 + 2.0)
    assert E.shear06(prof) == pytest.approx(want, rel=1e-14)
    assert o.model_diagnostics()["shear06"] == pytest.approx(want, rel=1e-14)
    # withShear (grid.ts:49-80) of a sounding: the difference from the ground wind scaled, the ground wind kept; a calm
    # one gets a straight westerly 0-6 km shear, a quarter of it per 1.5 km.
    for s in (0, 15, 40):
        scaled = E.with_shear(prof, s)
        assert E.shear06(scaled) == pytest.approx(s, abs=1e-12) and E.wind(scaled, 0) == pytest.approx(E.wind(prof, 0), abs=1e-12)
    calm = dict(C.profile(shear=0.0), wind_u=[1.0] * 4, wind_v=[-1.0] * 4)
    made = E.with_shear(calm, 20)
    for z in (0, 1500, 3000, 6000, 9000):
        assert E.wind(made, z) == pytest.approx((1 + 20 * min(z, 6000) / 6000, -1), abs=1e-12)


def test_the_thermal_moves_with_its_offset():
    # ensemble.test.ts:46-51: the warmest node of the thermal's level moves by the offset, (4 dx, 2 dy) -> (4, 2).
    g = C.small_grid()
    nx, layer = g["nx"], g["nx"] * g["ny"]
    dx, dy = g["width"] / nx, g["depth"] / g["ny"]

    def warmest(spec):
        o = OracleRun(g, C.profile(), spec)
        zs, _ = o.levels()
        k = int(np.argmin(np.abs(np.asarray(zs) - 1400)))
        c = int(np.argmax(o.field("theta")[k * layer:(k + 1) * layer]))
        return c % nx, c // nx

    plain = C.run(g, seed=3, noise_amp=0.0)
    x0, y0 = warmest(plain)
    x1, y1 = warmest(dict(plain, bubble_x=plain["bubble_x"] + 4 * dx, bubble_y=plain["bubble_y"] + 2 * dy))
    assert (x1 - x0, y1 - y0) == (4, 2)


def oracle_member(g, prof, spec, minutes, dt=3.0):
    """runMember (member.ts:103-117) on the oracle: samples every 60 s and the swaths."""
    o = OracleRun(g, prof, spec)
    o.set_swaths(True)
    samples, flashes = [], 0

    def take():
        nonlocal flashes
        d, rot, m = o.model_diagnostics(), o.rotation(), o.diag()
        total = int(m["flashesIC"] + m["flashesCG"])
        samples.append({"t": m["time"], "updraft": d["updraft"], "uh": rot["uh"], "rain": d["rainRate"], "gust": d["gust"],
                        "hail": d["hail"], "flashes": (total - flashes) * 60 / E.SAMPLE_EVERY, "rotationHeld": rot["persisted"],
                        "cores": d["cores"], "outflow": d["outflow"], "outflowLow": d["outflowLow"]})
        flashes = total

    take()
    for _ in range(int(minutes)):
        o.step(int(round(E.SAMPLE_EVERY / dt)), dt)
        take()
    return o, {"samples": samples, "nx": g["nx"], "ny": g["ny"], "width": g["width"], "depth": g["depth"],
               **{k: o.column(E.COLUMN_OF[k]) for k in ("rainTotal", "hailSize", "gustMax", "uhMax")}}


@pytest.mark.gpu
def test_members_stepped_together_sample_and_swath_as_the_oracle():
    import meteolab_amd as M
    g = C.small_grid()
    prof = C.profile()
    specs = [dict(C.ensemble_member(g, k), frame_u=4.0, frame_v=-2.0) for k in range(3)]
    minutes = 12
    with M.Context(0, 32) as ctx:
        pool = ctx.pool(g, 3)
        runs = [pool.run(prof, s) for s in specs]
        print("members on the device,", minutes, "minutes", flush=True)
        results = E.run_members(runs, minutes)
        for k, spec in enumerate(specs):
            _, want = oracle_member(g, prof, spec, minutes)
            got = results[k]
            assert [s["t"] for s in got["samples"]] == [60.0 * j for j in range(minutes + 1)]
            for a, b in zip(got["samples"], want["samples"]):
                for key in ("updraft", "uh", "rain", "gust", "outflow"):
                    assert a[key] == pytest.approx(b[key], rel=1e-3, abs=1e-6), (k, a["t"], key, a[key], b[key])
                assert a["flashes"] == b["flashes"] and a["rotationHeld"] == b["rotationHeld"], (k, a["t"])
            for key in ("rainTotal", "gustMax", "uhMax", "hailSize"):
                x, y = np.asarray(got[key], np.float64), np.asarray(want[key], np.float64)
                ny_ = np.linalg.norm(y)
                err = np.linalg.norm(x - y) / ny_ if ny_ > 0 else np.linalg.norm(x)
                assert err <= 1e-4, (k, key, err)
        assert max(np.max(r["gustMax"]) for r in results) > 0 and max(np.max(r["uhMax"]) for r in results) > 0
        for kind in E.MAPS:
            dev = E.probability_map(runs, kind)
            assert np.array_equal(dev, host_probability_map(results, kind)), kind


@pytest.mark.gpu
def test_a_seed_and_data_ensemble_steps_its_perturbed_members_as_the_oracle():
    # The control keeps its starting data; the others run the perturbed soundings and thermals as the oracle does.
    import meteolab_amd as M
    g, prof = C.small_grid(), C.profile()
    with M.Context(0, 32) as ctx:
        print("seed and data ensemble, 3 members, 2 minutes", flush=True)
        out = E.ensemble(g, 3, 2, ctx, data=True)
    for k in range(3):
        p, spec = (prof, C.ensemble_member(g, k)) if k == 0 else E.perturb_run(g, prof, C.ensemble_member(g, k))
        _, want = oracle_member(g, p, spec, 2)
        for a, b in zip(out["results"][k]["samples"], want["samples"]):
            for key in ("updraft", "rain", "gust", "outflow"):
                assert a[key] == pytest.approx(b[key], rel=1e-3, abs=1e-6), (k, a["t"], key)
    assert out["results"][1]["samples"][-1]["updraft"] != out["results"][0]["samples"][-1]["updraft"]


@pytest.mark.gpu
def test_device_sounding_with_wind_errors_is_the_oracles():
    import meteolab_amd as M
    g = C.small_grid()
    prof = dict(C.profile(), **WINDY)
    o = OracleRun(g, prof, C.run(g))
    with M.Context(0, 32) as ctx:
        print("device sounding with wind errors", flush=True)
        r = ctx.pool(g, 1).run(prof, C.run(g))
        lv, md = r.levels(), r.model_diagnostics()
    bs = o.base_state()
    for k in ("u", "v"):
        assert np.allclose(lv[k], bs[k], rtol=1e-12, atol=1e-12), k
    assert abs(bs["v"]).max() > 0
    assert md["shear06"] == pytest.approx(o.model_diagnostics()["shear06"], rel=1e-14)
