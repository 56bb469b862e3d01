"""The oracle's whole step against the reference's model-level tests of the charging and the water budget
(src/core/electrification.test.ts:19-67, src/core/invariants.test.ts:89-96), on the baseline configuration: the
charge separated in a mixed-phase cloud with none created, a charged block carried down by its graupel with its
charge kept, a strongly charged cloud that flashes within the step, no charge with the lightning off, and total water
kept to 0.01 % over a 30-minute ice storm. The reference's QUIET and STORM slider soundings are configurations this
step does not build; the analytic sounding stands in (without a thermal and noise where the reference's is quiet)."""
import math

import numpy as np
import pytest

from meteolab_amd import config as C
from oracle.binding import OracleRun

DT = 3.0


def quiet(g, lightning=True):
    o = OracleRun(g, C.profile(), dict(C.run(g), bubble_amp=0.0, noise_amp=0.0))
    o.set_lightning(lightning)
    return o


class Model:
    def __init__(self, o, g):
        self.o, self.g = o, g
        self.nx, self.ny, self.nz = g["nx"], g["ny"], g["nz"]
        self.layer = self.nx * self.ny
        self.dx, self.dy = g["width"] / self.nx, g["depth"] / self.ny
        self.zs, self.hz = o.levels()
        bs = o.base_state()
        self.rho, self.exner, self.p = bs["rho"], bs["exner"], bs["p"]

    def level_of(self, i):
        return i // self.layer

    def total_charge(self):
        """AtmosphereModel.totalCharge (model.ts): every carrier's charge over the nodes' air mass, C."""
        s = 0.0
        f = [self.o.field(k).astype(np.float64).reshape(self.nz, self.layer) for k in ("chargeIce", "chargeGraupel", "chargeRain", "chargeAir")]
        for z in range(self.nz):
            s += sum(float(a[z].sum()) for a in f) * self.rho[z] * self.dx * self.dy * self.hz[z]
        return s


def mixed_phase(g, lightning=True, z0=5500, z1=7500):
    """electrification.test.ts:7-18: a block of mixed-phase cloud (cloud water, ice, snow, graupel) between z0 and z1,
    saturated over water."""
    o = quiet(g, lightning)
    m = Model(o, g)
    f = {k: o.field(k) for k in ("cloud", "ice", "snow", "graupel", "q", "theta")}
    for z in range(m.nz):
        if not (z0 <= m.zs[z] <= z1):
            continue
        for y in range(m.ny // 2 - 3, m.ny // 2 + 3):
            for x in range(m.nx // 2 - 3, m.nx // 2 + 3):
                i = x + y * m.nx + z * m.layer
                f["cloud"][i], f["ice"][i], f["snow"][i], f["graupel"][i] = 1e-3, 5e-4, 5e-4, 2e-3
                qs = __import__("oracle.binding", fromlist=["lib"]).lib().orc_qsat(float(f["theta"][i]) * m.exner[z] - 273.15, m.p[z])
                f["q"][i] = max(float(f["q"][i]), qs)
    for k, v in f.items():
        o.set_field(k, v)
    return m


def test_collisions_separate_charge_without_creating_any():
    # electrification.test.ts:20-34
    m = mixed_phase(C.STANDARD)
    m.o.step(10, DT)
    cI, cG, cR = (m.o.field(k).astype(np.float64).reshape(m.nz, m.layer) for k in ("chargeIce", "chargeGraupel", "chargeRain"))
    separated = sum(float((np.abs(cI[z]) + np.abs(cG[z]) + np.abs(cR[z])).sum()) * m.rho[z] * m.dx * m.dy * m.hz[z] for z in range(m.nz))
    assert separated > 0
    graupel, ice = float(cG.sum()), float(cI.sum())
    assert abs(graupel) > 0 and math.copysign(1, graupel) == -math.copysign(1, ice)
    assert abs(m.total_charge() + m.o.diag()["chargeGround"]) < 1e-5 * separated


def test_a_charged_block_falls_with_its_carriers_and_keeps_its_charge():
    # electrification.test.ts:36-47
    m = mixed_phase(C.STANDARD, True, 8000, 9500)
    m.o.set_field("cloud", np.zeros(m.o.n, np.float32))
    graupel = m.o.field("graupel")
    charge = np.zeros(m.o.n, np.float32)
    for i in np.nonzero(graupel > 0)[0]:
        charge[i] = 1e-9 / m.rho[m.level_of(i)]
    m.o.set_field("chargeGraupel", charge)
    start = m.total_charge()

    def height():
        c = m.o.field("chargeGraupel").astype(np.float64).reshape(m.nz, m.layer)
        return float((c.sum(axis=1) * m.zs).sum() / c.sum())
    h0 = height()
    m.o.step(20, DT)
    assert abs(m.total_charge() + m.o.diag()["chargeGround"] - start) / start < 1e-4
    assert h0 - height() > 150


def test_a_strongly_charged_cloud_flashes_within_the_step():
    # electrification.test.ts:49-62: negative graupel charge at 6.5 km under positive ice charge at 9 km, 3 nC/m3.
    g = C.STANDARD
    o = quiet(g)
    m = Model(o, g)
    x, y = np.meshgrid(np.arange(m.nx), np.arange(m.ny), indexing="xy")
    r2 = (((x - m.nx / 2) * m.dx) ** 2 + ((y - m.ny / 2) * m.dy) ** 2).reshape(-1)
    ci, cg = np.zeros(o.n, np.float32), np.zeros(o.n, np.float32)
    for z in range(m.nz):
        sl = slice(z * m.layer, (z + 1) * m.layer)
        ci[sl] = 3e-9 * np.exp(-r2 / 2500 ** 2 - ((m.zs[z] - 9000) / 700) ** 2) / m.rho[z]
        cg[sl] = -3e-9 * np.exp(-r2 / 2500 ** 2 - ((m.zs[z] - 6500) / 700) ** 2) / m.rho[z]
    o.set_field("chargeIce", ci)
    o.set_field("chargeGraupel", cg)
    start = m.total_charge()
    o.step(1, DT)
    d = o.diag()
    assert d["flashesIC"] + d["flashesCG"] > 0
    assert d["firstFlash"] == DT
    assert abs(m.total_charge() + d["chargeGround"] - start) < 1e-3


def test_with_the_lightning_off_no_charge_appears():
    # electrification.test.ts:64-70
    m = mixed_phase(C.STANDARD, lightning=False)
    m.o.step(5, DT)
    for k in ("chargeGraupel", "chargeIce"):
        assert not m.o.field(k).any(), k


def test_total_water_is_kept_over_a_thirty_minute_ice_storm():
    # invariants.test.ts:89-96 on the default vertical grid's smaller box (20 x 16 x 24, 15 km, uniform), the storm of
    # the analytic sounding's thermal: ice and graupel form, total water (the air's plus the ground's) kept to 0.01 %.
    g = dict(nx=20, ny=16, nz=24, width=20 * 1200.0, depth=16 * 1125.0, height=15_000.0, bottom=0.0)
    o = OracleRun(g, C.profile(), C.run(g))
    before = o.diag()["water"]
    o.step(600, DT)
    assert o.field("ice").max() > 1e-4 and o.field("graupel").max() > 1e-4
    assert abs(o.diag()["water"] / before - 1) < 1e-4
