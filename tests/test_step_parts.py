"""The parts of a step through the C ABI (ml_dynamics, ml_physics, ml_mix, ml_after_step), split where the reference's
GPU step splits it (src/gpu/step.ts:12-31). The four parts in order are the step; a host that runs the physics and the
lightning itself between the device's dynamics and mixing (gpuStep's model-hook round trip: setCarried, landed,
physicsStep, mix, endGpuStep, afterStep) ends where the oracle's step ends; ml_fall_inputs gives fallInputs' sub-steps."""
import math

import numpy as np
import pytest

import meteolab_amd as M
from meteolab_amd import config as C
from oracle.binding import OracleRun
from test_gpu_step import compare, mature, oracle_restart, rel_l2, restart

pytestmark = pytest.mark.gpu

MIXED = ["u", "v", "w", "theta", "q", "cloud", "ice", "iceNumber"]  # AtmosphereModel.mixed() (model.ts:1338-1344)
FALLS = ["fallRain", "fallSnow", "fallGraupel", "fallHail", "fallChargeIce", "fallChargeGraupel"]


@pytest.fixture(scope="module")
def ctx():
    c = M.Context(0, 32)
    yield c
    c.close()


@pytest.fixture(scope="module")
def storm():
    # 32.5 min into the small-grid storm: every hydrometeor and charge present, hail falling.
    return mature(C.small_grid(), 650)


def test_the_parts_in_order_are_the_step(ctx, storm):
    g = C.small_grid()
    a, b = restart(ctx, storm, g), restart(ctx, storm, g)
    a.step(3)
    for _ in range(3):
        for part in M.STEP_PARTS:
            M.step_part([b], part)
    ctx.sync()
    for name in M.CARRIED + ["fieldZ"]:
        assert np.array_equal(b.field(name), a.field(name)), name
    for name in M.COLUMNS:
        assert np.array_equal(b.column(name), a.column(name)), name
    da, db = a.diagnostics(), b.diagnostics()
    assert db["time"] == da["time"] and db["flashesIC"] == da["flashesIC"] and db["flashesCG"] == da["flashesCG"]


def test_a_host_running_physics_and_lightning_between_the_parts_ends_as_the_oracle(ctx, storm):
    g = C.small_grid()
    ref = oracle_restart(storm, g)
    ref.step(3)
    host, dev = oracle_restart(storm, g), restart(ctx, storm, g)
    for _ in range(3):
        # GpuDynamics.step: the model's carried fields in, sedimentation and transport, fields and ground out.
        for name in M.CARRIED:
            dev.set_field(name, host.field(name))
        for name in M.COLUMNS:
            dev.set_column(name, host.column(name))
        M.step_part([dev], "dynamics")
        for name in M.CARRIED:                      # setCarried
            host.set_field(name, dev.field(name))
        for name in M.COLUMNS:                      # landed (the rain gauge) and the hail swath
            host.set_column(name, dev.column(name))
        for name in FALLS:                          # beginStep's fall speeds (the host's own in gpuStep)
            host.set_field(name, dev.field(name))
        host.step_part(1)                           # physicsStep on the host
        # GpuDynamics.mix: the fields after the physics and the condensate (cloud + ice, model.ts:1528) in.
        for name in M.CARRIED:
            dev.set_field(name, host.field(name))
        cond = (host.field("cloud").astype(np.float64) + host.field("ice")).astype(np.float32)
        dev.set_field("condensate", cond)
        M.step_part([dev], "mix")
        for name in MIXED:                          # endGpuStep
            host.set_field(name, dev.field(name))
        host.step_part(3)                           # afterStep on the host, time += dt
    compare(host, ref)
    assert host.diag()["time"] == ref.diag()["time"]


def test_fall_inputs_are_fallparts_of_the_step(ctx, storm):
    # fallParts (model.ts:575-580): max(1, ceil(max over non-zero nodes of dt rho vt / (levelWeight dz) / 0.8)), with
    # levelWeight dz = hz rho, from the fields at the step's start and the fall speeds the step computed from them.
    g = C.small_grid()
    r = restart(ctx, storm, g)
    start = {name: r.field(name) for name in ("rain", "snow", "graupel", "hail", "chargeIce", "chargeGraupel", "chargeRain")}
    M.step_part([r], "dynamics")
    parts = r.fall_parts()
    zs, hz = storm.levels()
    layer = g["nx"] * g["ny"]
    h = np.repeat(np.asarray(hz), layer)
    pairs = [("rain", "fallRain"), ("snow", "fallSnow"), ("graupel", "fallGraupel"), ("hail", "fallHail"),
             ("chargeIce", "fallChargeIce"), ("chargeGraupel", "fallChargeGraupel"), ("chargeRain", "fallRain")]
    want = []
    for name, fall in pairs:
        vt = r.field(fall).astype(np.float64)
        live = start[name] != 0
        c = float((3.0 * vt[live] / h[live]).max()) if live.any() else 0.0
        want.append(max(1, math.ceil(c / .8)))
    assert parts == want
