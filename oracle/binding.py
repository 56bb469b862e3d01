"""ctypes binding of the oracle library (liboracle.so).

TEST INFRASTRUCTURE ONLY: the checker the HIP step is compared with. Imported by tests/, by
``__graft_entry__.smoke()`` and by ``bench.py``'s CPU-baseline leg; the product package never imports it.

The oracle restates the reference's ``AtmosphereModel.step`` (src/core/model.ts:873-877) in C++ with float32
fields and float64 arithmetic; see stormlab_oracle.cpp for the file:line each part follows.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

# Field ids (oracle.h): the carried order of AtmosphereModel.carried() (model.ts:1319-1335), then extra arrays.
CARRIED = ["rain", "snow", "graupel", "hail", "u", "v", "w", "theta", "q", "cloud", "ice", "iceNumber", "cold",
           "chargeIce", "chargeGraupel", "chargeRain", "chargeAir"]
EXTRA = ["fieldZ", "fallRain", "fallSnow", "fallGraupel", "fallHail", "fallChargeIce", "fallChargeGraupel"]
FIELDS = CARRIED + EXTRA
FIELD_ID = {name: k for k, name in enumerate(FIELDS)}
COLUMNS = ["precipitation", "hailGround", "hailSize", "gustMax", "uhMax"]
INTERVENTIONS = ["thermal", "warm", "cool", "moisten", "dry", "coldPool", "vortex"]
# ModelDiagnostics (model.ts:161-180) in the order orc_model_diagnostics writes it.
MODEL_DIAGNOSTICS = ["updraft", "downdraft", "rainRate", "rainTotal", "cloudTop", "thermalTop", "maxCloud", "coldMax",
                     "cores", "shear06", "outflow", "outflowLow", "gust", "hail", "clipped"]
# RotationState (model.ts:142-158) in the order orc_rotation writes it.
ROTATION = ["uh", "x", "y", "anticyclonic", "persisted", "uh01", "uh03", "low", "coreZeta", "coreUH", "coreAnti"]
DIAG = ["time", "maxW", "maxThetaP", "rainTotal", "firstFlash", "flashesIC", "flashesCG", "clipped", "negativeFilled",
        "chargeGround", "minW", "water"]


class Grid(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int), ("width", ctypes.c_double),
                ("depth", ctypes.c_double), ("height", ctypes.c_double), ("bottom", ctypes.c_double)]


class Profile(ctypes.Structure):
    _fields_ = [(k, ctypes.c_double) for k in
                ("theta0", "dtheta", "ztr", "ttr", "rh_drop", "rh_top", "qv_max", "shear", "shear_depth")] + \
        [("wind_u", ctypes.c_double * 4), ("wind_v", ctypes.c_double * 4)]


class Run(ctypes.Structure):
    _fields_ = [("seed", ctypes.c_int), ("latitude", ctypes.c_double), ("turbulence", ctypes.c_double),
                ("tropopause_km", ctypes.c_double), ("roughness", ctypes.c_double), ("frame_u", ctypes.c_double),
                ("frame_v", ctypes.c_double), ("bubble_amp", ctypes.c_double), ("bubble_x", ctypes.c_double),
                ("bubble_y", ctypes.c_double), ("bubble_z", ctypes.c_double), ("bubble_rh", ctypes.c_double),
                ("bubble_rz", ctypes.c_double), ("noise_seed", ctypes.c_int), ("noise_amp", ctypes.c_double),
                ("noise_top", ctypes.c_double)]


_libs = {}


def lib():
    """Loads liboracle.so (built by ``make -C oracle``)."""
    return load(LIB_PATH)


def load(path):
    """An oracle library at `path` with its C ABI's signatures (the checks of tests/test_oracle_physics.py run on a
    library built with a planted mistake this way)."""
    if path not in _libs:
        if not os.path.exists(path):
            raise RuntimeError(f"oracle library missing: {path} (run make -C oracle)")
        L = ctypes.CDLL(path)
        vp, d, i = ctypes.c_void_p, ctypes.c_double, ctypes.c_int
        fp, dp = ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_double)
        L.orc_create.restype = vp
        L.orc_create.argtypes = [ctypes.POINTER(Grid), ctypes.POINTER(Profile), ctypes.POINTER(Run)]
        L.orc_destroy.argtypes = [vp]
        L.orc_step.argtypes = [vp, d, i]
        L.orc_step_part.argtypes = [vp, i, d]
        L.orc_step_part.restype = i
        L.orc_set_lightning.argtypes = [vp, i]
        for name in ("orc_get_field", "orc_get_column"):
            getattr(L, name).argtypes = [vp, i, fp]
            getattr(L, name).restype = i
        for name in ("orc_set_field", "orc_set_column"):
            getattr(L, name).argtypes = [vp, i, fp]
            getattr(L, name).restype = i
        L.orc_base_state.argtypes = [vp, dp]
        L.orc_grid_levels.argtypes = [vp, dp, dp]
        L.orc_diag.argtypes = [vp, dp]
        L.orc_pressure.argtypes = [vp, dp]
        L.orc_set_time.argtypes = [vp, d]
        L.orc_intervene.argtypes = [vp, i, d, d, d, d, dp]
        L.orc_intervene.restype = i
        L.orc_added.argtypes = [vp, dp]
        L.orc_rotation.argtypes = [vp, dp]
        L.orc_update_rotation.argtypes = [vp, d, i]
        L.orc_model_diagnostics.argtypes = [vp, dp]
        L.orc_set_swaths.argtypes = [vp, i]
        L.orc_add_swaths.argtypes = [vp]
        L.orc_uh_column.argtypes = [vp, fp]
        L.orc_uh_column.restype = i
        L.orc_outflow_deficit.argtypes = [vp, fp, dp]
        L.orc_outflow_deficit.restype = d
        L.orc_deficit_near.argtypes = [vp, fp, i]
        L.orc_deficit_near.restype = d
        for name in ("orc_weno", "orc_upwind5"):
            getattr(L, name).restype = d
            getattr(L, name).argtypes = [d] * 5
        L.orc_fft.argtypes = [i, dp, dp, i]
        L.orc_hdft_forward.argtypes = [i, i, i, dp, dp, dp]
        L.orc_hdft_inverse.argtypes = [i, i, i, dp, dp, dp]
        for name, n in (("orc_qsat", 2), ("orc_qsat_ice", 2), ("orc_fall_speed", 3), ("orc_snow_fall_speed", 3),
                        ("orc_graupel_fall_speed", 2), ("orc_hail_fall_speed", 2), ("orc_takahashi", 2),
                        ("orc_takahashi_collision", 4)):
            getattr(L, name).restype = d
            getattr(L, name).argtypes = [d] * n
        L.orc_rain_processes.restype = d
        L.orc_rain_processes.argtypes = [fp, fp, fp, d, d, d, d]
        L.orc_saturation_adjust.restype = d
        L.orc_saturation_adjust.argtypes = [fp, fp, fp, d, d]
        L.orc_flash_limit.restype = i
        L.orc_flash_limit.argtypes = [d]
        L.orc_stretch_ratio.restype = d
        L.orc_stretch_ratio.argtypes = [d, d, i]
        L.orc_project.restype = d
        L.orc_project.argtypes = [ctypes.POINTER(Grid), dp, fp, fp, fp, d, dp]
        L.orc_advect.argtypes = [ctypes.POINTER(Grid), dp, fp, fp, fp, fp, d, i, i]
        L.orc_mulberry32.restype = ctypes.c_uint32
        L.orc_mulberry32.argtypes = [ctypes.POINTER(ctypes.c_uint32)]
        L.orc_electric_solve.argtypes = [ctypes.POINTER(Grid), dp, dp]
        _physics_signatures(L)
        _libs[path] = L
    return _libs[path]


def _physics_signatures(L):
    """The column physics, turbulence, sedimentation and lightning pieces of oracle.h."""
    vp, d, i = ctypes.c_void_p, ctypes.c_double, ctypes.c_int
    fp, dp, ip = ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int)
    G = ctypes.POINTER(Grid)
    sig = {
        "orc_sat_adjust_mixed": (d, [fp, fp, fp, fp, d, d, fp]),
        "orc_freeze_rain": (None, [fp, fp, fp, d]),
        "orc_ice_processes": (d, [fp, fp, fp, fp, fp, d, d, d, d]),
        "orc_settle_crystals": (None, [fp, fp]),
        "orc_snow_processes": (d, [fp, fp, fp, fp, fp, fp, d, d, d, d, d]),
        "orc_graupel_processes": (d, [fp, fp, fp, fp, fp, fp, fp, d, d, d, d]),
        "orc_hail_processes": (d, [fp, fp, fp, fp, fp, fp, fp, fp, d, d, d, d]),
        "orc_hail_max_diameter": (d, [d, d]),
        "orc_wet_growing": (i, [i, d, d, d, d, d, d, d, d, d]),
        "orc_cooper_crystals": (d, [d]), "orc_splintering": (d, [d]), "orc_fragments_per_kg": (d, [d]),
        "orc_koenig_growth": (d, [d, d]), "orc_crystal_size": (d, [d, d, d]), "orc_constant": (d, [i]),
        "orc_charging_rate": (d, [d, d, d, d, d, d, d, d, i, i, d, i]),
        "orc_inductive_rate": (d, [d, d, d, d, d, d, d, i, i]),
        "orc_transfer_charge": (None, [dp, dp, dp]),
        "orc_crystal_number": (d, [d, d, d]), "orc_crystal_diameter": (d, [d, d, d]),
        "orc_viscosity": (None, [G, fp, fp, fp, fp, dp, d, d, fp, fp, dp, dp, dp, i]),
        "orc_mix": (None, [G, dp, dp, dp, fp, dp, d, d]),
        "orc_fall_parts": (i, [vp, i, i, d]), "orc_sediment": (i, [vp, i, i, d, i]), "orc_charge_ground": (d, [vp]),
        "orc_lightning_create": (vp, [G, dp, i]), "orc_lightning_destroy": (None, [vp]),
        "orc_lightning_solve": (None, [vp, fp, fp, fp, fp]), "orc_lightning_strongest": (d, [vp, ip]),
        "orc_lightning_discharge": (i, [vp, fp, fp, fp, fp, d, i, dp, i]),
        "orc_lightning_channel": (i, [vp, i, ip, ip, ip, i]), "orc_lightning_conduct": (d, [vp, fp, d]),
        "orc_lightning_set_cloudy": (None, [vp, fp]), "orc_lightning_set_corona": (None, [vp, i]),
        "orc_lightning_phi": (None, [vp, dp]), "orc_lightning_totals": (None, [vp, dp, dp]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype, f.argtypes = res, args


def fptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def dptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def make_grid(g):
    return Grid(g["nx"], g["ny"], g["nz"], g["width"], g["depth"], g["height"], g.get("bottom", 0.0))


def make_profile(p):
    """The sounding dict as orc_profile; the wind errors are zero when absent."""
    scalars = [float(p[k]) for k, t in Profile._fields_ if t is ctypes.c_double]
    winds = [(ctypes.c_double * 4)(*(float(x) for x in p.get(k, (0.0,) * 4))) for k in ("wind_u", "wind_v")]
    return Profile(*scalars, *winds)


def make_run(r):
    return Run(*(r[k] for k, _ in Run._fields_))


class OracleRun:
    """One oracle model: AtmosphereModel restated (model.ts:183-1697) for a grid, sounding and run spec (dicts)."""

    def __init__(self, grid, profile, run):
        self.grid = grid
        self.n = grid["nx"] * grid["ny"] * grid["nz"]
        self.layer = grid["nx"] * grid["ny"]
        self._g, self._p, self._r = make_grid(grid), make_profile(profile), make_run(run)
        self.h = lib().orc_create(ctypes.byref(self._g), ctypes.byref(self._p), ctypes.byref(self._r))

    def close(self):
        if self.h:
            lib().orc_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def step(self, n=1, dt=3.0):
        lib().orc_step(self.h, dt, n)

    def step_part(self, part, dt=3.0):
        """One part of a step as the reference's GPU step splits it: 0 sedimentation + transport, 1 physics,
        2 mixing + projection, 3 lightning and time += dt."""
        assert lib().orc_step_part(self.h, int(part), dt) == 0

    def set_lightning(self, on):
        lib().orc_set_lightning(self.h, 1 if on else 0)

    def field(self, name):
        out = np.empty(self.n, np.float32)
        assert lib().orc_get_field(self.h, FIELD_ID[name], fptr(out)) == 0
        return out

    def set_field(self, name, values):
        a = np.ascontiguousarray(values, dtype=np.float32)
        assert a.size == self.n
        assert lib().orc_set_field(self.h, FIELD_ID[name], fptr(a)) == 0

    def column(self, name):
        out = np.empty(self.layer, np.float32)
        assert lib().orc_get_column(self.h, COLUMNS.index(name), fptr(out)) == 0
        return out

    def set_column(self, name, values):
        a = np.ascontiguousarray(values, dtype=np.float32)
        assert lib().orc_set_column(self.h, COLUMNS.index(name), fptr(a)) == 0

    def base_state(self):
        nz = self.grid["nz"]
        out = np.empty(7 * nz, np.float64)
        lib().orc_base_state(self.h, dptr(out))
        return dict(zip(["p", "exner", "theta", "q", "u", "v", "rho"], out.reshape(7, nz)))

    def levels(self):
        nz = self.grid["nz"]
        zs, hz = np.empty(nz), np.empty(nz)
        lib().orc_grid_levels(self.h, dptr(zs), dptr(hz))
        return zs, hz

    def diag(self):
        out = np.empty(len(DIAG), np.float64)
        lib().orc_diag(self.h, dptr(out))
        return dict(zip(DIAG, out))

    def pressure(self):
        out = np.empty(self.layer * (self.grid["nz"] - 1), np.float64)
        lib().orc_pressure(self.h, dptr(out))
        return out

    def set_time(self, t):
        lib().orc_set_time(self.h, t)

    def intervene(self, kind, x, y, strength=1.0, thermal_radius=0.0):
        """intervene (intervene.ts:38-68): returns {"heat": J, "vapour": kg} put in."""
        out = np.empty(2, np.float64)
        assert lib().orc_intervene(self.h, INTERVENTIONS.index(kind), float(x), float(y), float(strength), float(thermal_radius), dptr(out)) == 0
        return {"heat": float(out[0]), "vapour": float(out[1])}

    def added(self):
        out = np.empty(3, np.float64)
        lib().orc_added(self.h, dptr(out))
        return {"heat": float(out[0]), "vapour": float(out[1]), "count": int(out[2])}

    def rotation(self):
        """The rotation diagnostics of the last step (updateRotation), by the names of ROTATION."""
        out = np.empty(len(ROTATION), np.float64)
        lib().orc_rotation(self.h, dptr(out))
        return dict(zip(ROTATION, out))

    def model_diagnostics(self):
        """AtmosphereModel.diagnostics() of the fields as they are, by the names of MODEL_DIAGNOSTICS."""
        out = np.empty(len(MODEL_DIAGNOSTICS), np.float64)
        lib().orc_model_diagnostics(self.h, dptr(out))
        return dict(zip(MODEL_DIAGNOSTICS, out))

    def outflow_deficit(self):
        """outflowDeficit (rkw.ts:72-80): (deficit per ground column as float32, warm, max)."""
        deficit, warm = np.empty(self.layer, np.float32), ctypes.c_double()
        mx = lib().orc_outflow_deficit(self.h, fptr(deficit), ctypes.byref(warm))
        return deficit, warm.value, mx

    def deficit_near(self, deficit, column):
        """deficitNear (rkw.ts:82-90) within 2 km of a ground column."""
        a = np.ascontiguousarray(deficit, dtype=np.float32)
        return lib().orc_deficit_near(self.h, fptr(a), int(column))

    def set_swaths(self, on):
        """runMember's addSwaths after every step (ensemble/member.ts:85-114)."""
        lib().orc_set_swaths(self.h, 1 if on else 0)

    def add_swaths(self):
        """addSwaths once at the model's time (ensemble/member.ts:85-100)."""
        lib().orc_add_swaths(self.h)

    def uh_column(self):
        """The last updateRotation's 2-5 km helicity of each column (model.ts uhColumn)."""
        out = np.empty(self.layer, np.float32)
        assert lib().orc_uh_column(self.h, fptr(out)) == 0
        return out

    def update_rotation(self, dt=0.0, keep=True):
        """updateRotation(dt) on the fields as they are; keep: refreshRotation (how long it held stays)."""
        lib().orc_update_rotation(self.h, dt, 1 if keep else 0)

    def state(self):
        return {name: self.field(name) for name in CARRIED}

    def load_state(self, state, columns=None, time=None):
        for name, v in state.items():
            self.set_field(name, v)
        for name, v in (columns or {}).items():
            self.set_column(name, v)
        if time is not None:
            self.set_time(time)
