"""StormLab model step on MI355X: the host side of the C ABI in include/meteolab.h (libmeteolab.so).

The reference's step is ``AtmosphereModel.step(dt)`` (src/core/model.ts:873-877) and its GPU plugin ``GpuDynamics``
(src/gpu/dynamics.ts:250-504). Here :class:`Context` owns a device, :class:`Pool` the resident runs of one grid size,
and :class:`Run` one model run with the reference's field names; :func:`step` advances a batch of runs of one pool.
The library must be built (``python -m meteolab_amd.build``); nothing here has a CPU path.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmeteolab.so")

# Field ids of include/meteolab.h: the carried order of AtmosphereModel.carried() (model.ts:1319-1335), then fieldZ,
# the fall speeds and the condensate.
CARRIED = ["rain", "snow", "graupel", "hail", "u", "v", "w", "theta", "q", "cloud", "ice", "iceNumber", "cold",
           "chargeIce", "chargeGraupel", "chargeRain", "chargeAir"]
FIELDS = CARRIED + ["fieldZ", "fallRain", "fallSnow", "fallGraupel", "fallHail", "fallChargeIce",
                    "fallChargeGraupel", "condensate"]
FIELD_ID = {name: k for k, name in enumerate(FIELDS)}
COLUMNS = ["precipitation", "hailGround", "hailSize", "gustMax", "uhMax"]
# InterventionKind (src/core/intervene.ts:13) in ml_run_intervene's order.
INTERVENTIONS = ["thermal", "warm", "cool", "moisten", "dry", "coldPool", "vortex"]
LEVELS = ["p", "exner", "theta", "q", "u", "v", "rho"]
DIAG = ["time", "maxW", "minW", "maxThetaP", "rainTotal", "water", "charge", "clipped", "negativeFilled",
        "chargeGround", "flashesIC", "flashesCG", "firstFlash", "error"]
# ModelDiagnostics (model.ts:161-180) in ml_model_diagnostics' order.
MODEL_DIAGNOSTICS = ["updraft", "downdraft", "rainRate", "rainTotal", "cloudTop", "thermalTop", "maxCloud", "coldMax",
                     "cores", "shear06", "outflow", "outflowLow", "gust", "hail", "clipped"]
# RotationState (model.ts:142-158) in ml_run_rotation's order.
ROTATION = ["uh", "x", "y", "anticyclonic", "persisted", "uh01", "uh03", "low", "coreZeta", "coreUH", "coreAnti"]
# STATE_FIELDS of src/core/state.ts:9 that this step carries (two-moment numbers are not part of the baseline step).
STATE_FIELDS = ["u", "v", "w", "theta", "q", "cloud", "rain", "cold", "ice", "snow", "graupel", "hail", "chargeIce",
                "chargeGraupel", "chargeRain", "chargeAir", "iceNumber"]


class MeteolabError(RuntimeError):
    """An error code of the C ABI, with ml_last_error()'s message."""


class Grid(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int), ("width", ctypes.c_double),
                ("depth", ctypes.c_double), ("height", ctypes.c_double), ("bottom", ctypes.c_double)]


class Profile(ctypes.Structure):
    _fields_ = [(k, ctypes.c_double) for k in
                ("theta0", "dtheta", "ztr", "ttr", "rh_drop", "rh_top", "qv_max", "shear", "shear_depth")] + \
        [("wind_u", ctypes.c_double * 4), ("wind_v", ctypes.c_double * 4)]


def profile_struct(profile):
    """The sounding dict as ml_profile; the wind errors (wind_u, wind_v at 0, 3, 6 and 10 km) are zero when absent."""
    scalars = [float(profile[k]) for k, t in Profile._fields_ if t is ctypes.c_double]
    winds = [(ctypes.c_double * 4)(*(float(x) for x in profile.get(k, (0.0,) * 4))) for k in ("wind_u", "wind_v")]
    return Profile(*scalars, *winds)


class RunConfig(ctypes.Structure):
    _fields_ = [("seed", ctypes.c_int), ("latitude", ctypes.c_double), ("turbulence", ctypes.c_double),
                ("tropopause_km", ctypes.c_double), ("roughness", ctypes.c_double), ("frame_u", ctypes.c_double),
                ("frame_v", ctypes.c_double), ("bubble_amp", ctypes.c_double), ("bubble_x", ctypes.c_double),
                ("bubble_y", ctypes.c_double), ("bubble_z", ctypes.c_double), ("bubble_rh", ctypes.c_double),
                ("bubble_rz", ctypes.c_double), ("noise_seed", ctypes.c_int), ("noise_amp", ctypes.c_double),
                ("noise_top", ctypes.c_double)]


# Every symbol include/meteolab.h declares, with (restype, argtypes).
_vp, _i, _d, _i64 = ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_int64
_ip, _fp, _dp, _cp = ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_double), ctypes.c_char_p
SIGNATURES = {
    "ml_ctx_create": (_i, [_i, _i, ctypes.POINTER(_vp)]),
    "ml_ctx_destroy": (None, [_vp]),
    "ml_last_error": (_cp, [_vp]),
    "ml_device_name": (_i, [_vp, _cp, _i]),
    "ml_pool_create": (_i, [_vp, ctypes.POINTER(Grid), _i, _i, _ip]),
    "ml_slot_bytes": (_i64, [_vp, ctypes.POINTER(Grid)]),
    "ml_pool_free": (_i, [_vp, _i]),
    "ml_pool_destroy": (_i, [_vp, _i]),
    "ml_run_create": (_i, [_vp, _i, ctypes.POINTER(Profile), ctypes.POINTER(RunConfig), _ip]),
    "ml_runs_create": (_i, [_vp, _i, _i, ctypes.POINTER(Profile), ctypes.POINTER(RunConfig), _ip]),
    "ml_pool_grow": (_i, [_vp, _i, _i]),
    "ml_mem_info": (_i, [_vp, ctypes.POINTER(_i64), ctypes.POINTER(_i64)]),
    "ml_ctx_query": (_i, [_vp]),
    "ml_run_release": (_i, [_vp, _i]),
    "ml_run_upload": (_i, [_vp, _i, _i, _fp]),
    "ml_run_download": (_i, [_vp, _i, _i, _fp]),
    "ml_run_upload_column": (_i, [_vp, _i, _i, _fp]),
    "ml_run_download_column": (_i, [_vp, _i, _i, _fp]),
    "ml_run_levels": (_i, [_vp, _i, _dp]),
    "ml_run_pressure": (_i, [_vp, _i, _dp]),
    "ml_run_diag": (_i, [_vp, _i, _dp]),
    "ml_run_set_time": (_i, [_vp, _i, _d]),
    "ml_run_set_totals": (_i, [_vp, _i, _d, _i64, _i64, _d]),
    "ml_run_rotation": (_i, [_vp, _i, _dp]),
    "ml_run_intervene": (_i, [_vp, _i, _i, _d, _d, _d, _d, _dp]),
    "ml_run_added": (_i, [_vp, _i, _dp]),
    "ml_run_set_added": (_i, [_vp, _i, _d, _d, _d]),
    "ml_run_refresh_rotation": (_i, [_vp, _i]),
    "ml_run_set_swaths": (_i, [_vp, _i, _i]),
    "ml_probability_map": (_i, [_vp, _ip, _i, _i, _d, _fp]),
    "ml_check_weno": (_i, [_vp, _fp, _i, _dp]),
    "ml_model_diagnostics": (_i, [_vp, _ip, _i, _dp]),
    "ml_model_diagnostics_noalloc": (_i, [_vp, _ip, _i, _dp]),
    "ml_run_flashes": (_i, [_vp, _i, _dp, _i]),
    "ml_step": (_i, [_vp, _ip, _i, _i, _d]),
    "ml_ctx_set_overlap": (_i, [_vp, _i]),
    "ml_dynamics": (_i, [_vp, _ip, _i, _d]),
    "ml_physics": (_i, [_vp, _ip, _i, _d]),
    "ml_mix": (_i, [_vp, _ip, _i, _d]),
    "ml_after_step": (_i, [_vp, _ip, _i, _d]),
    "ml_fall_inputs": (_i, [_vp, _i, _ip, _i]),
    "ml_sync": (_i, [_vp]),
    "ml_step_timed": (_i, [_vp, _ip, _i, _i, _d, _dp]),
    "ml_copy_probe": (_i, [_vp, _i64, _i, _dp]),
    "ml_batch_tuning": (_i, [_vp, _ip, _i, _ip, _fp]),
    "ml_batch_set_tuning": (_i, [_vp, _ip, _i, _ip]),
    "ml_profile_step": (_i, [_vp, _ip, _i, _i, _d, _dp, _i, _cp, _i]),
}

_lib = None


def lib():
    """Loads libmeteolab.so. There is no fallback: a missing library is an error."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"libmeteolab.so is not built ({LIB_PATH}); run python -m meteolab_amd.build")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _fptr(a):
    return a.ctypes.data_as(_fp)


def _dptr(a):
    return a.ctypes.data_as(_dp)


class Context:
    """A device and its stream (ml_ctx_create). precision 32: the production float32 kernels; 64: the same in float64."""

    def __init__(self, device=0, precision=32):
        self.precision = precision
        h = ctypes.c_void_p()
        rc = lib().ml_ctx_create(device, precision, ctypes.byref(h))
        if rc != 0:
            raise MeteolabError(f"ml_ctx_create failed ({rc}): no usable HIP device {device}")
        self.h = h

    def check(self, rc):
        if rc < 0:
            raise MeteolabError(f"meteolab error {rc}: {lib().ml_last_error(self.h).decode()}")
        return rc

    @property
    def adapter(self):
        buf = ctypes.create_string_buffer(256)
        self.check(lib().ml_device_name(self.h, buf, 256))
        return buf.value.decode()

    def pool(self, grid, capacity, electrified=True):
        return Pool(self, grid, capacity, electrified)

    def slot_bytes(self, grid):
        return self.check(lib().ml_slot_bytes(self.h, ctypes.byref(make_grid(grid))))

    def check_weno(self, stencils):
        """The transport kernels' WENO5 face values minus c of stencils [n, 5] (a, b, c, d, e), on the device
        (ml_check_weno)."""
        st = np.ascontiguousarray(stencils, np.float32)
        assert st.ndim == 2 and st.shape[1] == 5
        out = np.empty(len(st), np.float64)
        self.check(lib().ml_check_weno(self.h, _fptr(st), len(st), _dptr(out)))
        return out

    def sync(self):
        self.check(lib().ml_sync(self.h))

    def idle(self):
        """True when all the work given to the context has finished, without waiting (ml_ctx_query)."""
        return self.check(lib().ml_ctx_query(self.h)) == 1

    def mem_info(self):
        """(free, total) device memory in bytes (ml_mem_info)."""
        free, total = _i64(), _i64()
        self.check(lib().ml_mem_info(self.h, ctypes.byref(free), ctypes.byref(total)))
        return free.value, total.value

    def set_overlap(self, on):
        """A step's lightning beside the next step's dynamics (True, the default) or in line (ml_ctx_set_overlap)."""
        self.check(lib().ml_ctx_set_overlap(self.h, 1 if on else 0))

    def copy_bandwidth(self, nbytes=1 << 30, iters=20):
        """Achievable HBM bandwidth of a plain copy (read + write), GB/s (ml_copy_probe)."""
        out = ctypes.c_double()
        self.check(lib().ml_copy_probe(self.h, int(nbytes), int(iters), ctypes.byref(out)))
        return out.value

    def close(self):
        if self.h:
            lib().ml_ctx_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def make_grid(g):
    return Grid(int(g["nx"]), int(g["ny"]), int(g["nz"]), float(g["width"]), float(g["depth"]), float(g["height"]),
                float(g.get("bottom", 0.0)))


class Pool:
    """Slots for `capacity` runs of one grid (ml_pool_create); their state stays resident on the device."""

    def __init__(self, ctx, grid, capacity, electrified=True):
        self.ctx, self.grid, self.capacity = ctx, dict(grid), capacity
        self.n = grid["nx"] * grid["ny"] * grid["nz"]
        self.layer = grid["nx"] * grid["ny"]
        pid = ctypes.c_int()
        self._g = make_grid(grid)
        ctx.check(lib().ml_pool_create(ctx.h, ctypes.byref(self._g), capacity, 1 if electrified else 0, ctypes.byref(pid)))
        self.id = pid.value

    def free(self):
        return self.ctx.check(lib().ml_pool_free(self.ctx.h, self.id))

    def destroy(self):
        """Frees the pool's device memory (ml_pool_destroy); its runs become invalid. Also done when the pool (and
        every run holding it) is garbage-collected while its context is open."""
        if self.id is not None and self.ctx.h:
            self.ctx.check(lib().ml_pool_destroy(self.ctx.h, self.id))
        self.id = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass

    def run(self, profile, run):
        return Run(self, profile, run)

    def runs(self, starts):
        """Runs of (profile, run) pairs made at once (ml_runs_create: one launch of each start-up kernel, one wait)."""
        starts = list(starts)
        if not starts:
            return []
        profs = (Profile * len(starts))(*(profile_struct(p) for p, _ in starts))
        confs = (RunConfig * len(starts))(*(RunConfig(*(r[k] for k, _ in RunConfig._fields_)) for _, r in starts))
        ids = (ctypes.c_int * len(starts))()
        self.ctx.check(lib().ml_runs_create(self.ctx.h, self.id, len(starts), profs, confs, ids))
        return [Run(self, p, r, handle=h) for (p, r), h in zip(starts, ids)]

    def grow(self, capacity):
        """Grows the pool to `capacity` slots, its runs kept (ml_pool_grow)."""
        self.ctx.check(lib().ml_pool_grow(self.ctx.h, self.id, int(capacity)))
        self.capacity = int(capacity)


class Run:
    """One model run in a pool slot: its base state and initial state are built on the device (ml_run_create)."""

    def __init__(self, pool, profile, run, handle=None):
        self.pool, self.ctx = pool, pool.ctx
        self.n, self.layer, self.grid = pool.n, pool.layer, pool.grid
        if handle is None:
            p = profile_struct(profile)
            r = RunConfig(*(run[k] for k, _ in RunConfig._fields_))
            h = ctypes.c_int()
            self.ctx.check(lib().ml_run_create(self.ctx.h, pool.id, ctypes.byref(p), ctypes.byref(r), ctypes.byref(h)))
            handle = h.value
        self.h = handle
        self.spec = {"grid": dict(pool.grid), "profile": dict(profile), "run": dict(run)}

    def release(self):
        if self.h is not None:
            self.ctx.check(lib().ml_run_release(self.ctx.h, self.h))
            self.h = None

    def field(self, name):
        out = np.empty(self.n, np.float32)
        self.ctx.check(lib().ml_run_download(self.ctx.h, self.h, FIELD_ID[name], _fptr(out)))
        return out

    def set_field(self, name, values):
        a = np.ascontiguousarray(values, dtype=np.float32)
        if a.size != self.n:
            raise ValueError(f"{name}: {a.size} values for {self.n} nodes")
        self.ctx.check(lib().ml_run_upload(self.ctx.h, self.h, FIELD_ID[name], _fptr(a)))

    def column(self, name):
        out = np.empty(self.layer, np.float32)
        self.ctx.check(lib().ml_run_download_column(self.ctx.h, self.h, COLUMNS.index(name), _fptr(out)))
        return out

    def set_column(self, name, values):
        a = np.ascontiguousarray(values, dtype=np.float32)
        if a.size != self.layer:
            raise ValueError(f"{name}: {a.size} values for {self.layer} columns")
        self.ctx.check(lib().ml_run_upload_column(self.ctx.h, self.h, COLUMNS.index(name), _fptr(a)))

    def levels(self):
        nz = self.grid["nz"]
        out = np.empty(len(LEVELS) * nz, np.float64)
        self.ctx.check(lib().ml_run_levels(self.ctx.h, self.h, _dptr(out)))
        return dict(zip(LEVELS, out.reshape(len(LEVELS), nz)))

    def pressure(self):
        """The last projection's pressure cells, (nz - 1, ny, nx) float64 (model.ts:358)."""
        g = self.grid
        out = np.empty((g["nz"] - 1) * g["ny"] * g["nx"], np.float64)
        self.ctx.check(lib().ml_run_pressure(self.ctx.h, self.h, _dptr(out)))
        return out.reshape(g["nz"] - 1, g["ny"], g["nx"])

    def diagnostics(self):
        out = np.empty(len(DIAG), np.float64)
        self.ctx.check(lib().ml_run_diag(self.ctx.h, self.h, _dptr(out)))
        return dict(zip(DIAG, out))

    def flashes(self, max_count=64):
        out = np.zeros(max_count * 8, np.float64)
        k = self.ctx.check(lib().ml_run_flashes(self.ctx.h, self.h, _dptr(out), max_count))
        keys = ["time", "cg", "start", "nodes", "charge", "potential", "moment"]
        return [dict(zip(keys, out[8 * j:8 * j + 7])) for j in range(k)]

    def fall_parts(self):
        """The sub-steps of each fall of the last step or ml_dynamics (ml_fall_inputs): rain, snow, graupel, hail,
        then the charges of ice, graupel and rain."""
        out = (ctypes.c_int * 8)()
        k = self.ctx.check(lib().ml_fall_inputs(self.ctx.h, self.h, out, 8))
        return list(out[:k])

    def set_time(self, t):
        self.ctx.check(lib().ml_run_set_time(self.ctx.h, self.h, float(t)))

    def set_totals(self, charge_ground, flashes_ic, flashes_cg, persisted=0.0):
        """The totals of a saved state's header (ml_run_set_totals)."""
        self.ctx.check(lib().ml_run_set_totals(self.ctx.h, self.h, float(charge_ground), int(flashes_ic), int(flashes_cg), float(persisted)))

    def intervene(self, kind, x, y, strength=1.0, thermal_radius=0.0):
        """intervene (src/core/intervene.ts:38-68) on the run's fields (ml_run_intervene): `kind` one of INTERVENTIONS at
        (x, y) m from the domain's corner; returns {"heat": J, "vapour": kg} put in."""
        if kind not in INTERVENTIONS:
            raise ValueError(f"intervention kind {kind!r}: one of {INTERVENTIONS}")
        out = np.empty(2, np.float64)
        self.ctx.check(lib().ml_run_intervene(self.ctx.h, self.h, INTERVENTIONS.index(kind), float(x), float(y), float(strength),
                                              float(thermal_radius), _dptr(out)))
        return {"heat": float(out[0]), "vapour": float(out[1])}

    def added(self):
        """AtmosphereModel.added (ml_run_added): the heat (J) and vapour (kg) put in by interventions, and their count."""
        out = np.empty(3, np.float64)
        self.ctx.check(lib().ml_run_added(self.ctx.h, self.h, _dptr(out)))
        return {"heat": float(out[0]), "vapour": float(out[1]), "count": int(out[2])}

    def set_added(self, heat, vapour, count):
        """Sets those counts (a saved state's header.added; ml_run_set_added)."""
        self.ctx.check(lib().ml_run_set_added(self.ctx.h, self.h, float(heat), float(vapour), float(count)))

    def rotation(self):
        """The rotation diagnostics of the last step (ml_run_rotation), by the names of ROTATION."""
        out = np.empty(len(ROTATION), np.float64)
        self.ctx.check(lib().ml_run_rotation(self.ctx.h, self.h, _dptr(out)))
        return dict(zip(ROTATION, out))

    def set_swaths(self, on):
        """Fold each step's surface wind and helicity into the gustMax and uhMax columns (ml_run_set_swaths)."""
        self.ctx.check(lib().ml_run_set_swaths(self.ctx.h, self.h, 1 if on else 0))

    def refresh_rotation(self):
        """Recomputes the rotation diagnostics from the fields as they are, keeping how long the rotation has held
        (ml_run_refresh_rotation, refreshRotation of model.ts:1607)."""
        self.ctx.check(lib().ml_run_refresh_rotation(self.ctx.h, self.h))

    def state(self):
        return {name: self.field(name) for name in CARRIED}

    def load_state(self, state, columns=None, time=None):
        for name, v in state.items():
            self.set_field(name, v)
        for name, v in (columns or {}).items():
            self.set_column(name, v)
        if time is not None:
            self.set_time(time)

    def step(self, n=1, dt=3.0):
        step([self], n, dt)
        self.ctx.sync()

    def model_diagnostics(self):
        """AtmosphereModel.diagnostics() of the run's fields as they are (ml_model_diagnostics)."""
        return model_diagnostics([self])[0]


def model_diagnostics(runs):
    """AtmosphereModel.diagnostics() of runs of one pool in one launch (ml_model_diagnostics): a dict by the names of
    MODEL_DIAGNOSTICS per run."""
    ctx = runs[0].ctx
    ids = (ctypes.c_int * len(runs))(*[r.h for r in runs])
    out = np.empty((len(runs), len(MODEL_DIAGNOSTICS)), np.float64)
    ctx.check(lib().ml_model_diagnostics(ctx.h, ids, len(runs), _dptr(out)))
    return [dict(zip(MODEL_DIAGNOSTICS, row)) for row in out]


def model_diagnostics_noalloc(runs):
    """model_diagnostics through ml_model_diagnostics_noalloc: the same results, the scratch in the library's static
    device array (at most ML_NOALLOC_RUNS runs a call)."""
    ctx = runs[0].ctx
    ids = (ctypes.c_int * len(runs))(*[r.h for r in runs])
    out = np.empty((len(runs), len(MODEL_DIAGNOSTICS)), np.float64)
    ctx.check(lib().ml_model_diagnostics_noalloc(ctx.h, ids, len(runs), _dptr(out)))
    return [dict(zip(MODEL_DIAGNOSTICS, row)) for row in out]


NOALLOC_RUNS = 1024  # ML_NOALLOC_RUNS


def probability_map(runs, column, threshold):
    """probabilityMap over runs of one pool (ml_probability_map): per ground column the share of the runs whose column
    reached the threshold there (float32, layer)."""
    ctx = runs[0].ctx
    ids = (ctypes.c_int * len(runs))(*[r.h for r in runs])
    out = np.empty(runs[0].layer, np.float32)
    ctx.check(lib().ml_probability_map(ctx.h, ids, len(runs), COLUMNS.index(column), float(threshold), _fptr(out)))
    return out


def step(runs, n=1, dt=3.0):
    """Advances the runs (one pool) by n steps of dt (ml_step); asynchronous, Context.sync() waits."""
    if not runs:
        return
    ctx = runs[0].ctx
    ids = (ctypes.c_int * len(runs))(*[r.h for r in runs])
    ctx.check(lib().ml_step(ctx.h, ids, len(runs), int(n), float(dt)))


# The parts of a step (ml_dynamics, ml_physics, ml_mix, ml_after_step), in the order a step runs them.
STEP_PARTS = ["dynamics", "physics", "mix", "after_step"]


def step_part(runs, part, dt=3.0):
    """One part of a step of the runs (one pool), named as in STEP_PARTS; asynchronous. The four parts in order are
    one step: the split of the reference's GPU step (src/gpu/step.ts:12-31), for a host that runs some parts itself."""
    if part not in STEP_PARTS:
        raise ValueError(f"step part {part!r}: one of {STEP_PARTS}")
    ctx = runs[0].ctx
    ids = (ctypes.c_int * len(runs))(*[r.h for r in runs])
    ctx.check(getattr(lib(), "ml_" + part)(ctx.h, ids, len(runs), float(dt)))


def step_timed(runs, n, dt=3.0):
    """ml_step_timed: n graph-replayed steps with a HIP event after each; the device time of each step, ms."""
    ctx = runs[0].ctx
    ids = (ctypes.c_int * len(runs))(*[r.h for r in runs])
    ms = np.zeros(int(n))
    ctx.check(lib().ml_step_timed(ctx.h, ids, len(runs), int(n), float(dt), _dptr(ms)))
    return ms


def profile_step(runs, n=3, dt=3.0, max_launches=128):
    """Device time of each kernel launch of a step (HIP events between eager launches), ms averaged over n steps:
    a list of (kernel name, ms) in launch order."""
    ctx = runs[0].ctx
    ids = (ctypes.c_int * len(runs))(*[r.h for r in runs])
    ms = np.zeros(max_launches)
    names = ctypes.create_string_buffer(8192)
    k = ctx.check(lib().ml_profile_step(ctx.h, ids, len(runs), int(n), float(dt), _dptr(ms), max_launches, names, 8192))
    return list(zip(names.value.decode().split("\n")[:k], ms[:k]))


# The transport kernel's layouts (ml_batch_tuning's shape[0]).
LAYOUTS = ["strip", "level", "plane"]


def batch_tuning(runs):
    """The autotuned launch shape of this batch's transport kernel (ml_batch_tuning)."""
    ctx = runs[0].ctx
    ids = (ctypes.c_int * len(runs))(*[r.h for r in runs])
    shape = (ctypes.c_int * 4)()
    ms = ctypes.c_float()
    ctx.check(lib().ml_batch_tuning(ctx.h, ids, len(runs), shape, ctypes.byref(ms)))
    layout = LAYOUTS[shape[0]]
    return {"transport_layout": layout, "field_groups": shape[1] if shape[0] == 0 else 1,
            "levels_per_block": shape[3] if shape[0] else 1, "xcd_per_run": bool(shape[2]),
            "stage0_ms": round(ms.value, 4)}


def set_batch_tuning(runs, layout, groups=1, xcd=False, levels=1):
    """Forces this batch's transport launch shape (ml_batch_set_tuning): layout "level" (k_stage_level, `levels`
    levels per block) or "plane" (k_stage_plane, the same with each level in a padded LDS plane); the "strip" layout
    is not built and is refused."""
    ctx = runs[0].ctx
    ids = (ctypes.c_int * len(runs))(*[r.h for r in runs])
    shape = (ctypes.c_int * 4)(LAYOUTS.index(layout), int(groups), 1 if xcd else 0, int(levels))
    ctx.check(lib().ml_batch_set_tuning(ctx.h, ids, len(runs), shape))


def kernel_times(launches):
    """Per kernel name: (launches per step, total ms per step) from profile_step's list."""
    out = {}
    for name, ms in launches:
        c, t = out.get(name, (0, 0.0))
        out[name] = (c + 1, t + float(ms))
    return out
