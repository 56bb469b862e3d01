"""The .meteolab state file of the reference (src/core/state.ts:1-110): a run's prognostic fields at one moment, so a
run can be continued later or passed on, in the reference's layout:

    'MLAB', the header's length (uint32, little-endian), the JSON header (format, version, config, model time, grid
    size, lightning totals), padding to 4 bytes, then float32 fields in STATE_FIELDS order and columns in
    STATE_COLUMNS order, node i = x + nx (y + ny z).

save_state writes version 5 (state.ts:18): the two-moment numbers and volumes, which this step does not carry, are
zeros, and the water and charge budgets, which it does not keep, are zeros; the heat and vapour put in by
interventions (header.added) are the run's. read_state takes versions 1-5 and refuses
what the reference refuses, with its messages (state.ts:53-68). apply_state puts a state into a run of the same grid
(state.ts:74-110): the fields the file holds, zeros for those it does not, the columns, the model time, the
lightning totals, how long the rotation has held and the interventions' counts; the flash log starts empty. Everything else a step needs it recomputes from these, except the
vertical field of the last discharge, which a step computes at its end: a restored run's first step charges without
it, as the reference's applyState leaves it.
"""
import json
import math
import struct

import numpy as np

MAGIC = b"MLAB"
VERSION = 5
STATE_FIELDS = ["u", "v", "w", "theta", "q", "cloud", "rain", "cold", "ice", "snow", "graupel", "hail", "chargeIce",
                "chargeGraupel", "chargeRain", "chargeAir", "iceNumber", "rainNumber", "graupelNumber", "graupelVolume",
                "hailNumber", "hailVolume"]
STATE_COLUMNS = ["precipitation", "hailGround", "hailSize"]
# Fields a file of each version holds (state.ts:19, fieldsOf :32).
FIELD_COUNT = {1: 12, 2: 16, 3: 17, 4: 18, 5: 22}
# The two-moment numbers and volumes: fields of the file that this step does not carry.
UNCARRIED = {"rainNumber", "graupelNumber", "graupelVolume", "hailNumber", "hailVolume"}


def _number(x):
    """A float as JSON.stringify writes it: integral values without a fraction."""
    x = float(x)
    return int(x) if x.is_integer() and abs(x) < 2 ** 53 else x


class StateError(ValueError):
    """A state file the reference's readState or applyState would refuse, with its message."""


def save_state(run, config=None):
    """saveState (state.ts:34-45): the run's state as the bytes of a .meteolab file. `config` is the header's config
    (the reference writes its SimConfig; by default the run's grid, sounding and run settings)."""
    g = run.grid
    n, layer = g["nx"] * g["ny"] * g["nz"], g["nx"] * g["ny"]
    d = run.diagnostics()
    header = {"format": "meteolab-state", "version": VERSION, "config": run.spec if config is None else config,
              "time": _number(d["time"]), "persisted": _number(run.rotation()["persisted"]), "nx": g["nx"], "ny": g["ny"], "nz": g["nz"],
              "chargeGround": _number(d["chargeGround"]),
              "flashCount": {"ic": int(d["flashesIC"]), "cg": int(d["flashesCG"])},
              "water": {"condensed": 0, "evaporated": 0, "rainEvaporated": 0, "surface": 0, "start": 0},
              "added": {k: _number(v) if k != "count" else int(v) for k, v in run.added().items()},
              "charge": {"separated": 0, "ic": 0, "cgNegative": 0, "cgPositive": 0, "precipitation": 0, "conduction": 0}}
    text = json.dumps(header, separators=(",", ":")).encode()
    start = math.ceil((8 + len(text)) / 4) * 4
    data = np.zeros(len(STATE_FIELDS) * n + len(STATE_COLUMNS) * layer, np.float32)
    for k, name in enumerate(STATE_FIELDS):
        if name not in UNCARRIED:
            data[k * n:(k + 1) * n] = run.field(name)
    o = len(STATE_FIELDS) * n
    for k, name in enumerate(STATE_COLUMNS):
        data[o + k * layer:o + (k + 1) * layer] = run.column(name)
    return MAGIC + struct.pack("<I", len(text)) + text + b"\0" * (start - 8 - len(text)) + data.tobytes()


def read_state(buffer):
    """readState (state.ts:48-68): (header, data) of a .meteolab file, or StateError with the reference's message."""
    buffer = bytes(buffer)
    if len(buffer) < 8 or buffer[:4] != MAGIC:
        raise StateError("this is not a MeteoLab state file")
    length = struct.unpack("<I", buffer[4:8])[0]
    if 8 + length > len(buffer):
        raise StateError("the file is truncated")
    try:
        header = json.loads(buffer[8:8 + length].decode())
    except (UnicodeDecodeError, ValueError):
        raise StateError("the file’s header is damaged") from None
    if not isinstance(header, dict) or header.get("format") != "meteolab-state" or header.get("version") not in FIELD_COUNT:
        raise StateError("unknown file format or version")
    for k in ("time", "persisted", "nx", "ny", "nz"):
        v = header.get(k)
        if not isinstance(v, (int, float)) or isinstance(v, bool) or not math.isfinite(v) or v < 0:
            raise StateError("the file’s header is damaged")
    start = math.ceil((8 + length) / 4) * 4
    cells = int(header["nx"]) * int(header["ny"])
    size = FIELD_COUNT[header["version"]] * cells * int(header["nz"]) + len(STATE_COLUMNS) * cells
    if len(buffer) != start + size * 4:
        raise StateError("the file’s size does not match the grid")
    return header, np.frombuffer(buffer, np.float32, size, start)


def apply_state(run, state):
    """applyState (state.ts:74-110) into a run of the same grid: the file's fields (zeros for those it does not hold;
    the two-moment ones are not carried here), the columns, the model time and the lightning totals."""
    header, data = state
    g = run.grid
    if (header["nx"], header["ny"], header["nz"]) != (g["nx"], g["ny"], g["nz"]):
        raise StateError(f"the file’s grid is {header['nx']}×{header['ny']}×{header['nz']}, "
                         f"the model’s {g['nx']}×{g['ny']}×{g['nz']}")
    n, layer = g["nx"] * g["ny"] * g["nz"], g["nx"] * g["ny"]
    held = STATE_FIELDS[:FIELD_COUNT[header["version"]]]
    zero = np.zeros(n, np.float32)
    for k, name in enumerate(STATE_FIELDS):
        if name in UNCARRIED:
            continue
        run.set_field(name, data[k * n:(k + 1) * n] if name in held else zero)
    o = len(held) * n
    for k, name in enumerate(STATE_COLUMNS):
        run.set_column(name, data[o + k * layer:o + (k + 1) * layer])
    run.set_time(header["time"])
    count = header.get("flashCount") or {}
    run.set_totals(header.get("chargeGround") or 0.0, count.get("ic") or 0, count.get("cg") or 0, header["persisted"])
    added = header.get("added") or {}
    run.set_added(added.get("heat") or 0.0, added.get("vapour") or 0.0, added.get("count") or 0)
    # The rotation diagnostics are not in the file: recomputed from its fields (state.ts:73-74).
    run.refresh_rotation()
