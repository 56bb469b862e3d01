"""The .meteolab state file (src/core/state.ts): the layout, version handling and refusals of readState / applyState
(state.ts:48-110, tests as src/core/state.test.ts:8-85), the Python and Node hosts reading each other's files, and on
the GPU: a run continued from a saved state steps exactly as the original, lightning included."""
import json
import math
import os
import shutil
import struct
import subprocess

import numpy as np
import pytest

from meteolab_amd import config as C
from meteolab_amd.state import (FIELD_COUNT, STATE_COLUMNS, STATE_FIELDS, StateError, apply_state, read_state,
                                save_state)

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
NODE = shutil.which("node")


def file_of(header, data):
    """A state file laid out as saveState writes it (state.ts:37-44)."""
    text = json.dumps(header, separators=(",", ":")).encode()
    start = math.ceil((8 + len(text)) / 4) * 4
    return b"MLAB" + struct.pack("<I", len(text)) + text + b"\0" * (start - 8 - len(text)) + np.asarray(data, np.float32).tobytes()


class FakeRun:
    """What save_state and apply_state use of a run, on the host: a grid, fields, columns, diagnostics."""

    def __init__(self, grid, seed=0):
        self.grid, self.spec = dict(grid), {"grid": dict(grid)}
        n, layer = grid["nx"] * grid["ny"] * grid["nz"], grid["nx"] * grid["ny"]
        rng = np.random.default_rng(seed)
        self.fields = {k: rng.standard_normal(n).astype(np.float32) for k in STATE_FIELDS}
        self.columns = {k: rng.standard_normal(layer).astype(np.float32) for k in STATE_COLUMNS}
        self.time, self.counts = 612.0, (1.5, 3, 1, 90.0)
        self.added_counts = {"heat": 2.5e13, "vapour": -1.25e6, "count": 3}

    def diagnostics(self):
        return {"time": self.time, "chargeGround": self.counts[0], "flashesIC": self.counts[1], "flashesCG": self.counts[2]}

    def field(self, k):
        return self.fields[k]

    def column(self, k):
        return self.columns[k]

    def set_field(self, k, v):
        self.fields[k] = np.array(v, np.float32)

    def set_column(self, k, v):
        self.columns[k] = np.array(v, np.float32)

    def set_time(self, t):
        self.time = t

    def rotation(self):
        return {"persisted": self.counts[3]}

    def set_totals(self, q, ic, cg, persisted):
        self.counts = (q, ic, cg, persisted)
        self.refreshed = False

    def refresh_rotation(self):
        self.refreshed = True

    def added(self):
        return dict(self.added_counts)

    def set_added(self, heat, vapour, count):
        self.added_counts = {"heat": heat, "vapour": vapour, "count": count}


def test_saved_layout_is_the_references():
    g = C.small_grid()
    r = FakeRun(g)
    buf = save_state(r, config={"seed": 7})
    assert buf[:4] == b"MLAB"
    length = struct.unpack("<I", buf[4:8])[0]
    header = json.loads(buf[8:8 + length])
    assert header["format"] == "meteolab-state" and header["version"] == 5 and header["config"] == {"seed": 7}
    assert (header["nx"], header["ny"], header["nz"], header["time"]) == (g["nx"], g["ny"], g["nz"], 612.0)
    assert header["flashCount"] == {"ic": 3, "cg": 1} and header["chargeGround"] == 1.5 and header["persisted"] == 90
    assert header["added"] == {"heat": 2.5e13, "vapour": -1.25e6, "count": 3}
    start = math.ceil((8 + length) / 4) * 4
    n, layer = g["nx"] * g["ny"] * g["nz"], g["nx"] * g["ny"]
    assert len(buf) == start + (22 * n + 3 * layer) * 4
    data = np.frombuffer(buf, np.float32, offset=start)
    assert np.array_equal(data[:n], r.fields["u"]) and np.array_equal(data[6 * n:7 * n], r.fields["rain"])
    assert np.array_equal(data[16 * n:17 * n], r.fields["iceNumber"])
    assert not data[17 * n:22 * n].any()  # the two-moment numbers and volumes: not carried, zeros
    assert np.array_equal(data[22 * n:22 * n + layer], r.columns["precipitation"])
    h, d = read_state(buf)
    assert h == header and np.array_equal(d, data)


def test_apply_puts_fields_columns_time_and_counts_into_the_run():
    g = C.small_grid()
    src, dst = FakeRun(g, 1), FakeRun(g, 2)
    apply_state(dst, read_state(save_state(src)))
    for k in STATE_FIELDS[:17]:
        assert np.array_equal(dst.fields[k], src.fields[k]), k
    for k in STATE_COLUMNS:
        assert np.array_equal(dst.columns[k], src.columns[k]), k
    assert dst.time == 612.0 and dst.counts == (1.5, 3, 1, 90) and dst.refreshed
    assert dst.added_counts == {"heat": 2.5e13, "vapour": -1.25e6, "count": 3}


def test_opens_a_version_1_file_without_charge():
    # state.test.ts:55-66: 12 fields of ones; the charge fields and counts of a version 1 file are zero.
    g = C.small_grid()
    n, layer = g["nx"] * g["ny"] * g["nz"], g["nx"] * g["ny"]
    header = {"format": "meteolab-state", "version": 1, "config": {}, "time": 60, "persisted": 0, "nx": g["nx"], "ny": g["ny"], "nz": g["nz"]}
    buf = file_of(header, np.ones(12 * n + 3 * layer))
    run = FakeRun(g)
    run.fields["chargeIce"][:] = 5
    apply_state(run, read_state(buf))
    assert run.time == 60 and run.fields["hail"][0] == 1 and not run.fields["chargeIce"].any()
    assert run.counts == (0.0, 0, 0, 0)
    assert run.added_counts == {"heat": 0.0, "vapour": 0.0, "count": 0}
    assert FIELD_COUNT[1] == 12


def test_refuses_what_is_not_a_state_file_a_cut_one_and_a_different_grid():
    with pytest.raises(StateError, match="not a MeteoLab state file"):
        read_state(b"hello world, not a state")
    g = C.small_grid()
    buf = save_state(FakeRun(g))
    with pytest.raises(StateError, match="size does not match"):
        read_state(buf[:-4])
    with pytest.raises(StateError, match="truncated"):
        read_state(buf[:8] + b"{")
    bad = buf[:8] + b"{" * struct.unpack("<I", buf[4:8])[0] + buf[8 + struct.unpack("<I", buf[4:8])[0]:]
    with pytest.raises(StateError, match="header is damaged"):
        read_state(bad)
    h = {"format": "meteolab-state", "version": 9, "time": 0, "persisted": 0, "nx": 4, "ny": 4, "nz": 4}
    with pytest.raises(StateError, match="unknown file format"):
        read_state(file_of(h, []))
    other = FakeRun(dict(g, nx=20))
    with pytest.raises(StateError, match="grid is 16×12×24, the model’s 20×12×24"):
        apply_state(other, read_state(buf))


@pytest.mark.skipif(NODE is None, reason="node is absent")
def test_the_node_host_reads_and_writes_the_same_files(tmp_path):
    # The Node host's saveState / readState (node/meteolab.js) on the same layout: each reads the other's bytes.
    g = C.small_grid()
    r = FakeRun(g, 3)
    py = tmp_path / "py.meteolab"
    py.write_bytes(save_state(r, config={"seed": 3}))
    script = f"""
const m = require('./node/meteolab'), fs = require('fs')
const s = m.readState(fs.readFileSync('{py}'))
const g = s.header, n = g.nx * g.ny * g.nz, layer = g.nx * g.ny
const fake = {{ pool: {{ grid: {{ nx: g.nx, ny: g.ny, nz: g.nz }}, n, layer }}, spec: {{ seed: 3 }},
  diagnostics: () => ({{ time: g.time, chargeGround: g.chargeGround, flashesIC: g.flashCount.ic, flashesCG: g.flashCount.cg }}),
  rotation: () => ({{ persisted: g.persisted }}),
  added: () => g.added,
  field: k => s.data.slice(m.STATE_FIELDS.indexOf(k) * n, (m.STATE_FIELDS.indexOf(k) + 1) * n),
  column: k => s.data.slice(22 * n + m.STATE_COLUMNS.indexOf(k) * layer, 22 * n + (m.STATE_COLUMNS.indexOf(k) + 1) * layer) }}
fs.writeFileSync('{tmp_path / "node.meteolab"}', m.saveState(fake, {{ seed: 3 }}))
console.log(JSON.stringify({{ time: g.time, version: g.version, first: s.data[0] }}))
"""
    res = subprocess.run([NODE, "-e", script], capture_output=True, text=True, cwd=ROOT, timeout=120)
    assert res.returncode == 0, res.stderr
    info = json.loads(res.stdout)
    assert info["time"] == 612.0 and info["version"] == 5 and info["first"] == pytest.approx(float(r.fields["u"][0]))
    node_bytes = (tmp_path / "node.meteolab").read_bytes()
    h1, d1 = read_state(node_bytes)
    h0, d0 = read_state(py.read_bytes())
    assert h1 == h0 and np.array_equal(d1, d0)
    assert node_bytes == py.read_bytes()


@pytest.mark.gpu
def test_a_run_continued_from_its_saved_state_steps_exactly_as_the_original():
    # state.test.ts:8-23: 600 s into a run, saved, restored into a new run of the same settings, 5 more steps each.
    import meteolab_amd as M
    g = C.small_grid()
    with M.Context(0, 32) as ctx:
        original = ctx.pool(g, 1).run(C.profile(), C.run(g, seed=7))
        original.step(200)
        state = read_state(save_state(original))
        assert state[0]["time"] == 600.0
        restored = ctx.pool(g, 1).run(C.profile(), C.run(g, seed=7))
        apply_state(restored, state)
        assert restored.diagnostics()["time"] == 600.0
        # The rotation diagnostics are recomputed from the restored fields: those the last step left.
        assert restored.rotation() == original.rotation()
        original.step(5)
        restored.step(5)
        for name in ("w", "theta", "rain", "q"):
            assert np.array_equal(original.field(name), restored.field(name)), name
        assert np.array_equal(original.column("precipitation"), restored.column("precipitation"))


@pytest.mark.gpu
def test_a_run_continued_from_its_saved_state_flashes_as_the_original():
    # state.test.ts:25-52: a charged cloud above the breakdown threshold in dry air, a step so it has flashed, saved,
    # restored into a new run; three more steps each give the same flashes and charge.
    import meteolab_amd as M
    g = C.small_grid()
    dry = C.profile(qv_max=0.0)
    spec = C.run(g, seed=7, amp=0.0, noise_amp=0.0)
    with M.Context(0, 32) as ctx:
        pool = ctx.pool(g, 2)
        original = pool.run(dry, spec)
        nx, ny, nz = g["nx"], g["ny"], g["nz"]
        layer, dx, dy = nx * ny, g["width"] / nx, g["depth"] / ny
        rho = original.levels()["rho"]
        from oracle.binding import OracleRun
        zlev, _ = OracleRun(g, dry, spec).levels()  # the level heights
        x, y, z = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="xy")
        r2 = ((x - nx / 2) * dx) ** 2 + ((y - ny / 2) * dy) ** 2
        zz = zlev[z]
        ci = 3e-9 * np.exp(-r2 / 2500 ** 2 - ((zz - 9000) / 700) ** 2) / rho[z]
        cg = -3e-9 * np.exp(-r2 / 2500 ** 2 - ((zz - 6500) / 700) ** 2) / rho[z]
        original.set_field("chargeIce", np.transpose(ci, (2, 0, 1)).reshape(-1))
        original.set_field("chargeGraupel", np.transpose(cg, (2, 0, 1)).reshape(-1))
        original.step(1)
        before = original.diagnostics()
        assert before["flashesIC"] + before["flashesCG"] > 0
        buf = save_state(original)
        restored = pool.run(dry, spec)
        apply_state(restored, read_state(buf))
        after = restored.diagnostics()
        assert (after["flashesIC"], after["flashesCG"], after["chargeGround"]) == (before["flashesIC"], before["flashesCG"], before["chargeGround"])
        original.step(3)
        restored.step(3)
        a, b = original.diagnostics(), restored.diagnostics()
        assert (a["flashesIC"], a["flashesCG"], a["chargeGround"]) == (b["flashesIC"], b["flashesCG"], b["chargeGround"])
        fa = [(f["time"], f["cg"], f["charge"]) for f in original.flashes() if f["time"] > before["time"]]
        fb = [(f["time"], f["cg"], f["charge"]) for f in restored.flashes()]
        assert fa == fb
        assert np.array_equal(original.field("chargeGraupel"), restored.field("chargeGraupel"))
