'use strict'
// Node host of the MI355X StormLab step: the classes a caller of AtmosphereModel.step (src/core/model.ts:873-877) or of
// the GPU plugin GpuDynamics (src/gpu/dynamics.ts:250-504) uses, over the N-API addon of include/meteolab.h.
//
//   const { MeteolabDevice, STANDARD, profile, run } = require('./meteolab')
//   const dev = MeteolabDevice.create()            // GpuDynamics.create: a device and its stream
//   const pool = dev.pool(STANDARD, 64)            // slots for 64 standard runs, resident in HBM
//   const r = pool.run(profile(), run(STANDARD))   // new AtmosphereModel(...) for the step's state
//   pool.step([r], 20, 3)                          // 20 x (step(dt); time += dt), one launch per kernel for the batch
//   const w = r.field('w')                         // Float32Array, node i = x + nx (y + ny z)
//
// Errors of the device are thrown as Error and kept in `errors` (GpuDynamics.errors, gpu/dynamics.ts:252).
const path = require('path')
const addon = require(path.join(__dirname, 'meteolab.node'))

// The carried order of AtmosphereModel.carried() (model.ts:1319-1335), then the arrays a step recomputes; the ground
// columns of STATE_COLUMNS (src/core/state.ts:10).
const FIELDS = ['rain', 'snow', 'graupel', 'hail', 'u', 'v', 'w', 'theta', 'q', 'cloud', 'ice', 'iceNumber', 'cold',
  'chargeIce', 'chargeGraupel', 'chargeRain', 'chargeAir', 'fieldZ', 'fallRain', 'fallSnow', 'fallGraupel', 'fallHail',
  'fallChargeIce', 'fallChargeGraupel', 'condensate']
const COLUMNS = ['precipitation', 'hailGround', 'hailSize', 'gustMax', 'uhMax']
// InterventionKind (src/core/intervene.ts:13) in ml_run_intervene's order.
const INTERVENTIONS = ['thermal', 'warm', 'cool', 'moisten', 'dry', 'coldPool', 'vortex']
const DIAG = ['time', 'maxW', 'minW', 'maxThetaP', 'rainTotal', 'water', 'charge', 'clipped', 'negativeFilled',
  'chargeGround', 'flashesIC', 'flashesCG', 'firstFlash', 'error']
const LEVELS = ['p', 'exner', 'theta', 'q', 'u', 'v', 'rho']
// RotationState (model.ts:142-158) in ml_run_rotation's order.
// ModelDiagnostics (model.ts:161-180) in ml_model_diagnostics' order.
const MODEL_DIAGNOSTICS = ['updraft', 'downdraft', 'rainRate', 'rainTotal', 'cloudTop', 'thermalTop', 'maxCloud', 'coldMax',
  'cores', 'shear06', 'outflow', 'outflowLow', 'gust', 'hail', 'clipped']
const ROTATION = ['uh', 'x', 'y', 'anticyclonic', 'persisted', 'uh01', 'uh03', 'low', 'coreZeta', 'coreUH', 'coreAnti']
// The parts of a step where the reference's GPU step splits it (src/gpu/step.ts:12-31).
const PARTS = ['dynamics', 'physics', 'mix', 'afterStep']
const CARRIED = FIELDS.slice(0, 17)

// src/core/grid.ts:31 DEFAULT_GRID and :34 LARGE_GRID.
const STANDARD = { nx: 40, ny: 32, nz: 50, width: 48000, depth: 36000, height: 18900, bottom: 100 }
const LARGE = { nx: 80, ny: 64, nz: 50, width: 96000, depth: 72000, height: 18900, bottom: 100 }

// The Weisman-Klemp sounding (src/core/profiles.ts:61-73) with the wind u = shear tanh(z / shear depth).
function profile (qvMax = 0.014, shear = 25, shearDepth = 3000) {
  return { theta0: 300, dtheta: 43, ztr: 12000, ttr: 213, rh_drop: 0.75, rh_top: 0.25, qv_max: qvMax, shear, shear_depth: shearDepth }
}

// A run: the cos^2 thermal of model.ts:94 (ELLIPSOID) of `amp` K, a third of the way in from the west, shifted by
// `offset` m, and +-0.1 K theta noise below 1 km from mulberry32(seed).
function run (grid, seed = 1, amp = 2, offset = [0, 0]) {
  return { seed, latitude: 45, turbulence: 0.55, tropopause_km: 12, roughness: 0.03, frame_u: 0, frame_v: 0, bubble_amp: amp,
    bubble_x: grid.width / 3 + offset[0], bubble_y: grid.depth / 2 + offset[1], bubble_z: 1400, bubble_rh: 10000, bubble_rz: 1400,
    noise_seed: seed, noise_amp: 0.1, noise_top: 1000 }
}

class MeteolabDevice {
  constructor (ctx) { this.ctx = ctx; this.errors = []; this.adapter = addon.deviceName(ctx) }
  /** A context on HIP device `device`; precision 32 (production kernels) or 64 (the same in float64). */
  static create (device = 0, precision = 32) { return new MeteolabDevice(addon.ctxCreate(device, precision)) }
  /** Slots for `capacity` runs of one grid; electrified runs carry the charges and the lightning. */
  pool (grid, capacity, electrified = true) { return new MeteolabPool(this, grid, capacity, electrified) }
  /** Device bytes one run of this grid takes (admission by HBM footprint). */
  slotBytes (grid) { return addon.slotBytes(this.ctx, grid) }
  sync () { this.call(() => addon.sync(this.ctx)) }
  call (f) {
    try { return f() } catch (e) { this.errors.push(e.message); throw e }
  }
  destroy () { if (this.ctx) { addon.ctxDestroy(this.ctx); this.ctx = null } }
}

class MeteolabPool {
  constructor (device, grid, capacity, electrified) {
    this.device = device; this.grid = { ...grid }
    this.n = grid.nx * grid.ny * grid.nz; this.layer = grid.nx * grid.ny
    this.id = device.call(() => addon.poolCreate(device.ctx, grid, capacity, electrified ? 1 : 0))
  }
  run (prof, config) { return new MeteolabRun(this, prof, config) }
  /** Advances the runs (all of this pool) by `steps` steps of dt s; asynchronous, sync() waits. */
  step (runs, steps = 1, dt = 3) {
    const ids = Int32Array.from(runs.map(r => r.handle))
    this.device.call(() => addon.step(this.device.ctx, ids, steps, dt))
  }
  /** One part of a step of the runs: 'dynamics', 'physics', 'mix' or 'afterStep' (in that order they are a step). */
  part (runs, name, dt = 3) {
    if (!PARTS.includes(name)) throw new Error(`step part ${name}: one of ${PARTS.join(', ')}`)
    const ids = Int32Array.from(runs.map(r => r.handle))
    this.device.call(() => addon[name](this.device.ctx, ids, dt))
  }
  /** AtmosphereModel.diagnostics() of each run (one launch for all): an array of ModelDiagnostics objects. */
  diagnostics (runs) {
    const ids = Int32Array.from(runs.map(r => r.handle)), out = new Float64Array(runs.length * MODEL_DIAGNOSTICS.length)
    this.device.call(() => addon.modelDiagnostics(this.device.ctx, ids, out))
    return runs.map((_, k) => Object.fromEntries(MODEL_DIAGNOSTICS.map((name, j) => [name, out[k * MODEL_DIAGNOSTICS.length + j]])))
  }
  /**
   * probabilityMap (src/ensemble/combine.ts:70-75) on the device: per ground column the share of the runs whose swath
   * column ('precipitation' as the rain total, 'hailSize', 'gustMax', 'uhMax') reached the threshold there.
   */
  probabilityMap (runs, column, threshold) {
    const ids = Int32Array.from(runs.map(r => r.handle)), out = new Float32Array(this.layer)
    this.device.call(() => addon.probabilityMap(this.device.ctx, ids, COLUMNS.indexOf(column), threshold, out))
    return out
  }
  destroy () { if (this.id !== null) { this.device.call(() => addon.poolDestroy(this.device.ctx, this.id)); this.id = null } }
}

class MeteolabRun {
  constructor (pool, prof, config) {
    this.pool = pool; this.device = pool.device
    this.handle = this.device.call(() => addon.runCreate(this.device.ctx, pool.id, prof, config))
    this.spec = { grid: { ...pool.grid }, profile: { ...prof }, run: { ...config } }
  }
  /** A field (or fieldZ, a fall speed, the condensate) as a new Float32Array. */
  field (name) {
    const out = new Float32Array(this.pool.n)
    this.device.call(() => addon.download(this.device.ctx, this.handle, FIELDS.indexOf(name), out))
    return out
  }
  setField (name, values) {
    const a = Float32Array.from(values)
    if (a.length !== this.pool.n) throw new Error(`${name}: ${a.length} values for ${this.pool.n} nodes`)
    this.device.call(() => addon.upload(this.device.ctx, this.handle, FIELDS.indexOf(name), a))
  }
  column (name) {
    const out = new Float32Array(this.pool.layer)
    this.device.call(() => addon.downloadColumn(this.device.ctx, this.handle, COLUMNS.indexOf(name), out))
    return out
  }
  setColumn (name, values) {
    const a = Float32Array.from(values)
    if (a.length !== this.pool.layer) throw new Error(`${name}: ${a.length} values for ${this.pool.layer} columns`)
    this.device.call(() => addon.uploadColumn(this.device.ctx, this.handle, COLUMNS.indexOf(name), a))
  }
  /** AtmosphereModel.diagnostics() (model.ts:1672-1696) of the run's fields as they are: a ModelDiagnostics object. */
  modelDiagnostics () { return this.pool.diagnostics([this])[0] }
  /** The run metrics (ml_run_diag): time, max w, max theta', rain total, water, charge, flashes, first flash. */
  diagnostics () {
    const out = new Float64Array(DIAG.length)
    this.device.call(() => addon.diag(this.device.ctx, this.handle, out))
    return Object.fromEntries(DIAG.map((k, i) => [k, out[i]]))
  }
  /** The base state per level (environment.ts:124-130). */
  /** The last projection's pressure cells, (nz - 1) nx ny float64 (model.ts:358 pressure). */
  pressure () {
    const g = this.pool.grid, out = new Float64Array((g.nz - 1) * g.nx * g.ny)
    this.device.call(() => addon.pressure(this.device.ctx, this.handle, out))
    return out
  }
  levels () {
    const nz = this.pool.grid.nz, out = new Float64Array(LEVELS.length * nz)
    this.device.call(() => addon.levels(this.device.ctx, this.handle, out))
    return Object.fromEntries(LEVELS.map((k, i) => [k, out.subarray(i * nz, (i + 1) * nz)]))
  }
  /** The most recent flashes (lightning.ts Flash: time, kind, start, nodes, charge, potential, moment). */
  flashes (max = 64) {
    const out = new Float64Array(8 * max), n = this.device.call(() => addon.flashes(this.device.ctx, this.handle, out))
    const list = []
    for (let k = 0; k < n; k++) {
      const e = out.subarray(8 * k, 8 * k + 7)
      list.push({ time: e[0], kind: e[1] ? 'CG' : 'IC', start: e[2], nodes: e[3], charge: e[4], potential: e[5], moment: e[6] })
    }
    return list
  }
  /** fallInputs' sub-steps of each fall of the last step or dynamics part (rain, snow, graupel, hail, then charges). */
  fallParts () {
    const out = new Int32Array(8), k = this.device.call(() => addon.fallParts(this.device.ctx, this.handle, out))
    return Array.from(out.subarray(0, k))
  }
  set time (t) { this.device.call(() => addon.setTime(this.device.ctx, this.handle, t)) }
  /** A saved state's header totals: charge carried to the ground (C), IC and CG flash counts, rotation held (s). */
  setTotals (chargeGround, ic, cg, persisted = 0) { this.device.call(() => addon.setTotals(this.device.ctx, this.handle, chargeGround, ic, cg, persisted)) }
  /** The rotation diagnostics of the last step (model.ts RotationState): uh, x, y, anticyclonic, persisted, ... */
  rotation () {
    const out = new Float64Array(ROTATION.length)
    this.device.call(() => addon.rotation(this.device.ctx, this.handle, out))
    return Object.fromEntries(ROTATION.map((k, i) => [k, out[i]]))
  }
  /**
   * intervene (src/core/intervene.ts:38-68; the worker's 'intervene' message, src/simulation.worker.ts:219) on the
   * run's fields: { kind, x, y, strength } as the reference's Intervention; thermalRadius is SimConfig.bubbleRadius (km,
   * absent: 4.2). Returns { heat, vapour } put in (J, kg), also added to added().
   */
  intervene (it, thermalRadius = 0) {
    const kind = INTERVENTIONS.indexOf(it.kind)
    if (kind < 0) throw new Error(`intervention kind ${it.kind}: one of ${INTERVENTIONS.join(', ')}`)
    const out = new Float64Array(2)
    const strength = it.strength === undefined ? 1 : it.strength
    this.device.call(() => addon.intervene(this.device.ctx, this.handle, kind, it.x, it.y, strength, thermalRadius || 0, out))
    return { heat: out[0], vapour: out[1] }
  }
  /** AtmosphereModel.added: { heat, vapour, count } put in by interventions. */
  added () {
    const out = new Float64Array(3)
    this.device.call(() => addon.added(this.device.ctx, this.handle, out))
    return { heat: out[0], vapour: out[1], count: out[2] }
  }
  setAdded (a) { this.device.call(() => addon.setAdded(this.device.ctx, this.handle, a.heat || 0, a.vapour || 0, a.count || 0)) }
  /** addSwaths after every step (src/ensemble/member.ts:85-100) into the gustMax and uhMax columns. */
  setSwaths (on) { this.device.call(() => addon.setSwaths(this.device.ctx, this.handle, on ? 1 : 0)) }
  /** The rotation diagnostics of the fields as they are, keeping how long the rotation has held (refreshRotation). */
  refreshRotation () { this.device.call(() => addon.refreshRotation(this.device.ctx, this.handle)) }
  release () { if (this.handle !== null) { this.device.call(() => addon.runRelease(this.device.ctx, this.handle)); this.handle = null } }
}

/**
 * The reference's GPU plugin interface (GpuDynamics, src/gpu/dynamics.ts:250-504) over one resident run, so that
 * gpuStep (src/gpu/step.ts:12-31) and the worker's fallback (simulation.worker.ts:104-112) drive this device unchanged:
 * step() is the sedimentation and the transport (ml_dynamics), mix() the turbulence, friction and projection (ml_mix);
 * the physics and afterStep stay with the model on the host between them. Errors reject the promise and are kept in
 * `errors`, as the reference's are. stepN() is a whole step on the device (the state stays resident).
 *
 * The device computes the fall speeds and sub-steps of `falls` itself from the same fields (the model's fallInputs
 * are the same numbers); the ground deposits of all falls that land in the rain gauge come back as the first such
 * fall's array, the charge carried to the ground as the first charge fall's column 0 (landed() sums them the same).
 * The baseline carried set (17 fields, AtmosphereModel.carried() order) is the only one supported.
 */
class GpuDynamics {
  constructor (device, grid, env, prof, config) {
    this.device = device; this.grid = { ...grid }; this.env = env
    this.pool = device.pool(grid, 1); this.run = this.pool.run(prof, config)
    this.errors = []; this.adapter = device.adapter
  }
  /** GpuDynamics.create(grid, env): a device and one run slot; `options` names the run's sounding and thermal. */
  static async create (grid, env, options = {}) {
    const device = MeteolabDevice.create(options.device || 0, options.precision || 32)
    return new GpuDynamics(device, grid, env, options.profile || profile(), options.config || run(grid))
  }
  async step (fields, env, dt, timed = false, falls = []) {
    try {
      // The baseline's DynamicsEnv only: the device steps the base state it was created with; zones and nudging
      // (gpu/dynamics.ts:237-244) are not part of this step, and are refused rather than dropped.
      if (env && (env.zoneTheta || env.zoneVapour || env.nudge)) throw new Error('GpuDynamics.step: DynamicsEnv zones and nudging are not supported')
      if (env && env.shape && Array.from(env.shape).some(s => s !== 0)) throw new Error('GpuDynamics.step: DynamicsEnv zone shape is not supported')
      const t0 = Date.now(), r = this.run, layer = this.pool.layer
      this.load(fields)
      r.setColumn('precipitation', new Float32Array(layer))
      const charged = r.diagnostics().chargeGround
      this.pool.part([r], 'dynamics', dt)
      const out = fields.map((_, k) => r.field(CARRIED[k]))
      const rain = r.column('precipitation'), toGround = r.diagnostics().chargeGround - charged
      const deposit = falls.findIndex(f => f.deposit !== false), charge = falls.findIndex(f => f.deposit === false)
      const ground = falls.map((_, j) => {
        if (j === deposit) return rain
        const g = new Float32Array(layer)
        if (j === charge) g[0] = toGround / (this.grid.dx * this.grid.dy)
        return g
      })
      return { fields: out, ground, gpuMs: Date.now() - t0 }
    } catch (e) { this.errors.push(String(e && e.message ? e.message : e)); throw e }
  }
  async mix (fields, mix, dt) {
    try {
      const t0 = Date.now(), r = this.run
      for (const { k } of mix.mixed) r.setField(CARRIED[k], fields[k].q)
      r.setField('condensate', mix.condensate)
      this.pool.part([r], 'mix', dt)
      // The projection's pressure cells, as the model's pressure array (endGpuStep copies them, model.ts:1145-1150).
      return { fields: mix.mixed.map(({ k }) => r.field(CARRIED[k])), pressure: r.pressure(), gpuMs: Date.now() - t0 }
    } catch (e) { this.errors.push(String(e && e.message ? e.message : e)); throw e }
  }
  /** Whole steps on the device from the resident state (or `fields` first): pool.step. */
  stepN (n, dt, fields = null) {
    if (fields) this.load(fields)
    this.pool.step([this.run], n, dt)
    this.device.sync()
  }
  /** Whether this instance fits a grid and base state: the same dimensions and spacing, and the device's base state
   * equal to the model's within float64 rounding of the same formulas. */
  matches (grid, env) {
    const g = this.grid, lv = this.run.levels()
    const close = (a, b) => a.length === b.length && Array.from(a).every((x, k) => Math.abs(x - b[k]) <= 1e-9 * Math.abs(b[k]))
    return g.nx === grid.nx && g.ny === grid.ny && g.nz === grid.nz && g.width === grid.width && g.depth === grid.depth &&
      close(lv.rho, env.rho) && close(lv.theta, env.theta) && close(lv.exner, env.exner)
  }
  destroy () { this.run.release(); this.pool.destroy(); this.device.destroy() }
  load (fields) {
    if (fields.length !== CARRIED.length) throw new Error(`GpuDynamics: ${fields.length} carried fields, the baseline set has ${CARRIED.length}`)
    fields.forEach((f, k) => this.run.setField(CARRIED[k], f.q || f))
  }
}

// The .meteolab state file (src/core/state.ts:1-110): 'MLAB', the header's length (uint32 LE), the JSON header,
// padding to 4 bytes, float32 fields in STATE_FIELDS order, then the STATE_COLUMNS. Version 5 is written; the
// two-moment numbers and volumes (not carried by this step) and the water and charge budgets (not kept) are zeros; the
// interventions' counts (header.added) are the run's.
const STATE_FIELDS = ['u', 'v', 'w', 'theta', 'q', 'cloud', 'rain', 'cold', 'ice', 'snow', 'graupel', 'hail', 'chargeIce',
  'chargeGraupel', 'chargeRain', 'chargeAir', 'iceNumber', 'rainNumber', 'graupelNumber', 'graupelVolume', 'hailNumber', 'hailVolume']
const STATE_COLUMNS = ['precipitation', 'hailGround', 'hailSize']
const STATE_VERSION = 5
const FIELD_COUNT = { 1: 12, 2: 16, 3: 17, 4: 18, 5: 22 }
const UNCARRIED = new Set(['rainNumber', 'graupelNumber', 'graupelVolume', 'hailNumber', 'hailVolume'])

/** saveState (state.ts:34-45): a run's state as the bytes of a .meteolab file (a Buffer). */
function saveState (r, config = null) {
  const g = r.pool.grid, n = r.pool.n, layer = r.pool.layer, d = r.diagnostics()
  const header = { format: 'meteolab-state', version: STATE_VERSION, config: config === null ? r.spec : config, time: d.time, persisted: r.rotation().persisted,
    nx: g.nx, ny: g.ny, nz: g.nz, chargeGround: d.chargeGround, flashCount: { ic: d.flashesIC, cg: d.flashesCG },
    water: { condensed: 0, evaporated: 0, rainEvaporated: 0, surface: 0, start: 0 }, added: r.added(),
    charge: { separated: 0, ic: 0, cgNegative: 0, cgPositive: 0, precipitation: 0, conduction: 0 } }
  const json = Buffer.from(JSON.stringify(header)), start = Math.ceil((8 + json.length) / 4) * 4
  const buffer = Buffer.alloc(start + (STATE_FIELDS.length * n + STATE_COLUMNS.length * layer) * 4)
  buffer.write('MLAB', 0, 'latin1'); buffer.writeUInt32LE(json.length, 4); json.copy(buffer, 8)
  const data = new Float32Array(buffer.buffer, buffer.byteOffset + start, STATE_FIELDS.length * n + STATE_COLUMNS.length * layer)
  STATE_FIELDS.forEach((k, j) => { if (!UNCARRIED.has(k)) data.set(r.field(k), j * n) })
  STATE_COLUMNS.forEach((k, j) => data.set(r.column(k), STATE_FIELDS.length * n + j * layer))
  return buffer
}

/** readState (state.ts:48-68): { header, data } of a .meteolab file; throws with the reference's messages. */
function readState (buffer) {
  if (buffer.length < 8 || buffer.toString('latin1', 0, 4) !== 'MLAB') throw new Error('this is not a MeteoLab state file')
  const length = buffer.readUInt32LE(4)
  if (8 + length > buffer.length) throw new Error('the file is truncated')
  let header
  try { header = JSON.parse(buffer.toString('utf8', 8, 8 + length)) } catch (e) { throw new Error('the file’s header is damaged') }
  if (!header || header.format !== 'meteolab-state' || !(header.version in FIELD_COUNT)) throw new Error('unknown file format or version')
  for (const k of ['time', 'persisted', 'nx', 'ny', 'nz']) if (!Number.isFinite(header[k]) || header[k] < 0) throw new Error('the file’s header is damaged')
  const start = Math.ceil((8 + length) / 4) * 4, cells = header.nx * header.ny, size = FIELD_COUNT[header.version] * cells * header.nz + STATE_COLUMNS.length * cells
  if (buffer.length !== start + size * 4) throw new Error('the file’s size does not match the grid')
  const data = new Float32Array(size)
  for (let k = 0; k < size; k++) data[k] = buffer.readFloatLE(start + 4 * k)
  return { header, data }
}

/** applyState (state.ts:74-110) into a run of the same grid: fields, columns, model time, lightning totals. */
function applyState (r, state) {
  const g = r.pool.grid, h = state.header, n = r.pool.n, layer = r.pool.layer
  if (h.nx !== g.nx || h.ny !== g.ny || h.nz !== g.nz) throw new Error(`the file’s grid is ${h.nx}×${h.ny}×${h.nz}, the model’s ${g.nx}×${g.ny}×${g.nz}`)
  const held = FIELD_COUNT[h.version]
  STATE_FIELDS.forEach((k, j) => {
    if (UNCARRIED.has(k)) return
    r.setField(k, j < held ? state.data.subarray(j * n, (j + 1) * n) : new Float32Array(n))
  })
  STATE_COLUMNS.forEach((k, j) => r.setColumn(k, state.data.subarray(held * n + j * layer, held * n + (j + 1) * layer)))
  r.time = h.time
  const c = h.flashCount || {}
  r.setTotals(h.chargeGround || 0, c.ic || 0, c.cg || 0, h.persisted)
  r.setAdded(h.added || {})
  // The rotation diagnostics are not in the file: recomputed from its fields (state.ts:73-74).
  r.refreshRotation()
}

/**
 * runMember (src/ensemble/member.ts:103-117) for every run of a batch of one pool at once: the swaths folded in after
 * every step on the device, and every 60 model seconds a sample of each run's strongest values from one launch for the
 * batch. Returns MemberResult objects (member.ts:17-22), which src/ensemble/combine.ts (memberPeaks, probabilities,
 * probabilityMap) and src/regime/grid.ts (regimeOf) read as they are.
 */
function runMembers (pool, runs, minutes, dt = 3) {
  const SAMPLE_EVERY = 60, samples = runs.map(() => []), flashes = runs.map(() => 0)
  runs.forEach(r => r.setSwaths(true))
  const sample = () => {
    const ds = pool.diagnostics(runs)
    runs.forEach((r, k) => {
      const d = ds[k], rot = r.rotation(), m = r.diagnostics(), total = m.flashesIC + m.flashesCG
      samples[k].push({ t: m.time, updraft: d.updraft, uh: rot.uh, rain: d.rainRate, gust: d.gust, hail: d.hail, flashes: (total - flashes[k]) * 60 / SAMPLE_EVERY,
        rotationHeld: rot.persisted, cores: d.cores, outflow: d.outflow, outflowLow: d.outflowLow })
      flashes[k] = total
    })
  }
  sample()
  const per = Math.round(SAMPLE_EVERY / dt), end = minutes * 60
  let t = 0
  while (t < end - 1e-6) {
    const n = Math.min(per, Math.ceil((end - t) / dt - 1e-9))
    pool.step(runs, n, dt); t += n * dt
    pool.device.sync()
    sample()
  }
  const g = pool.grid
  return runs.map((r, k) => ({ seed: r.spec.run.seed, samples: samples[k], nx: g.nx, ny: g.ny, width: g.width, depth: g.depth,
    rainTotal: r.column('precipitation'), hailSize: r.column('hailSize'), gustMax: r.column('gustMax'), uhMax: r.column('uhMax') }))
}

// mulberry32 (the published 32-bit generator the reference uses, src/core/math.ts:7-14): uniforms in [0, 1).
function mulberry32 (seed) {
  let s = seed >>> 0
  return () => {
    s = (s + 0x6D2B79F5) >>> 0
    let t = Math.imul(s ^ (s >>> 15), 1 | s)
    t = (t + Math.imul(t ^ (t >>> 7), 61 | t)) ^ t
    return ((t ^ (t >>> 14)) >>> 0) / 4294967296
  }
}

// The starting data's uncertainty (src/ensemble/member.ts:33): K, % relative humidity, m/s, degrees, the thermal's
// relative strength, m.
const UNCERTAINTY = { temperature: 0.5, humidity: 4, wind: 1.5, direction: 8, thermal: 0.1, offset: 1500 }

// perturbConfig (src/ensemble/member.ts:45-65) on the analytic sounding and a run: new [profile, run] objects, the
// inputs untouched. Normal deviates by Box-Muller from mulberry32(seed ^ 0x5bd1e995), in the reference's order: the
// ground temperature (theta0), the humidity (the vapour cap times 1 + dRh / 100), the wind errors at 0, 3, 6 and 10 km
// (u, then v, added to the sounding's wind_u / wind_v), a speed and a direction error for each wind slider node in
// `sliders` (drawn, not applied), the thermal's strength and its place from the default one.
// The wind slider nodes SUMMER_DAY-based configurations carry (src/core/fixtures.ts:9: wind0, 3, 6, 10 and their
// directions).
const SLIDER_NODES = ['0', '3', '6', '10']
function perturbRun (grid, prof, spec, sliders = SLIDER_NODES) {
  const rng = mulberry32(spec.seed ^ 0x5bd1e995), U = UNCERTAINTY
  const n = () => { const a = rng(); return Math.sqrt(-2 * Math.log(1 - a)) * Math.cos(2 * Math.PI * rng()) }
  const dT = n() * U.temperature, dRh = n() * U.humidity
  const du = [0, 1, 2, 3].map(() => n() * U.wind), dv = [0, 1, 2, 3].map(() => n() * U.wind)
  // The slider wind nodes' speed and direction errors (member.ts:50-54), drawn and not applied: the analytic sounding
  // has no sliders, but the draws keep the thermal's in the reference's places of the stream.
  for (const k of ['0', '05', '1', '3', '6', '10']) if (sliders.includes(k)) { n(); n() }
  const zero = [0, 0, 0, 0], wu = prof.wind_u || zero, wv = prof.wind_v || zero
  const p = Object.assign({}, prof, { theta0: prof.theta0 + dT, qv_max: prof.qv_max * (1 + dRh / 100),
    wind_u: wu.map((x, k) => x + du[k]), wind_v: wv.map((x, k) => x + dv[k]) })
  const amp = spec.bubble_amp * Math.max(0.5, 1 + n() * U.thermal)
  const ox = n() * U.offset, oy = n() * U.offset
  return [p, Object.assign({}, spec, { bubble_amp: amp, bubble_x: grid.width / 3 + ox, bubble_y: grid.depth / 2 + oy })]
}

module.exports = { MeteolabDevice, runMembers, MeteolabPool, MeteolabRun, GpuDynamics, FIELDS, CARRIED, COLUMNS, INTERVENTIONS, DIAG, PARTS, STANDARD, LARGE, profile, run, addon,
  STATE_FIELDS, STATE_COLUMNS, ROTATION, MODEL_DIAGNOSTICS, saveState, readState, applyState, mulberry32, UNCERTAINTY, SLIDER_NODES, perturbRun }
