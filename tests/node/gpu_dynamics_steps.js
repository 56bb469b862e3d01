'use strict'
// Steps one run through the GpuDynamics interface as gpuStep (src/gpu/step.ts:12-31) calls it: step() for the
// sedimentation and transport, the model's physicsStep between (here the device's physics part stands in for the
// host's), mix() for the turbulence, friction and projection, then afterStep. Writes the carried fields (float32,
// carried order) and the last mix()'s pressure cells (float64) to files, and one JSON line: the errors, a forced error's
// rejection, the refusal of zones and nudging, and the diagnostics.
//   node tests/node/gpu_dynamics_steps.js <grid json> <steps> <out.bin> <pressure.bin>
const fs = require('fs')
const { GpuDynamics, CARRIED } = require('../../node/meteolab')

const MIXED = ['u', 'v', 'w', 'theta', 'q', 'cloud', 'ice', 'iceNumber']  // AtmosphereModel.mixed() (model.ts:1338-1344)

async function main () {
  const grid = JSON.parse(process.argv[2]), steps = Number(process.argv[3]), out = process.argv[4], pressureOut = process.argv[5]
  const levels = { rho: [], theta: [], exner: [] }
  const gd = await GpuDynamics.create(grid, levels)
  const r = gd.run, lv = r.levels()
  const env = { rho: lv.rho, theta: lv.theta, exner: lv.exner }
  const falls = ['rain', 'snow', 'graupel', 'hail', 'chargeIce', 'chargeGraupel', 'chargeRain'].map((f, j) => ({ k: CARRIED.indexOf(f), vt: null, parts: 1, deposit: j < 4 }))
  let precipitation = new Float32Array(gd.pool.layer), pressure = null
  for (let s = 0; s < steps; s++) {
    const model = CARRIED.map(name => ({ q: r.field(name), monotone: true, role: 'other' }))
    const first = await gd.step(model, {}, 3, false, falls)
    first.fields.forEach((f, k) => { model[k].q = f })                      // setCarried
    const landed = first.ground[0]; for (let c = 0; c < landed.length; c++) precipitation[c] += landed[c]  // landed
    gd.pool.part([r], 'physics', 3)                                        // physicsStep
    model.forEach((f, k) => { f.q = r.field(CARRIED[k]) })
    const mixing = { mixed: MIXED.map(name => ({ k: CARRIED.indexOf(name), base: null, factor: 1 })), condensate: r.field('condensate') }
    const second = await gd.mix(model, mixing, 3)
    mixing.mixed.forEach(({ k }, j) => { model[k].q = second.fields[j] })  // endGpuStep
    pressure = second.pressure
    gd.pool.part([r], 'afterStep', 3)                                      // afterStep (time += dt)
  }
  r.setColumn('precipitation', precipitation)
  const all = new Float32Array(CARRIED.length * gd.pool.n)
  CARRIED.forEach((name, k) => all.set(r.field(name), k * gd.pool.n))
  fs.writeFileSync(out, Buffer.from(all.buffer))
  fs.writeFileSync(pressureOut, Buffer.from(pressure.buffer))
  const info = { errors: gd.errors.slice(), matches: gd.matches(grid, env), diagnostics: r.diagnostics(), fallParts: r.fallParts() }
  // A forced error (a field set of the wrong size) rejects, and the error is kept, so gpuStep finishes on the CPU.
  info.rejected = await gd.step([{ q: new Float32Array(3) }], {}, 3).then(() => false, () => true)
  info.errorsAfter = gd.errors.length
  // A DynamicsEnv with nudging or zones is refused, not silently dropped.
  const model = CARRIED.map(name => ({ q: r.field(name) }))
  info.nudgeRefused = await gd.step(model, { nudge: { time: 0, top: 1, cx: 0, cy: 0 } }, 3).then(() => false, e => /nudging/.test(e.message))
  info.zonesRefused = await gd.step(model, { zoneTheta: new Float32Array(gd.pool.layer) }, 3).then(() => false, e => /zones/.test(e.message))
  gd.destroy()
  console.log(JSON.stringify(info))
}
main().catch(e => { console.error(e); process.exit(1) })
