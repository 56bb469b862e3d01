'use strict'
// Steps one run of a grid through the Node host and writes its carried fields (float32, carried order) to a file, and
// its diagnostics as one JSON line on stdout: the Node side of the boundary's parity test (tests/test_node_host.py).
//   node tests/node/step_fields.js <grid json> <steps> <out.bin>
const fs = require('fs')
const { MeteolabDevice, FIELDS, profile, run } = require('../../node/meteolab')

const grid = JSON.parse(process.argv[2]), steps = Number(process.argv[3]), out = process.argv[4]
const dev = MeteolabDevice.create()
const pool = dev.pool(grid, 1)
const r = pool.run(profile(), run(grid))
pool.step([r], steps, 3)
dev.sync()
const carried = FIELDS.slice(0, 17), n = pool.n, all = new Float32Array(carried.length * n)
carried.forEach((name, k) => all.set(r.field(name), k * n))
fs.writeFileSync(out, Buffer.from(all.buffer))
console.log(JSON.stringify({ adapter: dev.adapter, diagnostics: r.diagnostics(), errors: dev.errors }))
pool.destroy()
dev.destroy()
