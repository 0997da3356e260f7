'use strict';
/**
 * Checks of the N-API binding, run by the Python test suite:
 *   node check.js cpu <fixture.json>   load, version, cell order, named errors
 *   node check.js gpu <fixture.json>   the 10-particle OpenMM fixture through
 *                                      HipEngine: energies, a Promise run
 * <fixture.json>: {desc, positions, facts} written by the test.
 */
const assert = require('assert');
const fs = require('fs');
const { HipEngine, native } = require('./engine.js');

function typed(desc) {
  const out = {};
  for (const [k, v] of Object.entries(desc)) {
    if (Array.isArray(v)) {
      const ints = k.endsWith('_p1') || k.endsWith('_p2') || k.endsWith('_p3') || k.endsWith('_p4') ||
        k === 'torsion_periodicity' || k === 'restraint_particle' || k === 'grid';
      out[k] = ints ? Int32Array.from(v) : Float64Array.from(v);
    } else {
      out[k] = v;
    }
  }
  return out;
}

function expectError(fn, name) {
  let caught = null;
  try { fn(); } catch (e) { caught = e; }
  assert(caught, `expected ${name}`);
  assert.strictEqual(caught.name, name, `${caught.name}: ${caught.message}`);
  return caught;
}

async function main() {
  const [mode, file] = process.argv.slice(2);
  const fixture = JSON.parse(fs.readFileSync(file, 'utf8'));
  const desc = typed(fixture.desc);
  if (mode === 'cpu') {
    assert(native.version().includes('gfx950'));
    const cells = native.probeCellOrder(Int32Array.from([4, 4, 4]));
    assert.strictEqual(cells.length, 64);
    assert.deepStrictEqual(Array.from(cells).sort((a, b) => a - b), Array.from({ length: 64 }, (_, i) => i));
    // refusals raised on the host, before any device is touched
    const small = Object.assign({}, desc, { box: Float64Array.from([1.5, 0, 0, 0, 1.5, 0, 0, 0, 1.5]) });
    expectError(() => native.createSystem(small, { time_step: 0.002 }), 'CutoffTooLargeForBox');
    const odd = Object.assign({}, desc, { grid: Int32Array.from([34, 32, 32]) });
    expectError(() => native.createSystem(odd, { time_step: 0.002 }), 'NotSmooth');
    expectError(() => native.createSystem(Object.assign({}, desc, { n_particles: 0 }), { time_step: 0.002 }),
      'InvalidArgument');
    // the barostat binding checks its handle before touching the library
    expectError(() => native.attachBarostat({}, { frequency: 25 }), 'TypeError');
    // no GPU here: a valid System fails loudly, with no fallback
    expectError(() => new HipEngine(desc, { positions: Float64Array.from(fixture.positions) }), 'NoDevice');
    console.log('node binding (cpu): ok');
    return;
  }
  const engine = new HipEngine(desc, { positions: Float64Array.from(fixture.positions) });
  const e = engine.evaluate();
  const rel = (a, b) => Math.abs(a - b) / Math.abs(b);
  assert(rel(e.groups[3], fixture.facts.direct_group) < 1e-5, `direct ${e.groups[3]}`);
  assert(rel(e.groups[4], fixture.facts.reciprocal_group) < 1e-5, `reciprocal ${e.groups[4]}`);
  const s0 = engine.steps;
  await engine.run(25);
  assert.strictEqual(engine.steps, s0 + 25);
  const x = await engine.positions();
  assert(x.every(Number.isFinite));
  const sample = await engine.sample();
  assert(Number.isFinite(sample.totalEnergy) && sample.kineticEnergy > 0);
  // a box the cutoff does not fit is refused by name, and the engine goes on
  expectError(() => engine.setBox([1.5, 1.5, 1.5]), 'CutoffTooLargeForBox');
  await engine.run(5);
  assert.strictEqual(engine.steps, s0 + 30);
  // the Monte Carlo barostat on the device: every particle its own molecule
  const n = desc.n_particles;
  const start = Int32Array.from({ length: n + 1 }, (_, i) => i);
  const atoms = Int32Array.from({ length: n }, (_, i) => i);
  native.attachBarostat(engine.handle, { pressure: 1, temperature: 300, frequency: 5, moleculeStart: start,
    moleculeAtoms: atoms, seed: 3 });
  await engine.run(10);
  const st = native.barostatState(engine.handle);
  assert.strictEqual(st.attempts, 2);
  assert.strictEqual(st.stepsSinceAttempt, 0);
  assert(st.draws >= 2 && Number.isFinite(st.lastWork));
  expectError(() => native.attachBarostat(engine.handle, { pressure: 1, temperature: 300, frequency: 5,
    moleculeStart: Int32Array.from([0, 1]), moleculeAtoms: atoms, seed: 3 }), 'InvalidArgument');
  engine.destroy();
  console.log(`node binding (gpu): ok, direct ${e.groups[3]}, reciprocal ${e.groups[4]}, T ${sample.temperature}`);
}

main().catch((err) => { console.error(err); process.exit(1); });
