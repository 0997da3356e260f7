'use strict';
/**
 * HipEngine: upstream's ProtocolEngine (src/engine/protocol.ts:32-62) over the
 * N-API binding of libmdcore.so, for the TypeScript host under Node.  Same
 * members and meanings as LoopEngine (src/gpu/engine.ts:16-104): the device
 * owns the state, `run` resolves when the steps are done, `sample` reports
 * the potential (device, per force group), the kinetic energy (1/2 sum m v^2,
 * the LangevinMiddle convention) and the temperature over 3 N_moving - N_c
 * - 3 degrees of freedom (src/reference/kinetic.ts:86-113).
 *
 * `desc` is md_system_desc's fields (typed arrays for the arrays), as the
 * host's packSystem produces from a parsed OpenMM System.
 */
const native = require('./mdcore.node');

const BOLTZMANN = 0.008314462618153242; // kJ/mol/K (src/engine/units.ts)

class HipEngine {
  constructor(desc, state, options = {}) {
    const o = Object.assign({
      time_step: 0.002, temperature: 300, friction: 1, constraint_tolerance: 1e-6, seed: 0,
      integrator: 0, verlet_skin: -1, device: 0, max_batch: 0,
    }, options);
    this.options = o;
    this.n = desc.n_particles;
    this.masses = desc.masses;
    this.handle = native.createSystem(desc, o);
    let moving = 0;
    for (let i = 0; i < this.n; i++) if (desc.masses[i] > 0) moving++;
    this.degreesOfFreedom = 3 * moving - (desc.n_constraints || 0) - ((desc.cmm_frequency || 0) > 0 ? 3 : 0);
    native.setPositions(this.handle, state.positions);
    native.setVelocities(this.handle, state.velocities || new Float64Array(3 * this.n));
    if (state.stepCount) native.setStep(this.handle, state.stepCount);
  }

  get particleCount() { return this.n; }
  get steps() { return Number(native.step(this.handle)); }
  positions() { return Promise.resolve(native.positions(this.handle, this.n)); }
  velocities() { return Promise.resolve(native.velocities(this.handle, this.n)); }

  boxVectors() {
    const b = native.box(this.handle);
    return [[b[0], b[1], b[2]], [b[3], b[4], b[5]], [b[6], b[7], b[8]]];
  }

  boxLengths() {
    const b = native.box(this.handle);
    return [b[0], b[4], b[8]];
  }

  setPositions(x) { native.setPositions(this.handle, x); }
  setVelocities(v) { native.setVelocities(this.handle, v); }

  /** A diagonal (three lengths) or the whole cell (nine numbers). */
  setBox(box) {
    const b = box.length === 3 ? Float64Array.from([box[0], 0, 0, 0, box[1], 0, 0, 0, box[2]])
      : Float64Array.from(box);
    native.setBox(this.handle, b);
  }

  run(steps) { return native.run(this.handle, steps); }

  evaluate() { return native.evaluate(this.handle, this.n); }

  async sample() {
    const e = native.evaluate(this.handle, this.n);
    const v = native.velocities(this.handle, this.n);
    let ke = 0;
    for (let i = 0; i < this.n; i++) {
      const m = this.masses[i];
      ke += 0.5 * m * (v[3 * i] * v[3 * i] + v[3 * i + 1] * v[3 * i + 1] + v[3 * i + 2] * v[3 * i + 2]);
    }
    return {
      potentialEnergy: e.total,
      kineticEnergy: ke,
      totalEnergy: e.total + ke,
      temperature: this.degreesOfFreedom > 0 ? 2 * ke / (this.degreesOfFreedom * BOLTZMANN) : 0,
    };
  }

  destroy() {
    if (this.handle) native.destroySystem(this.handle);
    this.handle = null;
    return 0;
  }
}

module.exports = { HipEngine, native };
