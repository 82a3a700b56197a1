'use strict';
// The Node.js host of the MI355X MD step: the GpuSimulation surface of the
// reference (webgpu/simulation.js:82-1341) and its loaders
// (shared/amber/prmtop.js, inpcrd.js; shared/topology/system.js:40-236), over
// the C ABI through the N-API addon (addon.cc). Written for Node 12: no `??`,
// no `?.`, no private fields, CommonJS.
//
// There is no CPU fallback: requiring this module without the built addon, or
// creating a simulation without a HIP device, throws.

const fs = require('fs');
const path = require('path');

let addon;
try {
  addon = require(path.join(__dirname, 'rainmd.node'));
} catch (e) {
  throw new Error('the rainmd N-API addon is not built (rainmd_amd/node/rainmd.node): ' + e.message);
}

const STATUS = {
  INVALID: -1, PARSE: -2, SYSTEM: -3, DEVICE: -4, NONFINITE: -5,
  NEIGHBOUR_OVERFLOW: -6, STOPPED: -7, BUDGET: -8, CONSTRAINT: -9,
};

// ---- the reference's error classes, same names and fields ----

class NeighbourOverflowError extends Error {
  // webgpu/kernels/neighbours.js:172-183
  constructor(required, capacity, message) {
    super(message || ('the neighbour list needed ' + required + ' slots per atom and the capacity is ' + capacity));
    this.name = 'NeighbourOverflowError';
    this.required = required;
    this.capacity = capacity;
  }
}

class NonFiniteStateError extends Error {
  // shared/finite.js:63-95
  constructor(options) {
    super(options.message || ('the simulation is no longer a number: ' + options.what + ' went non-finite'));
    this.name = 'NonFiniteStateError';
    this.what = options.what;
    this.step = options.step;
    this.index = options.index;
  }
}

class SimulationStopped extends Error {
  // webgpu/runtime/yield.js:44-50
  constructor(reason, step) {
    super('the run was stopped: ' + reason);
    this.name = 'SimulationStopped';
    this.reason = reason;
    this.step = step;
  }
}

class GpuBudgetError extends Error {
  // webgpu/runtime/allocator.js:141-154
  constructor(message) {
    super(message);
    this.name = 'GpuBudgetError';
  }
}

class ConstraintFailure extends Error {
  // cpu/integrate/langevin.js:164
  constructor(message) {
    super(message);
    this.name = 'ConstraintFailure';
  }
}

// A cooperative stop flag, read between step blocks (webgpu/runtime/yield.js:40-51).
class StopSignal {
  constructor() { this._stopped = false; this._reason = null; this._targets = new Set(); }
  stop(reason) {
    this._stopped = true;
    this._reason = reason === undefined ? 'stopped by the user' : reason;
    for (const handle of this._targets) addon.requestStop(handle);
  }
  get stopped() { return this._stopped; }
  get reason() { return this._reason; }
}

function translate(error, context) {
  if (!error || typeof error.code !== 'number') return error;
  const message = error.message;
  switch (error.code) {
    case STATUS.NONFINITE: return new NonFiniteStateError({ what: 'the state', step: context.step, message: message });
    case STATUS.NEIGHBOUR_OVERFLOW: return new NeighbourOverflowError(context.required, context.capacity, message);
    case STATUS.STOPPED: return new SimulationStopped(context.reason || 'stopped by the user', context.step);
    case STATUS.BUDGET: return new GpuBudgetError(message);
    case STATUS.CONSTRAINT: return new ConstraintFailure(message);
    default: {
      const e = new Error(message);
      e.code = error.code;
      return e;
    }
  }
}

function call(fn, context) {
  try {
    return fn();
  } catch (e) {
    throw translate(e, context || {});
  }
}

// ---- loaders and the system builder ----

function readText(source) {
  if (Buffer.isBuffer(source)) return source;
  if (typeof source === 'string' && source.indexOf('\n') < 0 && fs.existsSync(source)) return fs.readFileSync(source);
  return source;
}

// parsePrmtop(text) (shared/amber/prmtop.js:111-385): a file path, its text or a Buffer.
function parsePrmtop(source) {
  return { handle: call(() => addon.parsePrmtop(readText(source))) };
}

// parseInpcrd(text, atomCount) (shared/amber/inpcrd.js:34-132).
function parseInpcrd(source, expectedAtomCount) {
  const handle = call(() => addon.parseInpcrd(readText(source), expectedAtomCount || 0));
  const view = addon.coordinatesView(handle);
  return Object.assign({ handle: handle }, view);
}

function boxVectorsFromLengthsAndAngles(lengths, angles) {
  return call(() => addon.boxVectors(lengths, angles));
}

function wrapSystem(handle) {
  const view = addon.systemView(handle);
  return Object.assign({ handle: handle, topology: { atomCount: view.atomCount, moleculeOfAtom: view.moleculeOfAtom } }, view);
}

// buildSystem(topology, coordinates, {constraints, rigidWater, cutoff}) (shared/topology/system.js:40-236).
function buildSystem(topology, coordinates, options) {
  const o = options || {};
  return wrapSystem(call(() => addon.buildSystem(topology.handle, coordinates ? coordinates.handle : null, o)));
}

// The seeded solvated-peptide builder of the synthetic workloads.
function buildSolvated(options) {
  const data = path.join(__dirname, '..', 'data');
  const prmtop = fs.readFileSync(path.join(data, 'ala-wat.prmtop'));
  const inpcrd = fs.readFileSync(path.join(data, 'ala-vac.inpcrd'));
  return wrapSystem(call(() => addon.buildSolvated(prmtop, inpcrd, options || {})));
}

function pmeParameters(options) {
  return call(() => addon.pmeParameters(options.boxVectors, options));
}

// ---- the counter-hash RNG and Maxwell-Boltzmann velocities ----

function mix(v) {
  v = (v ^ (v >>> 16)) >>> 0;
  v = Math.imul(v, 0x7feb352d) >>> 0;
  v = (v ^ (v >>> 15)) >>> 0;
  v = Math.imul(v, 0x846ca68b) >>> 0;
  return (v ^ (v >>> 16)) >>> 0;
}

// uniformAt (shared/random.js:59-64): ((h >> 9) + 0.5) * 2^-23, in (0, 1).
function uniformAt(seed, a, b, c) {
  let st = mix((seed ^ Math.imul(a, 0x9e3779b1)) >>> 0);
  st = mix((st ^ Math.imul(b, 0x85ebca6b)) >>> 0);
  st = mix((st ^ Math.imul(c, 0xc2b2ae35)) >>> 0);
  return ((st >>> 9) + 0.5) * (1.0 / 8388608.0);
}

// setVelocitiesToTemperature (cpu/integrate/langevin.js:211-230): per atom
// gaussianTriple(seed, atom, 0x5eed) scaled by sqrt(kT/m); massless atoms get
// zero.
function maxwellBoltzmann(system, temperature, seed) {
  const n = system.atomCount;
  const kT = 0.008314462618 * temperature;
  const v = new Float64Array(3 * n);
  for (let atom = 0; atom < n; atom += 1) {
    const m = system.masses[atom];
    if (!(m > 0)) continue;
    const u = [0, 1, 2, 3].map((k) => uniformAt(seed >>> 0, atom, 0x5eed, k));
    const r1 = Math.sqrt(-2 * Math.log(u[0])), r2 = Math.sqrt(-2 * Math.log(u[2]));
    const sigma = Math.sqrt(kT / m);
    v[3 * atom] = sigma * r1 * Math.cos(2 * Math.PI * u[1]);
    v[3 * atom + 1] = sigma * r1 * Math.sin(2 * Math.PI * u[1]);
    v[3 * atom + 2] = sigma * r2 * Math.cos(2 * Math.PI * u[3]);
  }
  return v;
}

// The synthetic workloads (the same builder arguments as the Python host).
const WORKLOADS = {
  'ala-389w': { waterCount: 389, boxEdge: 2.2816, seed: 101, includeSolute: true, cutoff: 0.9 },
  'ala-944w': { waterCount: 944, boxEdge: 3.0539, seed: 202, includeSolute: true, cutoff: 0.9 },
  'water-33400': { waterCount: 33400, boxEdge: 10.0, seed: 303, includeSolute: false, cutoff: 1.0 },
};

function workloadSystem(name) {
  if (!WORKLOADS[name]) throw new Error('unknown workload ' + name);
  return buildSolvated(WORKLOADS[name]);
}

// ---- GpuSimulation (webgpu/simulation.js:82-1341) ----

class GpuSimulation {
  // `device` is a HIP device ordinal (the reference takes a GPUDevice here).
  constructor(device, system, options) {
    const o = Object.assign({}, options || {});
    if (typeof device === 'number') o.device = device;
    GpuSimulation._attach(this, call(() => addon.create(system.handle, o)), system, o);
  }

  // The fields of a simulation over a context handle (its own, or a replica
  // of a batch).
  static _attach(sim, handle, system, o) {
    sim.system = system;
    sim.atomCount = system.atomCount;
    sim.timestep = o.timestep === undefined ? 0.002 : o.timestep;
    sim.pressure = o.pressure === undefined ? null : o.pressure;
    sim._handle = handle;
    sim._stopSignals = [];
    sim.step = 0;
    const diag = addon.diagnostics(handle);
    // public fields callers read (web/app.js:1141-1143)
    sim.pme = { grid: Array.from(diag.pmeGrid), alpha: diag.pmeAlpha };
    sim.plan = { counts: Array.from(diag.cellCounts), skin: diag.skin, maxNeighbours: diag.neighbourCapacity };
    sim.skin = diag.skin;
    sim.clusters = { layout: { clusterCount: diag.clusterUnits + diag.waterUnits } };
  }

  _context(extra) {
    const diag = this._handle ? addon.diagnostics(this._handle) : null;
    return Object.assign({ step: this.step, required: diag ? diag.neighbourRequired : undefined,
      capacity: diag ? diag.neighbourCapacity : undefined }, extra || {});
  }

  async _async(fn, extra) {
    try {
      return await fn();
    } catch (e) {
      throw translate(e, this._context(extra));
    }
  }

  // setState(positions, velocities, {step, suppliedNoise}) (simulation.js:357-370)
  setState(positions, velocities, options) {
    const o = options || {};
    call(() => addon.setState(this._handle, positions, velocities || null, o), this._context());
    this.step = o.step || 0;
  }

  // setNoise(noise) (simulation.js:371)
  setNoise(noise) { call(() => addon.setNoise(this._handle, noise), this._context()); }

  // initialise() (simulation.js:412-422)
  async initialise() { await this._async(() => addon.initialise(this._handle)); }

  // run(count, {blockSize, stopSignal}) -> step (simulation.js:456-527)
  async run(count, options) {
    const o = options || {};
    const signal = o.stopSignal || null;
    if (signal) {
      if (signal.stopped) throw new SimulationStopped(signal.reason, this.step);
      signal._targets.add(this._handle);
    }
    try {
      this.step = await this._async(() => addon.run(this._handle, count, o.blockSize || 0),
        { reason: signal ? signal.reason : undefined });
    } finally {
      if (signal) signal._targets.delete(this._handle);
    }
    return this.step;
  }

  // scaleVolume(s) -> potential (simulation.js:801-811)
  async scaleVolume(lengthScale) { return this._async(() => addon.scaleVolume(this._handle, lengthScale)); }

  // minimise(options) (simulation.js:880-979)
  async minimise(options) {
    const result = await this._async(() => addon.minimise(this._handle, options || {}));
    result.history = [];
    return result;
  }

  // attemptVolumeChange() (simulation.js:995-1070): null without a barostat.
  async attemptVolumeChange() { return this._async(() => addon.attemptVolumeChange(this._handle)); }

  // energy() (simulation.js:1073-1230)
  async energy() { return call(() => addon.energy(this._handle), this._context()); }

  // readPositions() (simulation.js:1239-1242): Float32Array of 3N
  async readPositions() { return call(() => addon.readPositions(this._handle, this.atomCount), this._context()); }

  // readState() (simulation.js:1244-1261)
  async readState() { return call(() => addon.readState(this._handle, this.atomCount), this._context()); }

  // diagnostics() (simulation.js:1264-1334)
  async diagnostics() {
    const d = call(() => addon.diagnostics(this._handle), this._context());
    if (this.pressure === null) {
      d.barostat = null;
    } else {
      d.barostat.pressure = this.pressure;
      d.barostat.acceptanceRate = d.barostat.attempts === 0 ? null : d.barostat.accepted / d.barostat.attempts;
    }
    return d;
  }

  // release() (simulation.js:1336-1340)
  release() {
    if (this._handle) addon.destroy(this._handle);
    this._handle = null;
  }
}

// The system in the live box a stage left behind: the page rebuilds its
// system with the current box vectors before every stage (web/app.js:354-364).
function reboxSystem(system, boxVectors) {
  return wrapSystem(call(() => addon.reboxSystem(system.handle, Array.from(boxVectors))));
}

// soluteHeavyAtoms(system, selections) (shared/topology/restraints.js:46-49):
// non-water, non-ion residues, atomic number other than 1.
function soluteHeavyAtoms(system) {
  return call(() => addon.soluteHeavyAtoms(system.handle));
}

// A replica batch (rmd_batch_*): many systems, each with its own box, PME
// grid, seed and barostat, stepped together by batched launches; every
// replica is bit for bit its own single simulation. No counterpart in the
// reference. replica(r) is the GpuSimulation surface of one replica (released
// with the batch); run() steps them all.
class Batch {
  constructor(device, systems, options, seeds) {
    const o = Object.assign({}, options || {});
    if (typeof device === 'number') o.device = device;
    this.systems = systems;
    this._options = o;
    this._handle = call(() => addon.batchCreate(systems.map((sys) => sys.handle), o, seeds || null));
    this._views = [];
    this.step = 0;
  }

  get replicas() { return this.systems.length; }

  replica(r) {
    if (!this._views[r]) {
      const view = Object.create(GpuSimulation.prototype);
      GpuSimulation._attach(view, call(() => addon.batchReplica(this._handle, r)), this.systems[r], this._options);
      view._batch = this;  // a view keeps its batch alive
      this._views[r] = view;
    }
    return this._views[r];
  }

  async _async(fn) {
    try {
      return await fn();
    } catch (e) {
      throw translate(e, { step: this.step });
    }
  }

  async initialise() { await this._async(() => addon.batchInitialise(this._handle)); }

  // run(count, {blockSize}) -> step
  async run(count, options) {
    const o = options || {};
    this.step = await this._async(() => addon.batchRun(this._handle, count, o.blockSize || 0));
    for (const view of this._views) if (view) view.step = this.step;
    return this.step;
  }

  info() { return call(() => addon.batchInfo(this._handle)); }
  setGraphs(enabled) { call(() => addon.batchSetGraphs(this._handle, !!enabled)); }
  setConcurrentForces(enabled) { call(() => addon.batchSetConcurrentForces(this._handle, !!enabled)); }
  requestStop() { addon.batchRequestStop(this._handle); }
  // the shared learned barostat screen: {enabled, trainingMoves, trainingSteps, learningRate, seed}
  configureScreen(options) { call(() => addon.batchScreenConfigure(this._handle, options || {})); }
  screenStats() { return call(() => addon.batchScreenStats(this._handle)); }

  release() {
    if (this._handle) addon.batchDestroy(this._handle);
  }
}

module.exports = {
  GpuSimulation, Batch, StopSignal,
  NeighbourOverflowError, NonFiniteStateError, SimulationStopped, GpuBudgetError, ConstraintFailure,
  parsePrmtop, parseInpcrd, boxVectorsFromLengthsAndAngles, buildSystem, buildSolvated, pmeParameters,
  reboxSystem, soluteHeavyAtoms,
  uniformAt, maxwellBoltzmann, WORKLOADS, workloadSystem,
  deviceInfo: (device) => addon.deviceInfo(device || 0),
  abiVersion: () => addon.abiVersion(),
  STATUS,
};
