'use strict';
// Trajectory and restart output of the Node host: the DCD writer (the byte
// layout openmm.app.DCDFile writes, as shared/traj/dcd.js:9-40 records it),
// the AMBER ASCII restart writer (shared/traj/rst7.js:70-101) and whole-
// molecule imaging into the primary cell (shared/traj/imaging.js:39-91).
// The file formats are fixed by their readers (cpptraj, MDAnalysis, VMD,
// this package's own inpcrd parser); the code is this package's own.

const AKMA_PS = 0.04888821;      // one AKMA time unit in ps (DCD's DELTA field)
const ANGSTROM_PER_NM = 10;
const AMBER_VELOCITY_SCALE = 20.455;  // (A/ps) per AMBER velocity unit

class NonFiniteValue extends Error {
  constructor(what, index, detail) {
    super('the simulation is no longer a number: ' + what + ' went non-finite first at component ' + index
      + ' (atom ' + Math.floor(index / 3) + ', ' + 'xyz'[index % 3] + '). ' + detail);
    this.name = 'NonFiniteStateError';
    this.what = what;
    this.index = index;
  }
}

function firstNonFinite(values, count) {
  for (let k = 0; k < count; k += 1) if (!Number.isFinite(values[k])) return k;
  return -1;
}

// Shift each molecule by the lattice vector that puts its geometric centre in
// the primary cell, reduced c, then b, then a (a triclinic cell in reduced
// form has a along x, b in the xy plane).
function imageMolecules(positions, boxVectors, molecules, out) {
  const n = Math.floor(positions.length / 3);
  const wrapped = out || new Float64Array(3 * n);
  for (let k = 0; k < 3 * n; k += 1) wrapped[k] = positions[k];
  if (!boxVectors) return wrapped;
  const ax = boxVectors[0], bx = boxVectors[3], by = boxVectors[4];
  const cx = boxVectors[6], cy = boxVectors[7], cz = boxVectors[8];
  if (!(ax > 0) || !(by > 0) || !(cz > 0)) return wrapped;
  const count = molecules.count;
  const centre = new Float64Array(3 * count);
  const size = new Float64Array(count);
  for (let atom = 0; atom < n; atom += 1) {
    const m = molecules.moleculeOfAtom[atom];
    for (let c = 0; c < 3; c += 1) centre[3 * m + c] += positions[3 * atom + c];
    size[m] += 1;
  }
  const shift = new Float64Array(3 * count);
  for (let m = 0; m < count; m += 1) {
    if (size[m] === 0) continue;
    const x = centre[3 * m] / size[m], y = centre[3 * m + 1] / size[m], z = centre[3 * m + 2] / size[m];
    const nc = Math.floor(z / cz);
    const nb = Math.floor((y - nc * cy) / by);
    const na = Math.floor((x - nc * cx - nb * bx) / ax);
    shift[3 * m] = -(na * ax + nb * bx + nc * cx);
    shift[3 * m + 1] = -(nb * by + nc * cy);
    shift[3 * m + 2] = -nc * cz;
  }
  for (let atom = 0; atom < n; atom += 1) {
    const m = molecules.moleculeOfAtom[atom];
    for (let c = 0; c < 3; c += 1) wrapped[3 * atom + c] = positions[3 * atom + c] + shift[3 * m + c];
  }
  return wrapped;
}

function moleculesOf(system) {
  return { moleculeOfAtom: system.moleculeOfAtom, count: system.moleculeCount };
}

// DCD unit-cell record: [a, cos(gamma), b, cos(beta), cos(alpha), c], lengths in A.
function cellFromVectors(v) {
  const a = [v[0], v[1], v[2]], b = [v[3], v[4], v[5]], c = [v[6], v[7], v[8]];
  const len = (u) => Math.hypot(u[0], u[1], u[2]);
  const dot = (u, w) => u[0] * w[0] + u[1] * w[1] + u[2] * w[2];
  const la = len(a), lb = len(b), lc = len(c);
  const cosine = (u, w, lu, lw) => (lu * lw > 0 ? dot(u, w) / (lu * lw) : 0);
  return [la * ANGSTROM_PER_NM, cosine(a, b, la, lb), lb * ANGSTROM_PER_NM, cosine(a, c, la, lc),
    cosine(b, c, lb, lc), lc * ANGSTROM_PER_NM];
}

function putText(bytes, offset, text, width) {
  for (let k = 0; k < width; k += 1) {
    const code = k < text.length ? text.charCodeAt(k) : 0;
    bytes[offset + k] = code < 128 ? code : 0x3f;
  }
}

class DcdWriter {
  // options: atomCount, timestep (ps), interval (steps per frame), periodic,
  // molecules ({moleculeOfAtom, count}: image whole molecules), firstStep, title.
  constructor(options) {
    this.atomCount = options.atomCount;
    this.timestep = options.timestep;
    this.interval = options.interval;
    this.periodic = options.periodic === true;
    this.molecules = options.molecules || null;
    this.firstStep = options.firstStep || 0;
    this.title = options.title === undefined ? 'Created by the RainMD MI355X host' : options.title;
    this.frameCount = 0;
    this.chunks = [];
    this._imaged = this.molecules ? new Float64Array(3 * this.atomCount) : null;
  }

  appendFrame(positions, boxVectors) {
    const n = this.atomCount;
    if (positions.length < 3 * n) throw new RangeError('frame has ' + positions.length + ' coordinates, expected ' + 3 * n);
    const box = boxVectors === undefined ? null : boxVectors;
    if (this.periodic && box === null) {
      throw new Error('this trajectory declares a periodic box, so every frame needs its box vectors');
    }
    const bad = firstNonFinite(positions, 3 * n);
    if (bad >= 0) {
      throw new NonFiniteValue('a coordinate handed to the DCD writer', bad, 'Frame ' + (this.frameCount + 1)
        + ' was refused; the ' + this.frameCount + ' frame(s) before it are kept.');
    }
    const xyz = this.molecules && box ? imageMolecules(positions, box, this.molecules, this._imaged) : positions;
    const cellBytes = this.periodic ? 56 : 0;
    const axisBytes = 4 * n + 8;
    const frame = Buffer.alloc(cellBytes + 3 * axisBytes);
    let offset = 0;
    if (this.periodic) {
      frame.writeInt32LE(48, 0);
      const cell = cellFromVectors(box);
      for (let k = 0; k < 6; k += 1) frame.writeDoubleLE(cell[k], 4 + 8 * k);
      frame.writeInt32LE(48, 52);
      offset = 56;
    }
    for (let c = 0; c < 3; c += 1) {
      frame.writeInt32LE(4 * n, offset);
      for (let atom = 0; atom < n; atom += 1) frame.writeFloatLE(xyz[3 * atom + c] * ANGSTROM_PER_NM, offset + 4 + 4 * atom);
      frame.writeInt32LE(4 * n, offset + 4 + 4 * n);
      offset += axisBytes;
    }
    this.chunks.push(frame);
    this.frameCount += 1;
    return this.frameCount;
  }

  // The 276-byte header: built last, once the frame count is known.
  header() {
    const h = Buffer.alloc(276);
    h.writeInt32LE(84, 0);
    h.write('CORD', 4, 'latin1');
    h.writeInt32LE(this.frameCount, 8);
    h.writeInt32LE(this.firstStep, 12);
    h.writeInt32LE(this.interval, 16);
    h.writeInt32LE(Math.max(0, this.frameCount - 1) * this.interval, 20);
    h.writeFloatLE(this.timestep / AKMA_PS, 44);
    h.writeInt32LE(this.periodic ? 1 : 0, 48);
    h.writeInt32LE(24, 84);
    h.writeInt32LE(84, 88);
    h.writeInt32LE(164, 92);
    h.writeInt32LE(2, 96);
    putText(h, 100, this.title, 80);
    putText(h, 180, 'Created ' + new Date().toISOString(), 80);
    h.writeInt32LE(164, 260);
    h.writeInt32LE(4, 264);
    h.writeInt32LE(this.atomCount, 268);
    h.writeInt32LE(4, 272);
    return h;
  }

  finish() { return Buffer.concat([this.header()].concat(this.chunks)); }
}

function columns(values, perLine) {
  const lines = [];
  for (let start = 0; start < values.length; start += perLine) {
    let line = '';
    for (let k = start; k < Math.min(start + perLine, values.length); k += 1) {
      if (!Number.isFinite(values[k])) {
        throw new NonFiniteValue('a coordinate handed to the restart writer', k,
          'No restart was written: a restart is read straight back into a simulation.');
      }
      const text = values[k].toFixed(7);
      if (text.length > 12) {
        throw new Error('coordinate ' + values[k] + ' does not fit the 12-column restart format: an atom has left the box');
      }
      line += text.padStart(12);
    }
    lines.push(line);
  }
  return lines;
}

// AMBER ASCII restart: title, "NATOM TIME", coordinates (A) six per line,
// optional velocities (A per AMBER time unit), optional box (A, degrees).
// options: positions (nm, 3N), velocities (nm/ps), box {lengths (nm), angles}, time (ps), title.
function writeRestart(options) {
  const n = Math.floor(options.positions.length / 3);
  const time = options.time === undefined ? 0 : options.time;
  const lines = [options.title === undefined ? 'RainMD restart' : options.title];
  lines.push(String(n).padStart(5) + time.toExponential(7).padStart(15));
  const x = new Array(3 * n);
  for (let k = 0; k < 3 * n; k += 1) x[k] = options.positions[k] * ANGSTROM_PER_NM;
  Array.prototype.push.apply(lines, columns(x, 6));
  if (options.velocities) {
    const v = new Array(3 * n);
    for (let k = 0; k < 3 * n; k += 1) v[k] = options.velocities[k] * ANGSTROM_PER_NM / AMBER_VELOCITY_SCALE;
    Array.prototype.push.apply(lines, columns(v, 6));
  }
  if (options.box) {
    const b = options.box;
    Array.prototype.push.apply(lines, columns([b.lengths[0] * ANGSTROM_PER_NM, b.lengths[1] * ANGSTROM_PER_NM,
      b.lengths[2] * ANGSTROM_PER_NM, b.angles[0], b.angles[1], b.angles[2]], 6));
  }
  return lines.join('\n') + '\n';
}

// Box lengths (nm) and angles (degrees) of a reduced box-vector matrix.
function boxLengthsAndAngles(v) {
  const cell = cellFromVectors(v);
  const deg = (c) => Math.acos(Math.max(-1, Math.min(1, c))) * 180 / Math.PI;
  return { lengths: [cell[0] / ANGSTROM_PER_NM, cell[2] / ANGSTROM_PER_NM, cell[5] / ANGSTROM_PER_NM],
    angles: [deg(cell[4]), deg(cell[3]), deg(cell[1])] };
}

module.exports = { DcdWriter, cellFromVectors, writeRestart, imageMolecules, moleculesOf, boxLengthsAndAngles,
  AKMA_PS, AMBER_VELOCITY_SCALE };
