'use strict';
// CPU checks of the Node host: loaders through the addon, the writers, the
// RNG. Prints one JSON object for tests/test_node.py.
const path = require('path');
const root = path.join(__dirname, '..', '..');
const rmd = require(path.join(root, 'rainmd_amd', 'node', 'index.js'));
const traj = require(path.join(root, 'rainmd_amd', 'node', 'traj.js'));
const data = path.join(root, 'rainmd_amd', 'data');
const out = {};

out.abi = rmd.abiVersion();
out.device = rmd.deviceInfo(0);

// test/amber-parse.test.js:33-46 counts
const vac = rmd.buildSystem(rmd.parsePrmtop(path.join(data, 'ala-vac.prmtop')), null,
  { constraints: 'none', rigidWater: false, cutoff: 0 });
out.vac = { atoms: vac.atomCount, residues: vac.residueCount, molecules: vac.moleculeCount,
  bonds: vac.bondCount, angles: vac.angleCount, torsions: vac.torsionCount, exceptions: vac.exceptionCount };
const wat = rmd.buildSystem(rmd.parsePrmtop(path.join(data, 'ala-wat.prmtop')),
  rmd.parseInpcrd(path.join(data, 'ala-wat.inpcrd'), 1190), { constraints: 'HBonds', cutoff: 1.0 });
out.wat = { atoms: wat.atomCount, constraints: wat.constraintCount, exceptions: wat.exceptionCount,
  dof: wat.degreesOfFreedom, waters: wat.waterCount, box: Array.from(wat.boxVectors) };

// refusals with the reference's error codes
function code(fn) { try { fn(); return null; } catch (e) { return { code: e.code, message: e.message }; } }
out.refuseMismatch = code(() => rmd.parseInpcrd(path.join(data, 'ala-wat.inpcrd'), 22));
out.refuseCutoff = code(() => rmd.buildSystem(rmd.parsePrmtop(path.join(data, 'ala-wat.prmtop')),
  rmd.parseInpcrd(path.join(data, 'ala-wat.inpcrd'), 1190), { cutoff: 2.0 }));
if (!out.device.ok) {
  out.createWithoutDevice = code(() => new rmd.GpuSimulation(0, wat, {}));
}

// the protocol's helpers: the restrained selection and the live-box rebuild
out.soluteHeavy = Array.from(rmd.soluteHeavyAtoms(wat));
const grown = Array.from(wat.boxVectors).map((v, k) => (k % 4 === 0 ? v * 1.02 : v));
const reboxed = rmd.reboxSystem(wat, grown);
out.rebox = { box: Array.from(reboxed.boxVectors), atoms: reboxed.atomCount, constraints: reboxed.constraintCount,
  original: Array.from(wat.boxVectors) };
out.refuseRebox = code(() => rmd.reboxSystem(wat, [1.9, 0, 0, 0, 3, 0, 0, 0, 3]));

// DCD layout (test/physics.test.js:253-287)
const dcd = new traj.DcdWriter({ atomCount: 3, timestep: 0.002, interval: 5, periodic: true });
const xyz = Float64Array.from([0, 0, 0, 0.1, 0.2, 0.3, -0.4, 0.5, -0.6]);
const box = Float64Array.from([2.5, 0, 0, 0.3, 2.6, 0, 0.4, 0.5, 2.7]);
dcd.appendFrame(xyz, box);
dcd.appendFrame(xyz, box);
const bytes = dcd.finish();
out.dcd = { length: bytes.length, marker: bytes.readInt32LE(0), cord: bytes.toString('latin1', 4, 8),
  frames: bytes.readInt32LE(8), interval: bytes.readInt32LE(16), last: bytes.readInt32LE(20),
  delta: bytes.readFloatLE(44), cellFlag: bytes.readInt32LE(48), version: bytes.readInt32LE(84),
  atoms: bytes.readInt32LE(268), cell: traj.cellFromVectors(box),
  firstCell: [0, 1, 2, 3, 4, 5].map((k) => bytes.readDoubleLE(276 + 4 + 8 * k)),
  firstX: [0, 1, 2].map((k) => bytes.readFloatLE(276 + 56 + 4 + 4 * k)) };
out.dcdNoBox = code(() => new traj.DcdWriter({ atomCount: 1, timestep: 0.002, interval: 1, periodic: true })
  .appendFrame(Float64Array.from([0, 0, 0])));
out.dcdNaN = code(() => new traj.DcdWriter({ atomCount: 1, timestep: 0.002, interval: 1, periodic: false })
  .appendFrame(Float64Array.from([0, NaN, 0])));

// restart round trip through the addon's own inpcrd reader (physics.test.js:300-323)
const crd = rmd.parseInpcrd(path.join(data, 'ala-wat.inpcrd'));
const vel = new Float64Array(crd.positions.length);
for (let k = 0; k < vel.length; k += 1) vel[k] = Math.sin(k * 0.7) * 0.5;
const text = traj.writeRestart({ positions: crd.positions, velocities: vel,
  box: { lengths: Array.from(crd.boxLengths), angles: Array.from(crd.boxAngles) }, time: 1.5 });
const again = rmd.parseInpcrd(Buffer.from(text), crd.atomCount);
let dx = 0, dv = 0, db = 0;
for (let k = 0; k < crd.positions.length; k += 1) dx = Math.max(dx, Math.abs(again.positions[k] - crd.positions[k]));
for (let k = 0; k < vel.length; k += 1) dv = Math.max(dv, Math.abs(again.velocities[k] - vel[k]));
for (let k = 0; k < 3; k += 1) db = Math.max(db, Math.abs(again.boxLengths[k] - crd.boxLengths[k]));
out.restart = { atoms: again.atomCount, time: again.time, dx: dx, dv: dv, db: db, head: text.split('\n').slice(0, 2) };
out.restartOverflow = code(() => traj.writeRestart({ positions: Float64Array.from([1e6, 0, 0]) }));
out.restartNaN = code(() => traj.writeRestart({ positions: Float64Array.from([1, NaN, 0]) }));

// imaging: every molecule's centre lands in the primary cell, by lattice shifts
const mol = { moleculeOfAtom: [0, 0, 1], count: 2 };
const raw = Float64Array.from([2.6, -0.1, 5.5, 2.7, 0.1, 5.5, -0.2, 0.3, 0.3]);
out.imaged = Array.from(traj.imageMolecules(raw, box, mol));

// RNG and velocities
out.uniform = [[7, 3, 0x5eed, 1], [123456, 99, 2, 3], [0, 0, 0, 0]].map((a) => rmd.uniformAt(a[0], a[1], a[2], a[3]));
out.velocities = Array.from(rmd.maxwellBoltzmann(wat, 300, 202).slice(0, 30));
process.stdout.write(JSON.stringify(out) + '\n');
