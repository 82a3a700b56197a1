#!/usr/bin/env node
'use strict';
// The run driver of the Node host (the counterpart of the reference's
// tools/gpu/write-trajectory.mjs:37-127): load a system, minimise, give it
// Maxwell-Boltzmann velocities, run step blocks, report energies, append DCD
// frames and write the final AMBER restart.
//
//   node rainmd_amd/node/run.js --workload ala-944w --steps 5000 --report 500 \
//        [--pressure 1] [--temperature 300] [--dcd out.dcd] [--rst out.rst7]
//   node rainmd_amd/node/run.js --prmtop x.prmtop --inpcrd x.inpcrd --cutoff 1.0 ...
//
// One JSON line per report on stdout; the last line is a summary.

const fs = require('fs');
const rmd = require('./index.js');
const traj = require('./traj.js');

function parseArgs(argv) {
  const a = { steps: 1000, report: 100, temperature: 300, pressure: null, seed: 0, minimise: 100,
    timestep: 0.002, block: 0, cutoff: 1.0, constraints: 'HBonds' };
  for (let k = 2; k < argv.length; k += 1) {
    const key = argv[k].replace(/^--/, '');
    const value = argv[k + 1];
    k += 1;
    if (['steps', 'report', 'temperature', 'pressure', 'seed', 'minimise', 'timestep', 'block', 'cutoff'].indexOf(key) >= 0) {
      a[key] = Number(value);
    } else {
      a[key] = value;
    }
  }
  return a;
}

async function main() {
  const args = parseArgs(process.argv);
  let system;
  if (args.workload) {
    system = rmd.workloadSystem(args.workload);
  } else {
    const topology = rmd.parsePrmtop(args.prmtop);
    const coordinates = args.inpcrd ? rmd.parseInpcrd(args.inpcrd) : null;
    system = rmd.buildSystem(topology, coordinates, { cutoff: args.cutoff, constraints: args.constraints });
  }
  const options = { temperature: args.temperature, seed: args.seed, timestep: args.timestep };
  if (args.pressure !== null && Number.isFinite(args.pressure)) options.pressure = args.pressure;
  const sim = new rmd.GpuSimulation(0, system, options);
  try {
    if (args.minimise > 0) {
      const m = await sim.minimise({ maxIterations: args.minimise, forceTolerance: 1000 });
      process.stdout.write(JSON.stringify({ minimise: m }) + '\n');
    }
    const positions = await sim.readPositions();
    sim.setState(Float64Array.from(positions), rmd.maxwellBoltzmann(system, args.temperature, args.seed), { step: 0 });
    const dcd = args.dcd ? new traj.DcdWriter({ atomCount: system.atomCount, timestep: args.timestep,
      interval: args.report, periodic: true, molecules: traj.moleculesOf(system) }) : null;
    const started = Date.now();
    let step = 0;
    while (step < args.steps) {
      const count = Math.min(args.report, args.steps - step);
      step = await sim.run(count, { blockSize: args.block || undefined });
      const e = await sim.energy();
      if (dcd) {
        const st = await sim.readState();
        dcd.appendFrame(st.positions, st.boxVectors);
      }
      process.stdout.write(JSON.stringify({ step: step, potential: e.potential, kinetic: e.kinetic,
        temperature: e.temperature, volume: e.volume }) + '\n');
    }
    const seconds = (Date.now() - started) / 1000;
    if (dcd) fs.writeFileSync(args.dcd, dcd.finish());
    if (args.rst) {
      const st = await sim.readState();
      fs.writeFileSync(args.rst, traj.writeRestart({ positions: st.positions, velocities: st.velocities,
        box: traj.boxLengthsAndAngles(st.boxVectors), time: step * args.timestep }));
    }
    const diag = await sim.diagnostics();
    process.stdout.write(JSON.stringify({ done: true, steps: step, seconds: seconds,
      nsPerDay: args.timestep * 1e-3 * step * 86400 / seconds, rebuilds: diag.rebuilds,
      barostat: diag.barostat }) + '\n');
  } finally {
    sim.release();
  }
}

main().catch((e) => {
  process.stderr.write((e && e.stack ? e.stack : String(e)) + '\n');
  process.exit(1);
});
