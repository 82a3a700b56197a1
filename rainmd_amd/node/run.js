#!/usr/bin/env node
'use strict';
// The run driver of the Node host: the page's protocol (web/app.js:1041-1124,
// docs/STATUS.md:286-316) and the trajectory writer
// (tools/gpu/write-trajectory.mjs:37-127) in one loop.
//
//   1. minimise the loaded structure;
//   2. equilibrate (optional): NPT (or NVT without --pressure) with the solute
//      heavy atoms restrained towards the stage's starting structure at
//      --restraint kJ/mol/nm^2;
//   3. production: the same ensemble, unrestrained.
//
// Every stage is a fresh GpuSimulation over the system rebuilt in the box the
// previous stage left (a barostat changes positions, velocities AND the box,
// web/app.js:1375-1393) and starts from that stage's whole state. Velocities
// are drawn from Maxwell-Boltzmann at the first dynamics stage only.
//
//   node rainmd_amd/node/run.js --workload ala-944w --steps 5000 --report 500 \
//        [--equilibrate 5000] [--restraint 500] [--pressure 1] [--temperature 300] \
//        [--dcd out.dcd] [--rst out.rst7]
//   node rainmd_amd/node/run.js --prmtop x.prmtop --inpcrd x.inpcrd --cutoff 1.0 ...
//
// One JSON line per report on stdout, tagged with its stage; the last line is
// a summary. With --dcd, the equilibration frames go to <name>.equil.dcd and
// the production frames to the name given.

const fs = require('fs');
const rmd = require('./index.js');
const traj = require('./traj.js');

const NUMERIC = ['steps', 'report', 'temperature', 'pressure', 'seed', 'minimise', 'timestep', 'block', 'cutoff',
  'equilibrate', 'restraint', 'barostatInterval'];

function parseArgs(argv) {
  const a = { steps: 1000, report: 100, temperature: 300, pressure: null, seed: 0, minimise: 100,
    timestep: 0.002, block: 0, cutoff: 1.0, constraints: 'HBonds', equilibrate: 0, restraint: 500,
    barostatInterval: 25 };
  for (let k = 2; k < argv.length; k += 1) {
    const key = argv[k].replace(/^--/, '');
    const value = argv[k + 1];
    k += 1;
    a[key] = NUMERIC.indexOf(key) >= 0 ? Number(value) : value;
  }
  return a;
}

function emit(o) { process.stdout.write(JSON.stringify(o) + '\n'); }

function loadSystem(args) {
  if (args.workload) return rmd.workloadSystem(args.workload);
  const topology = rmd.parsePrmtop(args.prmtop);
  const coordinates = args.inpcrd ? rmd.parseInpcrd(args.inpcrd) : null;
  return rmd.buildSystem(topology, coordinates, { cutoff: args.cutoff, constraints: args.constraints });
}

function dynamicsOptions(args) {
  const o = { temperature: args.temperature, seed: args.seed, timestep: args.timestep,
    barostatInterval: args.barostatInterval };
  if (args.pressure !== null && Number.isFinite(args.pressure)) o.pressure = args.pressure;
  return o;
}

// One stage of dynamics (web/app.js:1103-1270): a fresh simulation of
// `system` from `state`, `steps` steps in report chunks; returns the stage's
// whole final state.
async function runStage(name, system, state, args, stage) {
  const options = Object.assign(dynamicsOptions(args), {
    restraintForce: stage.restraintForce || 0,
    restraintAtoms: stage.restraintAtoms || [],
    // The restraint reference is the structure this stage starts from.
    restraintReference: state.positions,
  });
  const sim = new rmd.GpuSimulation(0, system, options);
  const dcd = stage.dcd ? new traj.DcdWriter({ atomCount: system.atomCount, timestep: args.timestep,
    interval: args.report, periodic: true, molecules: traj.moleculesOf(system) }) : null;
  let step = 0;
  const started = Date.now();
  try {
    sim.setState(state.positions, state.velocities, { step: 0 });
    while (step < stage.steps) {
      const count = Math.min(args.report, stage.steps - step);
      step = await sim.run(count, { blockSize: args.block || undefined });
      const e = await sim.energy();
      const st = await sim.readState();
      // The live box, not the system's starting box.
      if (dcd) dcd.appendFrame(st.positions, st.boxVectors);
      emit({ stage: name, step: step, potential: e.potential, kinetic: e.kinetic, temperature: e.temperature,
        volume: e.volume, positionRestraint: e.positionRestraint });
    }
    const seconds = (Date.now() - started) / 1000;
    const st = await sim.readState();
    const diag = await sim.diagnostics();
    if (dcd) fs.writeFileSync(stage.dcd, dcd.finish());
    return {
      positions: Float64Array.from(st.positions),
      velocities: Float64Array.from(st.velocities),
      boxVectors: Float64Array.from(st.boxVectors),
      steps: step, seconds: seconds, rebuilds: diag.rebuilds, barostat: diag.barostat,
    };
  } finally {
    sim.release();
  }
}

async function main() {
  const args = parseArgs(process.argv);
  let system = loadSystem(args);
  let state = { positions: Float64Array.from(system.positions), velocities: null,
    boxVectors: Float64Array.from(system.boxVectors) };

  if (args.minimise > 0) {
    const sim = new rmd.GpuSimulation(0, system, dynamicsOptions(args));
    try {
      sim.setState(state.positions, null, { step: 0 });
      const m = await sim.minimise({ maxIterations: args.minimise, forceTolerance: 1000 });
      emit({ stage: 'minimise', minimise: m });
      state.positions = Float64Array.from(await sim.readPositions());
    } finally {
      sim.release();
    }
  }
  state.velocities = rmd.maxwellBoltzmann(system, args.temperature, args.seed);

  const summary = { done: true, stages: [] };
  if (args.equilibrate > 0) {
    const restraintAtoms = rmd.soluteHeavyAtoms(system);
    emit({ stage: 'equilibration', restrainedAtoms: restraintAtoms.length, restraintForce: args.restraint });
    const out = await runStage('equilibration', system, state, args, {
      steps: args.equilibrate, restraintForce: Math.max(0, args.restraint), restraintAtoms: restraintAtoms,
      dcd: args.dcd ? args.dcd.replace(/\.dcd$/, '') + '.equil.dcd' : null,
    });
    summary.stages.push({ stage: 'equilibration', steps: out.steps, seconds: out.seconds, barostat: out.barostat });
    state = out;
    // The next stage runs in the box this one left.
    system = rmd.reboxSystem(system, state.boxVectors);
  }

  const out = await runStage('production', system, state, args, { steps: args.steps, dcd: args.dcd || null });
  summary.stages.push({ stage: 'production', steps: out.steps, seconds: out.seconds, barostat: out.barostat });
  if (args.rst) {
    fs.writeFileSync(args.rst, traj.writeRestart({ positions: out.positions, velocities: out.velocities,
      box: traj.boxLengthsAndAngles(out.boxVectors), time: out.steps * args.timestep }));
  }
  summary.steps = out.steps;
  summary.seconds = out.seconds;
  summary.nsPerDay = args.timestep * 1e-3 * out.steps * 86400 / out.seconds;
  summary.rebuilds = out.rebuilds;
  summary.barostat = out.barostat;
  emit(summary);
}

main().catch((e) => {
  process.stderr.write((e && e.stack ? e.stack : String(e)) + '\n');
  process.exit(1);
});
