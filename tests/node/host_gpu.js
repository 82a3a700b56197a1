'use strict';
// GPU checks of the Node host: a GpuSimulation through the addon. Prints one
// JSON object with the evaluated state for tests/test_node.py to compare with
// the FP64 oracle.
const path = require('path');
const root = path.join(__dirname, '..', '..');
const rmd = require(path.join(root, 'rainmd_amd', 'node', 'index.js'));

async function main() {
  const out = {};
  const system = rmd.workloadSystem('ala-389w');
  const sim = new rmd.GpuSimulation(0, system, { temperature: 300, seed: 11, pressure: 1.0 });
  out.pme = sim.pme;
  await sim.initialise();
  out.positions = Array.from(await sim.readPositions());
  out.energy = await sim.energy();
  out.energy.boxVectors = Array.from(out.energy.boxVectors);
  // dynamics: minimise, velocities, a few blocks, a stop signal, a volume move
  const m = await sim.minimise({ maxIterations: 150, forceTolerance: 1000 });
  out.minimise = m;
  sim.setState(Float64Array.from(await sim.readPositions()), rmd.maxwellBoltzmann(system, 300, 11), { step: 0 });
  out.step = await sim.run(100, { blockSize: 25 });
  const after = await sim.energy();
  out.after = { potential: after.potential, temperature: after.temperature, volume: after.volume, step: after.step };
  out.move = await sim.attemptVolumeChange();
  const signal = new rmd.StopSignal();
  signal.stop('test');
  try {
    await sim.run(25, { stopSignal: signal });
    out.stopped = null;
  } catch (e) {
    out.stopped = e.name;
  }
  // a synchronous call while a run is in flight is refused, not raced
  const pending = sim.run(200, { blockSize: 25 });
  out.busy = {};
  for (const [name, call] of [['energy', () => sim.energy()], ['readState', () => sim.readState()],
    ['readPositions', () => sim.readPositions()], ['diagnostics', () => sim.diagnostics()],
    ['setState', async () => sim.setState(new Float64Array(3 * sim.atomCount), null, { step: 0 })]]) {
    try {
      await call();
      out.busy[name] = null;
    } catch (e) {
      out.busy[name] = e.message;
    }
  }
  out.afterBusy = await pending;
  const diag = await sim.diagnostics();
  out.diagnostics = { rebuilds: diag.rebuilds, barostat: diag.barostat, pmePlanes: diag.pmePlanes };
  sim.release();
  let released = null;
  try { await sim.energy(); } catch (e) { released = e.message; }
  out.released = released;
  process.stdout.write(JSON.stringify(out) + '\n');
}
main().catch((e) => { process.stderr.write(String(e && e.stack ? e.stack : e) + '\n'); process.exit(1); });
