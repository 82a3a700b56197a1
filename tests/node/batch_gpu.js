'use strict';
// GPU checks of the Node host's replica batch: two workload systems in one
// batch, each replica minimised and given velocities through its view, then
// stepped together. Prints one JSON object for tests/test_node.py to compare
// with the same batch run through the Python host.
const path = require('path');
const root = path.join(__dirname, '..', '..');
const rmd = require(path.join(root, 'rainmd_amd', 'node', 'index.js'));

async function main() {
  const names = ['ala-389w', 'ala-944w'];
  const seeds = [31, 32];
  const systems = names.map((n) => rmd.workloadSystem(n));
  const batch = new rmd.Batch(0, systems, { temperature: 300, pressure: 1.0, barostatInterval: 10, lanesPerAtom: 16 },
    seeds);
  const out = { info: batch.info(), replicas: [] };
  for (let r = 0; r < systems.length; r += 1) {
    const sim = batch.replica(r);
    await sim.minimise({ maxIterations: 60, forceTolerance: 1000 });
    sim.setState(Float64Array.from(await sim.readPositions()), rmd.maxwellBoltzmann(systems[r], 300, seeds[r]),
      { step: 0 });
  }
  out.step = await batch.run(60);
  for (let r = 0; r < systems.length; r += 1) {
    const sim = batch.replica(r);
    const e = await sim.energy();
    const diag = await sim.diagnostics();
    out.replicas.push({ positions: Array.from(await sim.readPositions()), potential: e.potential, step: sim.step,
      attempts: diag.barostat.attempts });
  }
  out.infoAfter = batch.info();
  try {
    batch.replica(0).release();
    out.replicaRelease = null;
  } catch (e) {
    out.replicaRelease = e.message;
  }
  batch.release();
  try {
    await batch.replica(1).energy();
    out.afterRelease = null;
  } catch (e) {
    out.afterRelease = e.message;
  }
  process.stdout.write(JSON.stringify(out) + '\n');
}

main().catch((e) => {
  process.stderr.write((e && e.stack ? e.stack : String(e)) + '\n');
  process.exit(1);
});
