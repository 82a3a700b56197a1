"""The replica workloads at their configured shapes:

- the uneven batch: 256 replicas of ACE-ALA-NME + TIP3P with cube edges
  U(2.2, 5.0) nm (1,066-12,472 atoms, 1.42 M in all), NPT with a move every
  25 steps per replica, 200 steps: every replica finite, constraints held,
  moves attempted and accepted, the smallest and largest replicas' energies
  and forces against the FP64 oracle;
- the delayed-acceptance batch: the same 256 replicas plus the 2854-atom
  system, the shared network trained on the first 2,000 exact moves, then
  screening: both acceptance stages decide moves and a share of the moves
  skips the exact evaluation.

The reference runs one system per submit (webgpu/simulation.js:456-527);
its barostat is webgpu/simulation.js:995-1070.
"""
import numpy as np
import pytest

import oracle as O
import rainmd_amd as R
from helpers import rel, rel_rms

pytestmark = pytest.mark.gpu

COUNT = 256
SKIN = 0.15


@pytest.fixture(scope="module")
def replicas():
    """The 256 replica systems, their seeds and minimised starting states
    (positions and Maxwell-Boltzmann velocities at 300 K)."""
    work = R.replica_workload(COUNT)
    seeds = [s for _, _, s in work]
    systems = R.replica_systems(COUNT)
    batch = R.Batch(systems, seeds=seeds, temperature=300.0, pressure=1.0, barostat_interval=25, skin=SKIN)
    starts = []
    for r, system in enumerate(systems):
        sim = batch.replica(r)
        sim.minimise(max_iterations=60, force_tolerance=1000.0)
        starts.append((sim.read_positions().astype(np.float64), R.maxwell_boltzmann(system, 300.0, seeds[r])))
    batch.release()
    return systems, seeds, starts


def _constraint_error(system, positions):
    x = np.asarray(positions, dtype=np.float64).reshape(-1, 3)
    if len(system.constraint_a) == 0:
        return 0.0
    d = np.linalg.norm(x[system.constraint_a] - x[system.constraint_b], axis=1)
    return float(np.max(np.abs(d - system.constraint_distance) / system.constraint_distance))


def test_uneven_batch_of_256_replicas(replicas):
    systems, seeds, starts = replicas
    atoms = [s.atom_count for s in systems]
    assert len(systems) == COUNT and min(atoms) < 1100 and max(atoms) > 12000
    batch = R.Batch(systems, seeds=seeds, temperature=300.0, pressure=1.0, barostat_interval=25, skin=SKIN)
    for r, (x, v) in enumerate(starts):
        batch.replica(r).set_state(x, v, step=0)
    batch.run(200, 25)
    attempts = accepted = 0
    worst_constraint = 0.0
    for r, system in enumerate(systems):
        sim = batch.replica(r)
        d = sim.diagnostics()
        e = sim.energy()
        st = sim.read_state()
        assert d["steps"] == 200, r
        assert np.isfinite(e["potential"]) and np.isfinite(e["kinetic"]), r
        assert np.all(np.isfinite(st["positions"])) and np.all(np.isfinite(st["velocities"])), r
        worst_constraint = max(worst_constraint, _constraint_error(system, st["positions"]))
        attempts += d["barostat_attempts"]
        accepted += d["barostat_accepted"]
        assert d["barostat_attempts"] >= 7, r
    print("moves", accepted, "/", attempts, "worst constraint error", worst_constraint)
    assert worst_constraint < 2e-5
    assert 0 < accepted < attempts
    # the smallest and largest replicas against the FP64 oracle
    for r in (int(np.argmin(atoms)), int(np.argmax(atoms))):
        sim = batch.replica(r)
        gpu, forces = sim.evaluate()
        st = sim.read_state()
        d = sim.diagnostics()
        o = O.Oracle(systems[r], box=np.asarray(st["box_vectors"]))
        ref, ref_forces = o.potential(st["positions"].astype(np.float64), alpha=d["pme_alpha"],
                                      grid=list(d["pme_grid"]), f32_membership=True)
        terms = {"lennard_jones": ref["lennard_jones"], "coulomb_direct": ref["coulomb_direct_total"],
                 "coulomb_reciprocal": ref["reciprocal"], "bonded": ref["bonded"]}
        for key, want in terms.items():
            print("replica", r, atoms[r], "atoms", key, gpu[key], want, rel(gpu[key], want))
            # The reciprocal term of the smallest box (2.2 nm, 20^3 grid, ~1e2
            # kJ/mol) sits at the FP32 transform's floor: 1.2e-7 to 1.03e-6
            # across the states this test and the bench have ended in, where
            # the total is within 1e-8. Its bar is 2e-6; every other term's 1e-6.
            bar = 2e-6 if key == "coulomb_reciprocal" and atoms[r] < 1100 else 1e-6
            assert rel(gpu[key], want) < bar, (r, key)
        assert rel(gpu["potential"] - gpu["dispersion"], ref["total"]) < 1e-6, r
        assert rel_rms(forces, ref_forces) < 1e-4, r
    batch.release()


def test_delayed_acceptance_on_256_replicas_and_the_2854_atom_system(replicas):
    systems, seeds, starts = replicas
    big = R.workload_system("ala-944w")
    sim = R.GpuSimulation(big, temperature=300.0, seed=505, skin=SKIN)
    sim.minimise(max_iterations=200, force_tolerance=1000.0)
    big_start = (sim.read_positions().astype(np.float64), R.maxwell_boltzmann(big, 300.0, 505))
    sim.release()
    members = list(systems) + [big]
    batch = R.Batch(members, seeds=list(seeds) + [505], temperature=300.0, pressure=1.0, barostat_interval=25,
                    skin=SKIN)
    for r, (x, v) in enumerate(list(starts) + [big_start]):
        batch.replica(r).set_state(x, v, step=0)
    batch.run(100, 25)
    batch.configure_screen(training_moves=2000, training_steps=4000, learning_rate=1e-3, seed=505)
    collection = 0
    while batch.screen_stats()["phase"] == 0:
        batch.run(25, 25)
        collection += 25
        assert collection <= 400
    trained = batch.screen_stats()
    print("trained after", collection, "steps:", trained)
    assert trained["samples"] >= 2 * 2000
    batch.run(500, 25)
    st = batch.screen_stats()
    print("screening", st)
    # stage one turned moves down without an exact evaluation; stage two
    # decided the survivors on their exact energy change
    assert st["moves"] >= 18 * len(members)
    assert st["screened_out"] > 0
    assert st["exact_evaluations"] > 0 and st["accepted"] > 0
    assert st["exact_evaluations"] + st["screened_out"] <= st["moves"]
    skipped = 1.0 - st["exact_evaluations"] / st["moves"]
    print("skipped fraction", skipped)
    assert skipped > 0.1
    assert st["residual_rms"] < st["estimate_rms"]
    e = batch.replica(len(members) - 1).energy()
    assert np.isfinite(e["potential"]) and batch.replica(len(members) - 1).diagnostics()["steps"] > 0
    batch.release()
