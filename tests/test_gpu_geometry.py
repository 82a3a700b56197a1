"""Box shapes and Ewald parameters off the default paths.

- A triclinic (reduced) box runs the general minimum image of the direct
  space and the list search (rectangular boxes take a form with the zero
  off-diagonal terms left out): cells and pair sets bit-exact against the
  FP32 emulation, energies and forces against the FP64 oracle, and inside a
  replica batch beside rectangular replicas every replica bit for bit its own
  single run.
- An Ewald alpha beyond the rational force fit ((alpha rc)^2 > 12) takes the
  erfc force: energies and forces against the FP64 oracle.
- A small system's force-only evaluations fold the PME x lines into the
  plane gathers (a circular convolution along x), in a cube and in a
  triclinic box (where the convolution kernel is complex): reciprocal and
  total forces within FP32 rounding of the x lines' and within the force bar
  of the FP64 oracle; evaluations with energies are the x lines' own.
"""
import numpy as np
import pytest

import oracle as O
import rainmd_amd as R
from helpers import rel, rel_rms

pytestmark = pytest.mark.gpu

ENERGY_BAR = 1e-6
FORCE_BAR = 1e-4


def _triclinic(system):
    edge = system.box[0]
    box = np.array([[edge, 0.0, 0.0], [0.25 * edge, edge, 0.0], [-0.2 * edge, 0.15 * edge, edge]])
    return system.rebox(box)


def _check(label, system, sim, alpha, grid):
    gpu, forces = sim.evaluate()
    pos = sim.read_positions().astype(np.float64)
    oracle = O.Oracle(system)
    oracle.set_box(np.asarray(sim.read_state()["box_vectors"]).reshape(9))
    ref, ref_forces = oracle.potential(pos, alpha=alpha, grid=list(grid), f32_membership=True)
    pairs = {"lennard_jones": ref["lennard_jones"], "coulomb_direct": ref["coulomb_direct_total"],
             "coulomb_reciprocal": ref["reciprocal"], "harmonic_bond": ref["bond"],
             "harmonic_angle": ref["angle"], "periodic_torsion": ref["torsion"]}
    failures = []
    for key, value in pairs.items():
        r = rel(gpu[key], value)
        print(f"{label} {key:20s} gpu {gpu[key]:18.8f} oracle {value:18.8f} rel {r:.2e}")
        if r > ENERGY_BAR:
            failures.append((key, r))
    frel = rel_rms(forces, ref_forces)
    print(label, "forces rel rms", frel)
    assert not failures, failures
    assert frel < FORCE_BAR


def test_triclinic_box_cells_pairs_energies_and_forces():
    system = _triclinic(R.workload_system("ala-944w"))
    sim = R.GpuSimulation(system, temperature=300.0, seed=5)
    sim.minimise(max_iterations=100, force_tolerance=1000.0)
    diag = sim.diagnostics()
    box = np.asarray(sim.read_state()["box_vectors"]).reshape(9)
    assert box[3] != 0.0 and box[6] != 0.0
    sim.evaluate()
    pos = sim.read_positions()
    cell, _, _ = sim.read_cells()
    assert np.array_equal(cell, O.cells_f32(pos, box, diag["cell_counts"]))
    counts, lists = sim.read_neighbours()
    ref_counts, ref_lists = O.pairs_f32(pos, box, system.cutoff + diag["skin"], system.exclusion_start,
                                        system.exclusion_partners)
    assert np.array_equal(counts, ref_counts)
    for i in range(system.atom_count):
        assert np.array_equal(np.sort(lists[i, :counts[i]]), ref_lists[i, :ref_counts[i]]), i
    _check("triclinic", system, sim, diag["pme_alpha"], diag["pme_grid"])
    sim.release()


def test_batch_with_a_triclinic_replica_equals_single_runs():
    cube_small = R.workload_system("ala-389w")
    systems = [cube_small, _triclinic(R.workload_system("ala-389w")), R.workload_system("ala-944w")]
    seeds = [21, 22, 23]
    starts = []
    for sys_, seed in zip(systems, seeds):
        sim = R.GpuSimulation(sys_, temperature=300.0, seed=seed)
        sim.minimise(max_iterations=60, force_tolerance=1000.0)
        starts.append((sim.read_positions().astype(np.float64), R.maxwell_boltzmann(sys_, 300.0, seed)))
        sim.release()
    options = dict(temperature=300.0, pressure=1.0, barostat_interval=10, lanes_per_atom=16, cell_reach=3)
    batch = R.Batch(systems, seeds=seeds, **options)
    for r, (x, v) in enumerate(starts):
        batch.replica(r).set_state(x, v, step=0)
    batch.run(60)
    for r, (sys_, seed) in enumerate(zip(systems, seeds)):
        sim = R.GpuSimulation(sys_, seed=seed, **options)
        sim.set_pme_pipeline("planes")  # the batch members' x lines
        sim.set_state(*starts[r], step=0)
        sim.run(60)
        sb, ss = batch.replica(r).read_state(), sim.read_state()
        print("replica", r, "potential", batch.replica(r).energy()["potential"], sim.energy()["potential"])
        assert np.array_equal(sb["positions"], ss["positions"]), r
        assert np.array_equal(sb["velocities"], ss["velocities"]), r
        assert np.array_equal(sb["box_vectors"], ss["box_vectors"]), r
        sim.release()
    batch.release()


def test_alpha_beyond_the_rational_force_fit_takes_erfc():
    system = R.workload_system("ala-389w")
    alpha, grid = 4.2, (32, 32, 32)
    assert (alpha * system.cutoff) ** 2 > 12.0
    sim = R.GpuSimulation(system, alpha=alpha, grid=list(grid), temperature=300.0, seed=9)
    sim.minimise(max_iterations=60, force_tolerance=1000.0)
    _check("alpha 4.2", system, sim, alpha, grid)
    sim.release()


@pytest.mark.parametrize("shape", ["cube", "triclinic"])
def test_folded_x_lines(shape):
    system = R.workload_system("ala-944w")
    if shape == "triclinic":
        system = _triclinic(system)
    sim = R.GpuSimulation(system, temperature=300.0, seed=5)
    diag = sim.diagnostics()
    assert diag["pme_x_folded"] == 1 and diag["pme_grid"][0] <= 32
    recip = R._lib.TERM_RECIPROCAL
    _, folded = sim.evaluate(terms=recip, energies=False)
    _, folded_all = sim.evaluate(energies=False)
    e_folded, _ = sim.evaluate(terms=recip)
    sim.set_pme_pipeline("planes-scan")
    assert sim.diagnostics()["pme_x_folded"] == 0
    _, lines = sim.evaluate(terms=recip, energies=False)
    _, lines_all = sim.evaluate(energies=False)
    e_lines, _ = sim.evaluate(terms=recip)
    sim.set_pme_pipeline("auto")
    assert sim.diagnostics()["pme_x_folded"] == 1
    assert e_folded["coulomb_reciprocal"] == e_lines["coulomb_reciprocal"]
    pos = sim.read_positions().astype(np.float64)
    oracle = O.Oracle(system)
    oracle.set_box(np.asarray(sim.read_state()["box_vectors"]).reshape(9))
    _, ref = oracle.potential(pos, alpha=diag["pme_alpha"], grid=list(diag["pme_grid"]), terms=recip,
                              f32_membership=True)
    d_recip, d_all = rel_rms(folded, lines), rel_rms(folded_all, lines_all)
    o_folded, o_lines = rel_rms(folded, ref), rel_rms(lines, ref)
    print(shape, "folded vs x lines: reciprocal", d_recip, "all terms", d_all,
          "| reciprocal vs FP64: folded", o_folded, "x lines", o_lines)
    sim.release()
    assert d_recip < 2e-6 and d_all < 1e-6
    assert o_folded < FORCE_BAR and o_folded < 1.5 * o_lines + 1e-7
