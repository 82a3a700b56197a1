"""GPU parity: the HIP path through the C ABI against the FP64 oracle.

Bars (from the contract of this path): cell assignments and neighbour pair
sets bit-exact against the FP32 emulation; every energy term within 1e-6
relative of FP64 at the same f32 positions; per-atom forces within 1e-4
relative RMS.
"""
import numpy as np
import pytest

import oracle as O
import rainmd_amd as R
from helpers import ALPHA_24, fixture_system, rel, rel_rms

pytestmark = pytest.mark.gpu

ENERGY_BAR = 1e-6
FORCE_BAR = 1e-4


def _term_pairs(gpu, ref):
    return {
        "harmonic_bond": (gpu["harmonic_bond"], ref["bond"]),
        "harmonic_angle": (gpu["harmonic_angle"], ref["angle"]),
        "periodic_torsion": (gpu["periodic_torsion"], ref["torsion"]),
        "lennard_jones": (gpu["lennard_jones"], ref["lennard_jones"]),
        "coulomb_direct": (gpu["coulomb_direct"], ref["coulomb_direct_total"]),
        "coulomb_reciprocal": (gpu["coulomb_reciprocal"], ref["reciprocal"]),
        "coulomb_self": (gpu["coulomb_self"], ref["self"] + ref["net"]),
    }


def _check_terms(label, gpu, ref, bar=ENERGY_BAR):
    failures = []
    for name, (a, b) in _term_pairs(gpu, ref).items():
        r = rel(a, b)
        print(f"{label} {name:20s} gpu {a:18.8f} oracle {b:18.8f} rel {r:.2e}", flush=True)
        if r > bar:
            failures.append((name, r))
    total_gpu = gpu["potential"] - gpu["dispersion"]
    total_ref = ref["total"]
    print(f"{label} {'total':20s} gpu {total_gpu:18.8f} oracle {total_ref:18.8f} rel {rel(total_gpu, total_ref):.2e}")
    if rel(total_gpu, total_ref) > bar:
        failures.append(("total", rel(total_gpu, total_ref)))
    assert not failures, failures


@pytest.fixture(scope="module")
def ala_wat():
    system, coords = fixture_system("ala-wat", dict(constraints="HBonds", rigid_water=True, cutoff=1.0))
    sim = R.GpuSimulation(system, alpha=ALPHA_24, grid=[24, 24, 24], seed=7)
    yield system, sim
    sim.release()


def test_device_is_mi355x():
    info = R.device_info()
    print(info)
    assert info["ok"] and info["device_count"] >= 1
    assert info["arch"].startswith("gfx950")


def test_fft_matches_numpy(ala_wat):
    _, sim = ala_wat
    rng = np.random.default_rng(3)
    grid = sim.diagnostics()["pme_grid"]
    x = rng.standard_normal(grid).astype(np.float32)
    ours = sim.debug_fft_forward(x)
    # This build's forward transform carries exp(+2 pi i m k / K): the
    # conjugate of numpy's for real input.
    ref = np.conj(np.fft.rfftn(x.astype(np.float64)))
    err = np.abs(ours - ref).max() / np.abs(ref).max()
    print("fft rel err", err)
    assert err < 20 * np.sqrt(3 * np.log2(grid[0])) * 2.0 ** -24


@pytest.mark.parametrize("pipeline", ["planes", "lines"])
@pytest.mark.parametrize("grid", [(30, 30, 30), (28, 21, 25), (27, 32, 20), (35, 14, 21)])
def test_fft_mixed_radix(grid, pipeline):
    system, _ = fixture_system("ala-wat", dict(constraints="HBonds", rigid_water=True, cutoff=1.0))
    sim = R.GpuSimulation(system, alpha=ALPHA_24, grid=list(grid))
    sim.set_pme_pipeline(pipeline)
    assert sim.diagnostics()["pme_planes"] == (pipeline == "planes")
    rng = np.random.default_rng(sum(grid))
    x = rng.standard_normal(grid).astype(np.float32)
    ours = sim.debug_fft_forward(x)
    ref = np.conj(np.fft.rfftn(x.astype(np.float64)))
    err = np.abs(ours - ref).max() / np.abs(ref).max()
    print(grid, "fft rel err", err)
    sim.release()
    assert err < 20 * np.sqrt(3 * np.log2(max(grid))) * 2.0 ** -24


def test_cells_and_pairs_bit_exact(ala_wat):
    system, sim = ala_wat
    sim.evaluate()
    diag = sim.diagnostics()
    pos = sim.read_positions()
    cell, start, atoms = sim.read_cells()
    expected = O.cells_f32(pos, system.box, diag["cell_counts"])
    assert np.array_equal(cell, expected)
    # the cell list is a partition, each cell's atoms sorted ascending
    for c in range(len(start) - 1):
        seg = atoms[start[c]:start[c + 1]]
        assert np.all(np.diff(seg) > 0)
        assert np.all(cell[seg] == c)
    assert sorted(atoms.tolist()) == list(range(system.atom_count))
    counts, lists = sim.read_neighbours()
    rl = system.cutoff + diag["skin"]
    ref_counts, ref_lists = O.pairs_f32(pos, system.box, rl, system.exclusion_start, system.exclusion_partners)
    assert np.array_equal(counts, ref_counts)
    for i in range(system.atom_count):
        assert np.array_equal(np.sort(lists[i, :counts[i]]), ref_lists[i, :ref_counts[i]]), i


@pytest.mark.parametrize("pipeline", ["planes", "lines"])
def test_energies_and_forces_ala_wat(ala_wat, pipeline):
    system, sim = ala_wat
    sim.set_pme_pipeline(pipeline)
    gpu, forces = sim.evaluate()
    sim.set_pme_pipeline("auto")
    pos = sim.read_positions().astype(np.float64)
    ref, ref_forces = O.Oracle(system).potential(pos, alpha=ALPHA_24, grid=[24, 24, 24], f32_membership=True)
    _check_terms("ala-wat", gpu, ref)
    frel = rel_rms(forces, ref_forces)
    print("forces rel rms", frel)
    assert frel < FORCE_BAR


@pytest.mark.parametrize("workload", ["ala-389w", "ala-944w"])
def test_energies_and_forces_workloads(workload):
    system = R.workload_system(workload)
    sim = R.GpuSimulation(system, temperature=300.0, seed=11)
    pme = sim.diagnostics()
    gpu, forces = sim.evaluate()
    pos = sim.read_positions().astype(np.float64)
    ref, ref_forces = O.Oracle(system).potential(pos, alpha=pme["pme_alpha"], grid=pme["pme_grid"],
                                                 f32_membership=True)
    # per-piece breakdown: direct pairs and exception pairs on their own
    oracle = O.Oracle(system)
    for mask, keys in [(R._lib.TERM_DIRECT, ("lj_direct", "coulomb_direct")),
                       (R._lib.TERM_EXCEPTIONS, ("lj_exception", "coulomb_exception"))]:
        g, _ = sim.evaluate(terms=mask, forces=False)
        r, _ = oracle.potential(pos, alpha=pme["pme_alpha"], grid=pme["pme_grid"], terms=mask, f32_membership=True,
                                forces=False)
        print(workload, keys, "lj rel %.2e coul rel %.2e" % (rel(g["lennard_jones"], r[keys[0]]),
                                                             rel(g["coulomb_direct"], r[keys[1]])))
    _check_terms(workload, gpu, ref)
    frel = rel_rms(forces, ref_forces)
    print(workload, "forces rel rms", frel)
    sim.release()
    assert frel < FORCE_BAR


@pytest.mark.parametrize("concurrent, graphs", [(True, True), (False, True), (True, False)])
def test_steps_are_deterministic(ala_wat, concurrent, graphs):
    # the same 20 steps twice, and against the default mode (overlapped force
    # branches in captured graphs: for this small system the direct space,
    # bonded terms and PME forward planes in one launch; concurrent=False runs
    # the separate kernels): bit-identical states, forces and energies
    system, _ = ala_wat
    runs = []
    for mode in [(True, True), (concurrent, graphs)]:
        sim = R.GpuSimulation(system, alpha=ALPHA_24, grid=[24, 24, 24], seed=5, temperature=298.0)
        sim.set_concurrent_forces(mode[0])
        sim.set_graphs(mode[1])
        sim.run(20)
        st = sim.read_state()
        e = sim.energy()
        runs.append((st["positions"].copy(), st["velocities"].copy(), st["forces"].copy(),
                     np.array([e["potential"], e["kinetic"], e["coulomb_reciprocal"]])))
        sim.release()
    for a, b in zip(runs[0], runs[1]):
        assert np.array_equal(a, b)


def test_molecular_scaling_derivative_matches_central_differences():
    # The barostat scales molecules about their centroids with their geometry
    # kept; dscale_molecular is the energy's derivative along that move
    # (scaleVolume performs exactly that move). Checked by central differences
    # on the terms without a cutoff - PME reciprocal, bonded, exceptions -
    # whose energy is smooth in the scale; the direct space's cutoff makes its
    # finite differences jump as pairs cross rc. The atomic (r dU/dr) derivative
    # of the same terms is printed beside it: it is not the barostat's.
    from rainmd_amd import _lib as L
    smooth = L.TERM_RECIPROCAL | L.TERM_BONDED | L.TERM_EXCEPTIONS
    system = R.workload_system("ala-944w")
    sim = R.GpuSimulation(system, temperature=300.0, seed=3)
    sim.minimise(max_iterations=100, force_tolerance=1000.0)
    e0, _ = sim.evaluate(terms=smooth, forces=False)
    delta = 1e-3
    sim.scale_volume(np.exp(delta))
    e_up, _ = sim.evaluate(terms=smooth, forces=False)
    sim.scale_volume(np.exp(-2 * delta))
    e_down, _ = sim.evaluate(terms=smooth, forces=False)
    energy = lambda e: e["potential"] - e["dispersion"]  # noqa: E731 (self energy is constant)
    finite = (energy(e_up) - energy(e_down)) / (2 * delta)
    molecular = e0["dscale_molecular"]
    print("central difference", finite, "molecular", molecular, "atomic reciprocal", e0["dscale_reciprocal"])
    assert abs(molecular - finite) < 1e-3 * abs(finite)
    sim.release()
