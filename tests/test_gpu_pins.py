"""GPU parity of the pieces around the force field, each against the FP64
oracle or the reference's own vectors:

- the thermostat's device noise (uniforms bit for bit, normals in FP32) and
  one BAOAB step drawing it (tools/gpu/check-integrator.mjs:44-127);
- kinetic energy and temperature (cpu/integrate/langevin.js:192-208);
- the cell plan and skin (shared/topology/cells.js:77-181);
- neighbour pair sets on both list-search paths, at 26 cells per axis, on an
  uneven-batch replica and on a sheared box the stencil does not cover whole
  (tools/gpu/check-nonbonded.mjs:50-339);
- the displacement test: the device's worst displacement, a tiny skin that
  must rebuild every few steps and leave the energies exact, a run at the
  derived skin that never misses a pair, and a non-finite state mid-run
  (tools/gpu/check-displacement.mjs:62-175, webgpu/simulation.js:544-623);
- the steepest-descent minimiser's path (webgpu/simulation.js:880-979).
"""
import json
import os

import numpy as np
import pytest

import oracle as O
import rainmd_amd as R
from helpers import ALPHA_24, fixture_system, rel, rel_rms

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
BOLTZMANN = 0.008314462618
SEED, STEP = 20261106, 41  # the reference's generator arm (check-integrator.mjs:45)


def _ala_wat(**options):
    system, coords = fixture_system("ala-wat", dict(constraints="HBonds", rigid_water=True, cutoff=1.0))
    sim = R.GpuSimulation(system, alpha=ALPHA_24, grid=[24, 24, 24], **options)
    return system, coords, sim


def _on_manifold(system, x, v):
    o = O.Oracle(system)
    x, _ = o.shake(x, x, 1e-12)
    v, _ = o.rattle(x, v, 1e-12)
    return x.astype(np.float32).astype(np.float64), v.astype(np.float32).astype(np.float64)


# ---- noise -------------------------------------------------------------------

def test_device_noise_is_the_reference_stream():
    system, _, sim = _ala_wat(seed=SEED)
    n = system.atom_count
    u, g = sim.debug_noise(STEP)
    ref_u = np.array([[O.uniform(SEED, a, STEP, k) for k in range(4)] for a in range(n)])
    # ((h >> 9) + 0.5) 2^-23 is exact in FP32: the device's uniforms are the
    # host's bit for bit
    assert np.array_equal(u.astype(np.float64), ref_u)
    with open(os.path.join(HERE, "golden", "random", "random_vectors.json")) as f:
        vectors = json.load(f)
    theirs = np.array([t[3:] for t in vectors["triples"] if t[0] == SEED and t[2] == STEP])
    assert len(theirs) == 64
    ref_g = np.array([O.gaussian(SEED, a, STEP) for a in range(n)])
    assert np.allclose(ref_g[:64], theirs, rtol=1e-14, atol=1e-15)
    # FP32 Box-Muller: log, sqrt and sincospi to a few ulps of values up to ~5.8
    err = np.abs(g - ref_g)
    print("device normals: max |err| %.3e, rms %.3e" % (err.max(), np.sqrt(np.mean(err ** 2))))
    assert err.max() < 4e-6
    assert abs(g.mean()) < 0.05 and abs(g.std() - 1) < 0.05
    sim.release()


def test_one_step_with_device_noise_matches_the_host_integrator():
    # tools/gpu/check-integrator.mjs:68-127 with the noise drawn on both sides
    # from (seed, atom, step) instead of supplied
    system, coords, sim = _ala_wat(tolerance=1e-10, temperature=298.0, friction=1.0, seed=SEED)
    x0, v0 = _on_manifold(system, coords.positions.astype(np.float32).astype(np.float64),
                          R.maxwell_boltzmann(system, 298.0, 99).astype(np.float32).astype(np.float64))
    sim.set_state(x0, v0, step=STEP, supplied_noise=False)
    _, f0 = sim.evaluate()
    sim.step_once()
    st = sim.read_state()
    x1, v1 = st["positions"].astype(np.float64), st["velocities"].astype(np.float64)
    o = O.Oracle(system)
    xo, vo, rc = o.langevin_step(x0, v0, f0.reshape(-1), timestep=0.002, temperature=298.0, friction=1.0,
                                 tolerance=1e-12, max_iterations=5000, seed=SEED, step=STEP)
    assert rc == 0
    mid = o.kinetic(vo)
    _, f1 = sim.evaluate()
    vo, rc = o.langevin_finish(xo, vo, f1.reshape(-1), tolerance=1e-12, max_iterations=5000)
    dx = np.abs(x1 - xo).max()
    dv = np.abs(v1 - vo).max()
    print("one step, device noise: max |dx| %.3e nm, max |dv| %.3e nm/ps" % (dx, dv))
    assert dx < 1e-5
    assert dv < 1e-3 * np.abs(vo).max()
    # the same step without noise differs by far more than the bars: the
    # noise is not a no-op the bars could hide
    xn, vn, _ = o.langevin_step(x0, v0, f0.reshape(-1), timestep=0.002, temperature=298.0, friction=1.0,
                                tolerance=1e-12, max_iterations=5000, noise=np.zeros(3 * system.atom_count))
    assert np.abs(xn - xo).max() > 20 * 1e-5
    # the kinetic energy sampled before the closing half kick (the
    # temperature's sample) is the host's at that point
    e = sim.energy()
    assert e["kinetic_sampled_in_step"] == 1
    print("mid-step kinetic: gpu %.6f oracle %.6f" % (e["kinetic_for_temperature"], mid))
    assert rel(e["kinetic_for_temperature"], mid) < 1e-4
    sim.release()


# ---- kinetic energy and temperature --------------------------------------------

def test_kinetic_energy_and_temperature_match_the_oracle():
    system, coords, sim = _ala_wat(temperature=298.0, seed=3)
    o = O.Oracle(system)
    dof = system.degrees_of_freedom
    # test/physics.test.js:160-175: unit velocities give (1/2) sum(m) per axis
    sim.set_state(coords.positions, np.ones(3 * system.atom_count), step=0)
    e = sim.energy()
    want = o.kinetic(np.ones(3 * system.atom_count))
    assert rel(e["kinetic"], want) < 1e-6
    assert e["kinetic_sampled_in_step"] == 0
    assert rel(e["temperature"], 2 * want / (dof * BOLTZMANN)) < 1e-6
    # thermal velocities, and after a run: the step-end kinetic energy of the
    # velocities the device holds
    sim.set_state(coords.positions, o.velocities(298.0, 3), step=0)
    sim.run(50)
    e = sim.energy()
    v = sim.read_state()["velocities"].astype(np.float64)
    k = o.kinetic(v)
    print("kinetic: gpu %.8f oracle %.8f rel %.2e" % (e["kinetic"], k, rel(e["kinetic"], k)))
    assert rel(e["kinetic"], k) < 1e-6
    assert rel(e["temperature"], 2 * e["kinetic_for_temperature"] / (dof * BOLTZMANN)) < 1e-12
    sim.release()


# ---- cell plan -------------------------------------------------------------------

@pytest.mark.parametrize("name", ["ala-389w", "ala-944w"])
def test_cell_plan_and_skin_follow_the_reference_rules(name):
    system = R.workload_system(name)
    sim = R.GpuSimulation(system, temperature=300.0)
    d = sim.diagnostics()
    o = O.Oracle(system)
    skin = o.skin_for(0.002, 300.0)
    plan = o.plan_cells(skin)
    print(name, d["skin"], skin, list(d["cell_counts"]), plan)
    assert abs(d["skin"] - skin) < 1e-12
    assert list(d["cell_counts"]) == plan["counts"]
    assert d["neighbour_capacity"] == plan["max_neighbours"]
    # a box under four list cutoffs: the reference's plan, all-pairs search
    assert d["cell_reach"] == 1
    sim.release()


def test_cell_plan_beyond_the_reference_reach():
    # The 10 nm water box: at reach 1 the reference's 8^3 cells; the default
    # plan takes cells of (rc + skin)/3 with a 7-cell stencil per axis
    system = R.workload_system("water-33400")
    o = O.Oracle(system)
    plan = o.plan_cells(0.15)
    one = R.GpuSimulation(system, skin=0.15, grid=[112, 112, 112], cell_reach=1)
    d1 = one.diagnostics()
    assert list(d1["cell_counts"]) == plan["counts"] == [8, 8, 8] and d1["cell_reach"] == 1
    assert d1["neighbour_capacity"] == plan["max_neighbours"]
    one.release()
    wide = R.GpuSimulation(system, skin=0.15, grid=[112, 112, 112])
    d3 = wide.diagnostics()
    assert d3["cell_reach"] == 3
    assert list(d3["cell_counts"]) == [int(np.floor(10.0 * 3 / 1.15))] * 3 == [26, 26, 26]
    assert d3["search_by_cell"] == 1
    wide.release()


# ---- pair sets on both search paths ------------------------------------------------

def _pairs_equal(sim, system, box, rl):
    sim.evaluate()
    pos = sim.read_positions()
    counts, lists = sim.read_neighbours()
    ref_counts, ref_lists = O.pairs_f32_binned(pos, box, rl, system.exclusion_start, system.exclusion_partners)
    assert np.array_equal(counts, ref_counts)
    bad = [i for i in range(system.atom_count)
           if not np.array_equal(np.sort(lists[i, :counts[i]]), ref_lists[i, :ref_counts[i]])]
    assert not bad, bad[:10]
    cell, _, _ = sim.read_cells()
    assert np.array_equal(cell, O.cells_f32(pos, box, sim.diagnostics()["cell_counts"]))
    return int(counts.sum())


def _truncated_octahedron(d):
    # the reduced truncated octahedron of vector length d, filled at 1 g/cm^3
    # from a cube of the same volume (the cube's atoms rewrapped by the box)
    box = np.array([[d, 0.0, 0.0], [d / 3, 2 * np.sqrt(2) * d / 3, 0.0], [-d / 3, np.sqrt(2) * d / 3, np.sqrt(6) * d / 3]])
    volume = box[0, 0] * box[1, 1] * box[2, 2]
    edge = volume ** (1 / 3)
    waters = int(round((602.214076 * volume - 144.17) / 18.016))
    system = R.System.solvated(water_count=waters, box_edge=edge, seed=77, include_solute=True, cutoff=0.9)
    return system.rebox(box)


@pytest.mark.parametrize("case", ["replica", "octahedron"])
def test_pair_sets_bit_exact_on_both_search_paths(case):
    if case == "replica":
        # the largest-box replica of the uneven batch at its 0.15 nm skin
        work = R.replica_workload(256)
        edge, waters, seed = max(work)
        system = R.System.solvated(water_count=waters, box_edge=edge, seed=seed, include_solute=True, cutoff=0.9)
        options = dict(skin=0.15)
    else:
        # a truncated octahedron (d = 4.4 nm): its face-to-face widths, 3.59 nm
        # on every axis, give 8 cells of (rc + skin)/3 per axis, where its
        # diagonal would give 10 x 9 x 8 cells too thin for the 7-cell stencil
        system = _truncated_octahedron(4.4)
        options = {}
    totals = {}
    for mode in (1, 2):
        sim = R.GpuSimulation(system, temperature=300.0, seed=5, search_mode=mode, **options)
        d = sim.diagnostics()
        box = np.asarray(sim.read_state()["box_vectors"]).reshape(9)
        rl = system.cutoff + d["skin"]
        print(case, "mode", mode, "cells", list(d["cell_counts"]), "reach", d["cell_reach"], "by cell",
              d["search_by_cell"])
        assert d["search_by_cell"] == (mode == 2)
        assert d["cell_reach"] > 1 and min(d["cell_counts"]) > 2 * d["cell_reach"] + 1
        if case == "octahedron":
            assert list(d["cell_counts"]) == [8, 8, 8]
        totals[mode] = _pairs_equal(sim, system, box, rl)
        sim.release()
    assert totals[1] == totals[2]


def test_pair_sets_bit_exact_at_26_cells_per_axis():
    # the 100,200-atom water box, where the plan's occupancy selects the
    # per-cell search
    system = R.workload_system("water-33400")
    sim = R.GpuSimulation(system, skin=0.15, grid=[112, 112, 112])
    d = sim.diagnostics()
    assert d["search_by_cell"] == 1 and list(d["cell_counts"]) == [26, 26, 26]
    total = _pairs_equal(sim, system, system.box, system.cutoff + 0.15)
    print("100,200 atoms: %d list entries" % total)
    sim.release()


# ---- displacement, stale lists and non-finite states ------------------------------------

def test_device_displacement_matches_the_host():
    # check-displacement.mjs arm 1: one step from a fresh list
    system, coords, sim = _ala_wat(temperature=298.0, friction=10.0, seed=4242)
    x0, v0 = _on_manifold(system, coords.positions, R.maxwell_boltzmann(system, 298.0, 17))
    sim.set_state(x0, v0, step=0)
    sim.initialise()
    start = sim.read_positions().astype(np.float64).reshape(-1, 3)
    sim.run(1, 1)
    after = sim.read_positions().astype(np.float64).reshape(-1, 3)
    d = sim.diagnostics()
    host = np.linalg.norm(after - start, axis=1).max()
    print("worst displacement: device %.8e host %.8e" % (d["worst_displacement"], host))
    assert d["rebuilds"] == 1
    assert abs(d["worst_displacement"] - host) / host < 1e-5
    sim.release()


def test_tiny_skin_rebuilds_and_keeps_the_energies_exact():
    # check-displacement.mjs arm 2: a 0.004 nm skin is overrun within a step
    # or two; the device's test must rebuild every time, so the list is never
    # stale and the energies stay the oracle's
    system, coords, sim = _ala_wat(temperature=298.0, friction=10.0, seed=4242, skin=0.004)
    x0, v0 = _on_manifold(system, coords.positions, R.maxwell_boltzmann(system, 298.0, 17))
    sim.set_state(x0, v0, step=0)
    sim.run(100, 25)
    d = sim.diagnostics()
    print("skin", d["skin"], "rebuilds", d["rebuilds"], "worst displacement", d["worst_displacement"])
    assert d["skin"] == 0.004
    assert d["rebuilds"] >= 50
    assert d["worst_displacement"] <= 0.5 * 0.004
    gpu, forces = sim.evaluate()
    pos = sim.read_positions().astype(np.float64)
    ref, ref_forces = O.Oracle(system).potential(pos, alpha=ALPHA_24, grid=[24, 24, 24], f32_membership=True)
    for key, want in (("lennard_jones", ref["lennard_jones"]), ("coulomb_direct", ref["coulomb_direct_total"]),
                      ("coulomb_reciprocal", ref["reciprocal"])):
        assert rel(gpu[key], want) < 1e-6, key
    assert rel_rms(forces, ref_forces) < 1e-4
    sim.release()


def test_derived_skin_never_misses_a_pair():
    # check-displacement.mjs arm 3: the shipped skin on a minimised start;
    # after 500 steps every pair inside the cutoff is in the list the step
    # used, and rebuilds stay occasional
    system = R.workload_system("ala-389w")
    sim = R.GpuSimulation(system, temperature=298.0, seed=4242)
    sim.minimise(max_iterations=200, force_tolerance=500.0)
    sim.set_state(sim.read_positions(), R.maxwell_boltzmann(system, 298.0, 17), step=0)
    sim.run(500, 10)
    d = sim.diagnostics()
    pos = sim.read_positions()
    counts, lists = sim.read_neighbours()
    inside_counts, inside = O.pairs_f32(pos, system.box, system.cutoff, system.exclusion_start,
                                        system.exclusion_partners)
    missing = 0
    for i in range(system.atom_count):
        missing += len(set(inside[i, :inside_counts[i]].tolist()) - set(lists[i, :counts[i]].tolist()))
    print("rebuilds", d["rebuilds"], "worst displacement", d["worst_displacement"], "skin", d["skin"])
    assert missing == 0
    assert d["rebuilds"] <= 100
    assert d["worst_displacement"] <= 0.5 * d["skin"]
    sim.release()


def test_a_state_that_turns_non_finite_mid_run_is_an_error():
    # webgpu/simulation.js:544-623: a blow-up is reported, never stepped
    # through. One water oxygen's supplied noise is NaN: its velocity and then
    # its position turn non-finite inside the block.
    system, coords, sim = _ala_wat(temperature=298.0, seed=8)
    x0, v0 = _on_manifold(system, coords.positions, R.maxwell_boltzmann(system, 298.0, 8))
    noise = np.random.default_rng(3).standard_normal(3 * system.atom_count).astype(np.float32)
    water_oxygen = int(system.constraint_a[-1])  # the last constraint is a water's
    noise[3 * water_oxygen] = np.nan
    sim.set_state(x0, v0, step=0, supplied_noise=True)
    sim.set_noise(noise)
    with pytest.raises(R.NonFiniteStateError):
        sim.run(10, 10)
    # a fresh state runs again
    sim.set_state(x0, v0, step=0, supplied_noise=False)
    sim.run(10, 10)
    assert np.isfinite(sim.energy()["potential"])
    sim.release()


# ---- minimiser -----------------------------------------------------------------

def test_minimiser_follows_the_oracle_path():
    # the same steepest descent from the same lattice start in FP64: the same
    # accepted and rejected trials, potentials and positions
    system = R.workload_system("ala-389w")
    sim = R.GpuSimulation(system, temperature=300.0)
    d = sim.diagnostics()
    start = sim.read_positions().astype(np.float64)
    result = sim.minimise(max_iterations=8, force_tolerance=1.0, max_step=0.002, max_displacement=0.05)
    x_gpu = sim.read_positions().astype(np.float64)
    o = O.Oracle(system)
    x_ref, ref = o.minimise(start, max_iterations=8, force_tolerance=1.0, max_step=0.002, max_displacement=0.05,
                            tolerance=1e-10, restore_forces=True, alpha=d["pme_alpha"], grid=list(d["pme_grid"]),
                            f32_membership=True)
    gpu_potential = result["potential"] - system.dispersion_energy(d["volume"])
    print("gpu", result, "oracle", ref, "max |dx| %.3e" % np.abs(x_gpu - x_ref).max())
    assert (result["iterations"], result["rejected"], result["stalled"]) == (ref["iterations"], ref["rejected"],
                                                                            ref["stalled"])
    assert abs(result["max_step"] - ref["max_step"]) < 1e-12
    assert rel(result["max_force"], ref["max_force"]) < 1e-4
    # the paths: FP32 positions against FP64 ones after 8 trials
    assert np.abs(x_gpu - x_ref).max() < 1e-5
    # the potential the GPU reports is the FP64 energy of its own final
    # positions; the oracle's path ends ~1e-6 nm away, where a potential of
    # -2.8e3 kJ/mol with forces of ~1.8e3 kJ/mol/nm moves by ~1e-2 kJ/mol
    at_gpu, _ = o.potential(x_gpu, alpha=d["pme_alpha"], grid=list(d["pme_grid"]), f32_membership=True,
                            forces=False)
    print("potential at the GPU's positions: gpu %.6f oracle %.6f" % (gpu_potential, at_gpu["total"]))
    assert rel(gpu_potential, at_gpu["total"]) < 1e-6
    assert rel(gpu_potential, ref["potential"]) < 1e-4
    sim.release()
