"""Pins for the oracle's host-side pieces: the counter-hash noise against the
reference's own vectors, kinetic energy and temperature, Maxwell-Boltzmann
velocities, the cell plan and skin, and the steepest-descent minimiser.

Each check is one a plausible mistake fails: a wrong hash constant or key
order, a dropped 1/2 or a mass in place of its inverse, a transposed plan
width, a minimiser that steps along the wrong direction or stops early.
"""
import json
import os

import numpy as np
import pytest

import oracle as O
import rainmd_amd as R
from helpers import fixture_system

HERE = os.path.dirname(os.path.abspath(__file__))
BOLTZMANN = 0.008314462618


def _vectors():
    with open(os.path.join(HERE, "golden", "random", "random_vectors.json")) as f:
        return json.load(f)


def test_uniforms_and_normals_match_the_reference_vectors():
    # uniformAt / gaussianTriple (shared/random.js:59-87), vectors written by
    # the reference's own module (tests/golden/random/make_random_vectors.mjs)
    data = _vectors()
    for seed, a, b, c, u in data["uniforms"]:
        assert O.uniform(seed, a, b, c) == u, (seed, a, b, c)
    seeds, a, b, c, u = (np.array(col) for col in zip(*data["uniforms"]))
    for s in np.unique(seeds):
        k = seeds == s
        assert np.array_equal(R.uniform_at(int(s), a[k], b[k], c[k]), u[k])
    for seed, atom, step, *g in data["triples"]:
        # both in FP64; the reference's Math.log/cos/sin and C's libm may
        # differ in the last bits
        assert np.allclose(O.gaussian(seed, atom, step), g, rtol=1e-14, atol=1e-15), (seed, atom, step)


def _lowbias32_inverse(x):
    # the published inverse of the lowbias32 mix (C. Wellons, hash-prospector)
    x &= 0xffffffff
    x ^= x >> 16
    x = (x * 0x43021123) & 0xffffffff
    x ^= (x >> 15) ^ (x >> 30)
    x = (x * 0x1d69e2a5) & 0xffffffff
    x ^= x >> 16
    return x


def test_mix_is_lowbias32():
    # the oracle's mix undone by the published inverse: a wrong shift or
    # multiplier in the forward hash breaks the round trip
    rng = np.random.default_rng(7)
    values = [0, 1, 0xffffffff, 0x80000000, 0x5eed] + rng.integers(0, 2 ** 32, 2000, dtype=np.uint64).tolist()
    for v in values:
        assert _lowbias32_inverse(O.mix(int(v))) == int(v)


def test_kinetic_energy_and_temperature_of_unit_velocities():
    # test/physics.test.js:160-175: unit velocities give K = (1/2) sum(m) per
    # axis, so the temperature is a statement about the denominator
    system, _ = fixture_system("ala-wat", dict(constraints="HBonds", rigid_water=True, cutoff=1.0))
    o = O.Oracle(system)
    k = o.kinetic(np.ones(3 * system.atom_count))
    assert abs(k - 1.5 * system.masses.sum()) < 1e-9
    assert system.degrees_of_freedom == 3 * 1190 - 1176 - 3
    v = np.zeros(3 * system.atom_count)
    v[3 * 5 + 1] = 2.0  # one atom, one axis: (1/2) m v^2
    assert abs(o.kinetic(v) - 0.5 * system.masses[5] * 4.0) < 1e-12


def test_velocities_follow_maxwell_boltzmann_per_atom_type():
    # setVelocitiesToTemperature (cpu/integrate/langevin.js:211-230): each
    # component of an atom of mass m has variance kT/m and mean zero
    system = R.workload_system("ala-944w")
    o = O.Oracle(system)
    T = 300.0
    v = o.velocities(T, 202).reshape(-1, 3)
    for mass in (15.999, 1.008):
        atoms = np.nonzero(np.abs(system.masses - mass) < 1e-3)[0]
        assert len(atoms) > 900
        comp = v[atoms].reshape(-1)
        expected = BOLTZMANN * T / mass
        sigma_var = expected * np.sqrt(2.0 / (len(comp) - 1))
        assert abs(comp.var() - expected) < 3 * sigma_var, mass
        assert abs(comp.mean()) < 3 * np.sqrt(expected / len(comp)), mass
    # seeded: the same draw again, and the product's mirror draws the same
    assert np.array_equal(o.velocities(T, 202), v.reshape(-1))
    assert np.allclose(R.maxwell_boltzmann(system, T, 202), v.reshape(-1), rtol=1e-15, atol=0)


def test_cell_plan_reference_rules():
    # test/physics.test.js:178-204: at cutoffs 1.0, 0.7, 0.5 and skin 0.1 on
    # the solvated fixture, the deduplicated stencil visits distinct cells and
    # a cell is at least the list cutoff wide
    t = R.Topology.load("ala-wat.prmtop")
    c = R.Coordinates.load("ala-wat.inpcrd", 0)
    for cutoff in (1.0, 0.7, 0.5):
        system = R.System.build(t, c, constraints="HBonds", rigid_water=True, cutoff=cutoff)
        plan = O.Oracle(system).plan_cells(0.1)
        width = np.array([system.box[0], system.box[4], system.box[8]])
        counts = np.array(plan["counts"])
        assert np.all(width / counts >= cutoff + 0.1 - 1e-12)
        # one cell more per axis would be thinner than the list cutoff
        assert np.all(width / (counts + 1) < cutoff + 0.1)
        offsets = [[-1, 0, 1] if k >= 3 else [0, 1] if k == 2 else [0] for k in counts]
        cells = {(dx % counts[0], dy % counts[1], dz % counts[2])
                 for dx in offsets[0] for dy in offsets[1] for dz in offsets[2]}
        assert plan["stencil"] == len(cells) == len(offsets[0]) * len(offsets[1]) * len(offsets[2])


def test_cell_plan_and_skin_at_the_workload_sizes():
    # The sizes the reference's formulas give for the 1189-, 2854- and
    # 100,200-atom workloads: skins 0.217 (capped by the 2.28 nm box), 0.40
    # scaled to 300 K, and the configured 0.15; 2, 2 and 8 cells per axis;
    # capacities of about 1169, 1845 and 1277 neighbours.
    cases = [("ala-389w", None, 0.2167, 2, 1169), ("ala-944w", None, 0.40 * np.sqrt(300 / 298), 2, 1845),
             ("water-33400", 0.15, 0.15, 8, 1277)]
    for name, skin, want_skin, cells, capacity in cases:
        system = R.workload_system(name)
        o = O.Oracle(system)
        s = skin if skin else o.skin_for(0.002, 300.0)
        assert abs(s - want_skin) < 2e-4, name
        # the product's own skinFor is the same function
        if skin is None:
            assert abs(system.skin_for(0.002, 300.0) - s) < 1e-12, name
        plan = o.plan_cells(s)
        assert plan["counts"] == [cells] * 3, name
        assert abs(plan["max_neighbours"] - capacity) <= 0.02 * capacity, (name, plan)


def test_skin_is_capped_by_the_box_and_floored():
    system = R.workload_system("ala-389w")
    o = O.Oracle(system)
    box = np.array(system.box, dtype=np.float64)
    room = 0.5 * box[0] - system.cutoff
    assert abs(o.skin_for(0.002, 300.0) - 0.9 * room) < 1e-12
    # a hot long step asks for more than the room; a tiny box leaves none
    assert abs(o.skin_for(0.004, 600.0) - 0.9 * room) < 1e-12
    small = box * (2 * (system.cutoff + 0.01) / box[0])
    assert o.skin_for(0.002, 300.0, small) == 0.1


def test_minimiser_first_trial_moves_the_largest_force_by_max_step():
    # webgpu/simulation.js:888-951: the step is sized so the largest force
    # moves exactly max_step, along the force
    system, coords = fixture_system("ala-vac", dict(constraints="none", rigid_water=False, cutoff=0))
    o = O.Oracle(system)
    x0 = coords.positions.astype(np.float64)
    _, f0 = o.potential(x0, method="NoCutoff")
    x1, r = o.minimise(x0, max_iterations=1, force_tolerance=1.0, max_step=0.002, method="NoCutoff")
    d = (x1 - x0.reshape(-1)).reshape(-1, 3)
    mags = np.linalg.norm(f0, axis=1)
    top = int(np.argmax(mags))
    assert abs(np.linalg.norm(d[top]) - 0.002) < 1e-15
    cosines = np.einsum("ij,ij->i", d, f0) / (np.linalg.norm(d, axis=1) * mags)
    assert np.all(cosines > 1 - 1e-12)
    assert r["iterations"] == 1 and r["rejected"] == 0 and abs(r["max_step"] - 0.0026) < 1e-15


def test_minimiser_displacement_cap_and_monotone_descent():
    system, coords = fixture_system("ala-vac", dict(constraints="none", rigid_water=False, cutoff=0))
    o = O.Oracle(system)
    x0 = coords.positions.astype(np.float64)
    x1, _ = o.minimise(x0, max_iterations=1, max_step=0.5, max_displacement=0.01, method="NoCutoff")
    assert np.linalg.norm((x1 - x0.reshape(-1)).reshape(-1, 3), axis=1).max() <= 0.01 + 1e-15
    energies = [o.minimise(x0, max_iterations=k, method="NoCutoff")[1]["potential"] for k in range(1, 30)]
    assert np.all(np.diff(energies) <= 0)


@pytest.mark.parametrize("restore_forces", [True, False])
def test_minimiser_reaches_the_lbfgs_minimum(restore_forces):
    # Steepest descent to a tight force tolerance lands in the minimum an
    # independent quasi-Newton minimisation of the same FP64 energy finds.
    from scipy.optimize import minimize

    system, coords = fixture_system("ala-vac", dict(constraints="none", rigid_water=False, cutoff=0))
    o = O.Oracle(system)
    x0 = coords.positions.astype(np.float64)
    x, r = o.minimise(x0, max_iterations=40000, force_tolerance=0.5, method="NoCutoff",
                      restore_forces=restore_forces)
    assert r["converged"] and r["max_force"] <= 0.5

    def fun(y):
        e, f = o.potential(y, method="NoCutoff")
        return e["total"], -f.reshape(-1)

    res = minimize(fun, x0.reshape(-1), jac=True, method="L-BFGS-B",
                   options=dict(gtol=1e-6, maxiter=100000, ftol=1e-15))
    print("steepest descent", r["potential"], "L-BFGS", res.fun)
    assert res.fun - 1e-6 <= r["potential"] < res.fun + 1e-2


def test_minimiser_constraints_are_projected():
    system, coords = fixture_system("ala-wat", dict(constraints="HBonds", rigid_water=True, cutoff=1.0))
    o = O.Oracle(system)
    x0, _ = o.shake(coords.positions, coords.positions, 1e-12)
    x, r = o.minimise(x0, max_iterations=3, alpha=2.6282608939112583, grid=[24, 24, 24])
    p = x.reshape(-1, 3)
    d = np.linalg.norm(p[system.constraint_a] - p[system.constraint_b], axis=1)
    assert np.abs(d / system.constraint_distance - 1).max() < 1e-9
    assert r["iterations"] == 3
