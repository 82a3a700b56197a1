"""Pins for every oracle function beyond the OpenMM fixtures: textbook
routines, closed forms, invariants and brute force, each chosen so a
plausible mistake (a dropped term, a wrong sign, a transposed operand) fails.
"""
import numpy as np
import pytest
from scipy import special

import oracle as O
import rainmd_amd as R
from helpers import ALPHA_24, fixture_system


def test_erfc_matches_scipy():
    # test/physics.test.js:19-35, including erfc(2.628261)
    xs = np.concatenate([np.linspace(-3, 6, 901), [2.628261, 0.46875, 4.0, 1e-10]])
    for x in xs:
        ref = special.erfc(x)
        assert abs(O.erfc(x) - ref) <= 1e-13 * max(abs(ref), 1e-300) + 1e-300, x
        assert abs(O.erf(x) - special.erf(x)) <= 1e-13 + 1e-13 * abs(special.erf(x))


def test_alpha_known_answer():
    assert abs(O.alpha_from_tolerance(1.0, 5e-4) - 2.628261) < 1e-6


def test_bspline_partition_of_unity_and_derivatives():
    # test/physics.test.js:43-61
    for u in np.linspace(0, 0.999, 37):
        w, d = O.bspline(u)
        assert abs(w.sum() - 1) < 1e-14
        assert abs(d.sum()) < 1e-14
        # the derivative is the gradient of the weights in the offset
        h = 1e-6
        wp, _ = O.bspline(min(u + h, 0.9999999))
        wm, _ = O.bspline(max(u - h, 0.0))
        step = min(u + h, 0.9999999) - max(u - h, 0.0)
        assert np.allclose((wp - wm) / step, d, atol=1e-6)


def test_moduli_closed_form():
    # integer-point values of the order-5 B-spline are [1, 11, 11, 1]/24
    K = 24
    m = np.arange(K)
    f = np.abs(sum(c * np.exp(2j * np.pi * m * k / K) for k, c in enumerate([1 / 24, 11 / 24, 11 / 24, 1 / 24]))) ** 2
    ours = O.moduli(K)
    nyq = K // 2
    keep = m != nyq
    assert np.allclose(ours[keep], f[keep], rtol=1e-13)
    assert abs(ours[nyq] - 0.5 * (f[nyq - 1] + f[nyq + 1])) < 1e-14
    assert np.allclose(R.pme_moduli(K), ours, rtol=1e-15)


def test_rng_properties():
    # shared/random.js: integer hash, u in (0, 1), the 9-bit shift makes the
    # uniform exact in f32 as well as f64 (40000 of 40000 draws)
    draws = np.array([O.uniform(123, a, 7, c) for a in range(10000) for c in range(4)])
    assert draws.min() > 0 and draws.max() < 1
    assert np.array_equal(draws.astype(np.float32).astype(np.float64), draws)
    assert abs(draws.mean() - 0.5) < 0.01
    g = np.array([O.gaussian(5, a, 11) for a in range(20000)]).reshape(-1)
    assert abs(g.mean()) < 0.02 and abs(g.std() - 1) < 0.02
    # the host product path's vectorised hash is the same function
    assert np.array_equal(R.uniform_at(123, np.arange(100), 7, 2), [O.uniform(123, a, 7, 2) for a in range(100)])


def test_metropolis_weight_signs():
    # test/physics.test.js:541-556
    kT = 0.008314462618 * 298
    assert O.metropolis_weight(-1.0, 0.0, 1.0, 0.06, kT, 10) < 0
    assert O.metropolis_weight(1.0, 0.0, 1.0, 0.06, kT, 10) > 0
    # expansion against pressure costs P dV; the volume entropy favours it
    w = O.metropolis_weight(0.0, 0.1, 1.01, 0.06, kT, 100)
    assert abs(w - (0.06 * 0.1 - 100 * kT * np.log(1.01))) < 1e-15


@pytest.mark.parametrize("box_kind", ["rectangular", "sheared"])
def test_pme_reciprocal_matches_plain_ewald_sum(box_kind):
    # test/physics.test.js:391-519: the grid sum against an exact lattice sum,
    # on a rectangular box and on the truncated octahedron (ala-oct), where a
    # transposed wave vector would fail.
    if box_kind == "rectangular":
        system, coords = fixture_system("ala-wat", dict(constraints="none", rigid_water=False, cutoff=1.0))
    else:
        system, coords = fixture_system("ala-oct", dict(constraints="none", rigid_water=False, cutoff=1.0))
    o = O.Oracle(system)
    pme, f_pme = o.potential(coords.positions, alpha=ALPHA_24, grid=[48, 48, 48], terms=8)
    ew, f_ew = o.ewald_reciprocal(coords.positions, ALPHA_24, 12, forces=True)
    assert abs(pme["reciprocal"] - ew) / abs(ew) < 1e-5
    rms = np.sqrt(np.mean(f_ew ** 2))
    assert np.sqrt(np.mean((f_pme - f_ew) ** 2)) / rms < 1e-3


def test_ewald_sum_matches_openmm_ewald_fixture(golden):
    # ala-wat-ewald.json: OpenMM's Ewald total (dispersion on). OpenMM stops its
    # k-sum at a kmax derived from the 5e-4 tolerance, so the converged sum here
    # can differ from it by that truncation, measured at 6.1e-6 of the total.
    ref = golden("ala-wat-ewald.json")
    system, coords = fixture_system("ala-wat", dict(constraints="none", rigid_water=False, cutoff=1.0))
    o = O.Oracle(system)
    e, _ = o.potential(coords.positions, alpha=ALPHA_24, grid=[24, 24, 24], forces=False)
    ew, _ = o.ewald_reciprocal(coords.positions, ALPHA_24, 14)
    volume = system.box[0] * system.box[4] * system.box[8]
    total = e["total"] - e["reciprocal"] + ew + o.dispersion(volume, 1.0)
    assert abs(total - ref["energy"]["total"]) / abs(ref["energy"]["total"]) < 1e-5


@pytest.mark.parametrize("term,mask", [("bonded", 1), ("direct", 2), ("exceptions", 4), ("reciprocal", 8)])
def test_forces_are_the_gradient(term, mask):
    # tools/check-force-gradients.mjs: every force against a central difference
    system, coords = fixture_system("ala-wat", dict(constraints="none", rigid_water=False, cutoff=1.0))
    o = O.Oracle(system)
    x = coords.positions.copy()
    e0, f = o.potential(x, alpha=ALPHA_24, grid=[24, 24, 24], terms=mask)
    rms = np.sqrt(np.mean(f ** 2))
    h = 2e-6 if mask == 8 else 1e-7
    rng = np.random.default_rng(mask)
    for idx in rng.choice(len(x), 12, replace=False):
        xp = x.copy()
        xm = x.copy()
        xp[idx] += h
        xm[idx] -= h
        ep, _ = o.potential(xp, alpha=ALPHA_24, grid=[24, 24, 24], terms=mask, forces=False)
        em, _ = o.potential(xm, alpha=ALPHA_24, grid=[24, 24, 24], terms=mask, forces=False)
        num = -(ep["total"] - em["total"]) / (2 * h)
        assert abs(num - f.reshape(-1)[idx]) <= 1e-5 * max(rms, 1.0), (term, idx)


def test_bonded_forces_sum_to_zero():
    system, coords = fixture_system("ala-vac", dict(constraints="none", rigid_water=False, cutoff=0))
    _, f = O.Oracle(system).potential(coords.positions, method="NoCutoff", terms=1)
    assert np.abs(f.sum(0)).max() < 1e-9


def test_settle_is_the_limit_of_shake():
    system, coords = fixture_system("ala-wat", dict(constraints="HBonds", rigid_water=True, cutoff=1.0))
    o = O.Oracle(system)
    rng = np.random.default_rng(2)
    x0 = coords.positions
    x1 = x0 + rng.normal(scale=0.003, size=x0.shape)
    shaken, it = o.shake(x1, x0, 1e-13, 100000)
    assert it > 0
    d = dict(((min(a, b), max(a, b)), dd) for a, b, dd in
             zip(system.constraint_a, system.constraint_b, system.constraint_distance))
    # a water: its H-H constraint is the 0.15136 nm one; the oxygen is the
    # partner both hydrogens are constrained to
    h1, h2 = next(k for k, dd in d.items() if abs(dd - 0.15136) < 1e-9 and k[0] > 40)
    w = next(a for (a, b) in d if b == h1 and (a, h2) in d)
    assert h1 == w + 1 and h2 == w + 2
    m = system.masses
    old = x0[3 * w:3 * w + 9]
    new = x1[3 * w:3 * w + 9]
    dHH = d[(h1, h2)]
    dOH = d[(w, h1)]
    settled = O.settle(old, new, m[w], m[h1], dOH, dHH)
    assert np.abs(settled - shaken[3 * w:3 * w + 9]).max() < 1e-10
    p = settled.reshape(3, 3)
    assert abs(np.linalg.norm(p[0] - p[1]) - dOH) < 1e-12
    assert abs(np.linalg.norm(p[0] - p[2]) - dOH) < 1e-12
    assert abs(np.linalg.norm(p[1] - p[2]) - dHH) < 1e-12


def test_one_step_closed_form_without_constraints():
    # B A O A with zero friction and no constraints is x + dt v + dt^2/(2m) F
    system, coords = fixture_system("ala-vac", dict(constraints="none", rigid_water=False, cutoff=0))
    o = O.Oracle(system)
    rng = np.random.default_rng(0)
    v = rng.normal(size=3 * system.atom_count)
    f = rng.normal(size=3 * system.atom_count) * 100
    m = np.repeat(system.masses, 3)
    dt = 0.002
    x1, v1, rc = o.langevin_step(coords.positions, v, f, timestep=dt, friction=0.0, noise=np.zeros_like(v))
    assert rc == 0
    assert np.allclose(x1, coords.positions + dt * v + dt * dt / (2 * m) * f, rtol=0, atol=1e-14)
    assert np.allclose(v1, v + dt / (2 * m) * f, rtol=0, atol=1e-13)


def test_one_step_thermostat_closed_form():
    # O alone: v' = a v + sqrt(1 - a^2) sqrt(kT/m) xi with a = exp(-gamma dt)
    system, coords = fixture_system("ala-vac", dict(constraints="none", rigid_water=False, cutoff=0))
    o = O.Oracle(system)
    n3 = 3 * system.atom_count
    xi = np.random.default_rng(1).normal(size=n3)
    dt, gamma, T = 0.002, 5.0, 310.0
    _, v1, _ = o.langevin_step(coords.positions, np.zeros(n3), np.zeros(n3), timestep=dt, friction=gamma,
                               temperature=T, noise=xi)
    a = np.exp(-gamma * dt)
    m = np.repeat(system.masses, 3)
    assert np.allclose(v1, np.sqrt(1 - a * a) * np.sqrt(0.008314462618 * T / m) * xi, rtol=1e-13)


def test_f32_cells_and_pairs_agree_with_fp64_away_from_boundaries():
    system, coords = fixture_system("ala-wat", dict(constraints="HBonds", rigid_water=True, cutoff=1.0))
    pos32 = coords.positions.astype(np.float32)
    box = system.box
    counts = [2, 2, 2]
    cells = O.cells_f32(pos32, box, counts)
    x = pos32.astype(np.float64).reshape(-1, 3)
    L = np.array([box[0], box[4], box[8]])
    frac = x / L
    frac -= np.floor(frac)
    c3 = np.minimum((frac * counts).astype(int), 1)
    ref = (c3[:, 0] * 2 + c3[:, 1]) * 2 + c3[:, 2]
    margin = np.abs(frac * 2 - np.round(frac * 2)).min(1) > 1e-5
    assert np.array_equal(cells[margin], ref[margin])
    rl = 1.2
    counts_f32, lists = O.pairs_f32(pos32, box, rl, system.exclusion_start, system.exclusion_partners)
    for i in range(0, system.atom_count, 97):
        d = x - x[i]
        d -= L * np.floor(d / L + 0.5)
        r = np.sqrt((d ** 2).sum(1))
        excl = set(system.exclusion_partners[system.exclusion_start[i]:system.exclusion_start[i + 1]]) | {i}
        sure = {j for j in np.nonzero(r < rl - 1e-5)[0] if j not in excl}
        maybe = {j for j in np.nonzero(r <= rl + 1e-5)[0] if j not in excl}
        got = set(lists[i, :counts_f32[i]].tolist())
        assert sure <= got <= maybe


def test_binned_pair_finder_equals_the_all_pairs_one():
    """oracle_pairs_f32_binned (coarse-cell candidates, used for boxes too
    large for the all-pairs loop) returns the all-pairs finder's sets, in a
    rectangular and a triclinic box, at list cutoffs that give 2 to 5 coarse
    cells per axis."""
    system, coords = fixture_system("ala-wat", dict(constraints="HBonds", rigid_water=True, cutoff=1.0))
    pos32 = coords.positions.astype(np.float32)
    es, ep = system.exclusion_start, system.exclusion_partners
    boxes = [np.asarray(system.box, dtype=np.float64)]
    sheared = np.array(boxes[0]).reshape(3, 3).copy()
    sheared[1, 0] = 0.3 * sheared[0, 0]
    sheared[2, 0], sheared[2, 1] = -0.2 * sheared[0, 0], 0.25 * sheared[1, 1]
    boxes.append(sheared.reshape(9))
    for box in boxes:
        for rl in (0.7, 1.2):
            counts, lists = O.pairs_f32(pos32, box, rl, es, ep)
            bcounts, blists = O.pairs_f32_binned(pos32, box, rl, es, ep)
            assert np.array_equal(counts, bcounts), rl
            assert counts.sum() > 0
            for i in range(system.atom_count):
                assert np.array_equal(np.sort(lists[i, :counts[i]]), blists[i, :bcounts[i]]), (rl, i)


def test_centroid_scaling_keeps_molecules_rigid():
    system, coords = fixture_system("ala-wat", dict(constraints="HBonds", rigid_water=True, cutoff=1.0))
    o = O.Oracle(system)
    x = coords.positions
    y = o.scale_molecules(x, 1.02).reshape(-1, 3)
    x3 = x.reshape(-1, 3)
    for m in range(0, system.molecule_count, 37):
        atoms = system.molecule_atoms[system.molecule_start[m]:system.molecule_start[m + 1]]
        w = system.masses[atoms]
        cx = (x3[atoms] * w[:, None]).sum(0) / w.sum()
        cy = (y[atoms] * w[:, None]).sum(0) / w.sum()
        assert np.allclose(cy, 1.02 * cx, atol=1e-12)
        assert np.allclose(y[atoms] - cy, x3[atoms] - cx, atol=1e-12)
