"""Host setup through the C ABI (no GPU): the AMBER readers and buildSystem
against the reference's own parse tests (test/amber-parse.test.js) and the
OpenMM fixtures' constraint and exception tables."""
import numpy as np
import pytest

import rainmd_amd as R
from rainmd_amd.simulation import read_data


def test_vacuum_topology_counts():
    # test/amber-parse.test.js:33-46
    s = R.System.build(R.Topology.load("ala-vac.prmtop"), None, constraints="none", rigid_water=False, cutoff=0)
    assert s.atom_count == 22
    assert s.residue_count == 3
    assert s.molecule_count == 1
    assert len(s.bond_a) == 21
    assert len(s.angle_a) == 36
    assert len(s.torsion_a) == 67
    assert len(s.exception_a) == 98
    assert abs(s.charges.sum()) < 1e-6
    # 41 of the 98 exceptions are scaled 1-4 pairs
    assert np.count_nonzero(s.exception_lj_a) + np.count_nonzero(
        (s.exception_lj_a == 0) & (s.exception_charge_product != 0)) == 41


def test_water_is_three_constraints_and_no_angle():
    s = R.System.build(R.Topology.load("ala-wat.prmtop"), R.Coordinates.load("ala-wat.inpcrd", 1190))
    assert s.water_count == 388
    waters = set(np.nonzero(s.residue_of_atom >= 3)[0])
    assert not any(int(b) in waters for b in s.angle_b)


def test_constraints_and_exceptions_match_openmm(golden):
    # test/amber-parse.test.js:83-95 and the fixtures' own tables, as sets
    ref = golden("ala-wat-pme.json")
    s = R.System.build(R.Topology.load("ala-wat.prmtop"), R.Coordinates.load("ala-wat.inpcrd", 1190),
                       constraints="HBonds", rigid_water=True, cutoff=1.0)
    assert len(s.constraint_a) == ref["constraintCount"] == 1176
    assert len(s.exception_a) == ref["exceptionCount"] == 1262
    assert s.degrees_of_freedom == 3 * 1190 - 1176 - 3
    ours = {(min(a, b), max(a, b)): d for a, b, d in zip(s.constraint_a, s.constraint_b, s.constraint_distance)}
    theirs = {(min(a, b), max(a, b)): d for a, b, d in ref["constraints"]}
    assert ours.keys() == theirs.keys()
    assert max(abs(ours[k] - theirs[k]) for k in ours) < 1e-12
    # OpenMM exceptions: [a, b, chargeProd, sigma, epsilon]
    ex = {(min(int(e[0]), int(e[1])), max(int(e[0]), int(e[1]))): e for e in ref["exceptions"]}
    assert set(zip(s.exception_a.tolist(), s.exception_b.tolist())) == set(ex)
    for a, b, qq in zip(s.exception_a, s.exception_b, s.exception_charge_product):
        assert abs(qq - ex[(a, b)][2]) < 1e-9


def test_coordinate_count_mismatch_is_refused():
    with pytest.raises(R.ParseError, match="not from the same tleap run"):
        R.Coordinates.load("ala-wat.inpcrd", 22)


def test_netcdf_restart_is_refused_by_name():
    with pytest.raises(R.ParseError, match="NetCDF"):
        R.Coordinates(b"CDF\x01rest of a binary file")


def test_rectangular_box_reduces_to_exact_zeros():
    v = R.box_vectors_from_lengths_and_angles([2.71882, 2.97181, 2.5397], [90, 90, 90])
    for k in (1, 2, 3, 5, 6, 7):
        assert v[k] == 0.0
    assert abs(v[0] - 2.71882) < 1e-12


def test_truncated_octahedron_is_reduced():
    v = R.box_vectors_from_lengths_and_angles([2.8011, 2.8011, 2.8011], [109.4712190] * 3).reshape(3, 3)
    assert v[0, 1] == 0 and v[0, 2] == 0 and v[1, 2] == 0
    assert abs(v[1, 0]) <= 0.5 * v[0, 0] + 1e-12 and abs(v[2, 0]) <= 0.5 * v[0, 0] + 1e-12
    assert abs(v[2, 1]) <= 0.5 * v[1, 1] + 1e-12


def test_cutoff_over_half_the_box_is_refused():
    with pytest.raises(R.SystemError_, match="does not fit in a box"):
        R.System.build(R.Topology.load("ala-wat.prmtop"), R.Coordinates.load("ala-wat.inpcrd", 1190), cutoff=2.0)


def test_hangles_is_refused():
    with pytest.raises(R.SystemError_, match="HAngles"):
        R.System.build(R.Topology.load("ala-vac.prmtop"), None, constraints="HAngles")


def test_unknown_energy_flag_is_refused():
    text = read_data("ala-vac.prmtop").decode()
    cmap = "%FLAG CMAP_COUNT\n%FORMAT(2I8)\n       1       1\n"
    with pytest.raises(R.ParseError, match="CMAP"):
        R.Topology(text + cmap)
    other = "%FLAG SOMETHING_NEW\n%FORMAT(1I8)\n       1\n"
    with pytest.raises(R.ParseError, match="does not know"):
        R.Topology(text + other)


def test_fixed_width_fields_are_sliced_not_split():
    # a 10I8 line with two adjacent negative numbers (test/amber-parse.test.js:20-24)
    text = read_data("ala-vac.prmtop").decode()
    marker = "%FLAG RADIUS_SET"
    extra = "%FLAG JOIN_ARRAY\n%FORMAT(10I8)\n-1234567-7654321\n"
    R.Topology(text.replace(marker, extra + marker))


def test_pme_parameters_known_answers():
    # test/physics.test.js:37-41: alpha(1.0 nm, 5e-4) = 2.628261
    p = R.pme_parameters(np.diag([2.7188215, 2.9718097, 2.5397026]).reshape(-1), 1.0)
    assert abs(p["alpha"] - 2.628261) < 1e-6
    # OpenMM's rule gives [22, 24, 21] for this box, rounded up to 2/3/5-smooth sizes
    assert p["grid"] == [24, 24, 24]
    # radix-7 sizes are accepted when forced; 11 is not transformable
    assert R.pme_parameters(np.eye(3).reshape(-1) * 10, 1.0, grid=[112, 112, 112])["grid"] == [112, 112, 112]
    with pytest.raises(R.RainmdError):
        R.pme_parameters(np.eye(3).reshape(-1) * 10, 1.0, grid=[22, 24, 24])


def test_moduli_nyquist_patch():
    # test/physics.test.js:43-61: the 24^3 Nyquist modulus is the neighbours' mean
    m = R.pme_moduli(24)
    assert abs(m[12] - 0.5 * (m[11] + m[13])) < 1e-12
    assert np.all(m > 0)


def test_solvated_builder_is_seeded_and_clean():
    a = R.workload_system("ala-389w")
    b = R.workload_system("ala-389w")
    assert np.array_equal(a.positions, b.positions)
    assert a.atom_count == 22 + 3 * 389 == 1189
    assert a.water_count == 389 and a.molecule_count == 390
    x = a.positions.reshape(-1, 3)
    # rigid water geometry exactly at the constraint lengths (the solute keeps
    # its file coordinates, written to 1e-7 A)
    water = a.constraint_a >= 22
    d = np.linalg.norm(x[a.constraint_a[water]] - x[a.constraint_b[water]], axis=1)
    assert np.max(np.abs(d - a.constraint_distance[water])) < 1e-12
    # every oxygen at least the clearance away from the solute (minimum image)
    L = a.box[0]
    oxy = x[22::3]
    sol = x[:22]
    dd = oxy[:, None, :] - sol[None, :, :]
    dd -= L * np.floor(dd / L + 0.5)
    assert np.sqrt((dd ** 2).sum(-1)).min() > 0.25 - 0.02 * np.sqrt(3)
    # 1.0 g/cm^3 at this edge: 7152.4 g/mol in a 2.2816 nm cube
    mass = a.masses.sum()
    density = mass / 6.02214076e23 / (L * 1e-7) ** 3
    assert abs(density - 1.0) < 0.01
    c = R.System.solvated(389, 2.2816, seed=102)
    assert not np.array_equal(a.positions, c.positions)


def test_solute_heavy_atoms_are_the_restraint_fixture_selection(golden):
    # ala-wat-restraint-500.json restrains soluteHeavyAtoms (shared/topology/restraints.js:46-49)
    fixture = golden("ala-wat-restraint-500.json")["restraint"]
    s = R.System.build(R.Topology.load("ala-wat.prmtop"), R.Coordinates.load("ala-wat.inpcrd", 1190),
                       constraints="HBonds", cutoff=1.0)
    assert s.solute_heavy_atoms().tolist() == fixture["atoms"]
    # a pure water box has no solute
    water = R.System.solvated(64, 1.9, 3, include_solute=False, cutoff=0.9)
    assert water.solute_heavy_atoms().size == 0


def test_rebox_keeps_the_topology_and_checks_the_cutoff():
    s = R.System.build(R.Topology.load("ala-wat.prmtop"), R.Coordinates.load("ala-wat.inpcrd", 1190),
                       constraints="HBonds", cutoff=1.0)
    box = s.box.copy()
    box[[0, 4, 8]] *= 1.013
    t = s.rebox(box)
    assert np.array_equal(t.box, box)
    assert t.atom_count == s.atom_count and np.array_equal(t.constraint_a, s.constraint_a)
    assert np.array_equal(t.exclusion_partners, s.exclusion_partners)
    with pytest.raises(R.SystemError_, match="cutoff"):
        s.rebox(np.diag([1.9, 3.0, 3.0]).reshape(9))
    with pytest.raises(R.RainmdError):
        s.rebox([3.0, 0.1, 0, 0, 3.0, 0, 0, 0, 3.0])
