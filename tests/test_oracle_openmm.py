"""The FP64 oracle against the reference's OpenMM fixtures, at the
reference's own bars (tools/check-against-openmm.mjs:137-173): bonded 1e-10,
LJ 1e-9, Coulomb and total 1e-7, forces 1e-6 of the rms force."""
import numpy as np
import pytest

import oracle as O
from helpers import CASES, fixture_system, group_energy, rel


@pytest.mark.parametrize("case", list(CASES))
def test_oracle_matches_openmm(case, golden):
    topology, reference, build, evaluate = CASES[case]
    system, coords = fixture_system(topology, build)
    ref = golden(reference)
    positions = coords.positions
    restraint = None
    if ref.get("restraint", {}).get("force", 0) > 0:
        positions = np.array(ref["positions"])
        atoms = np.array(ref["restraint"]["atoms"])
        restraint = dict(force=ref["restraint"]["force"], atoms=atoms,
                         reference=coords.positions.reshape(-1, 3)[atoms].reshape(-1))
    e, f = O.Oracle(system).potential(positions, restraint=restraint, **evaluate)
    for key, cls in [("bond", "HarmonicBondForce"), ("angle", "HarmonicAngleForce"),
                     ("torsion", "PeriodicTorsionForce"), ("restraint", "CustomExternalForce")]:
        theirs = group_energy(ref, cls)
        if theirs is not None:
            assert rel(e[key], theirs) <= 1e-10, key
    assert rel(e["lennard_jones"], ref["energy"]["ljOnly"]) <= 1e-9
    assert rel(e["coulomb"], ref["energy"]["coulombOnly"]) <= 1e-7
    assert rel(e["total"], ref["energy"]["total"]) <= 1e-7
    if "forces" in ref:
        theirs = np.array(ref["forces"]["total"])
        rms = np.sqrt(np.mean(theirs ** 2))
        assert np.abs(f - theirs).max() / rms <= 1e-6


def test_dispersion_known_answer_and_tail_fixture(golden):
    # shared/dispersion.js: -41.537158 kJ/mol on ala-wat (test/physics.test.js:521-539)
    system, coords = fixture_system("ala-wat", dict(constraints="none", rigid_water=False, cutoff=1.0))
    o = O.Oracle(system)
    vol = system.box[0] * system.box[4] * system.box[8]
    disp = o.dispersion(vol, 1.0)
    assert abs(disp - (-41.537158)) < 1e-6 * 41.537158 + 1e-6
    # scales as 1/V
    assert abs(o.dispersion(2 * vol, 1.0) - disp / 2) < 1e-12
    tail = golden("ala-wat-pme-tail.json")
    free = golden("ala-wat-pme-free.json")
    # the two dumps differ only by the correction: their totals differ by it
    assert rel(tail["energy"]["total"] - free["energy"]["total"], disp) < 1e-9
