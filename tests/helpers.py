"""Shared builders for the tests: the reference fixtures' systems."""
import numpy as np

import rainmd_amd as R

ALPHA_24 = 2.6282608939112583  # the fixtures' forced alpha (cutoff 1.0 nm, tolerance 5e-4)

CASES = {
    # name: (topology, reference json, build options, evaluation options)
    "ala-vac": ("ala-vac", "ala-vac-nocutoff.json", dict(constraints="none", rigid_water=False, cutoff=0),
                dict(method="NoCutoff")),
    "ala-wat-free": ("ala-wat", "ala-wat-pme-free.json", dict(constraints="none", rigid_water=False, cutoff=1.0),
                     dict(method="PME", alpha=ALPHA_24, grid=[24, 24, 24])),
    "ala-wat": ("ala-wat", "ala-wat-pme.json", dict(constraints="HBonds", rigid_water=True, cutoff=1.0),
                dict(method="PME", alpha=ALPHA_24, grid=[24, 24, 24])),
    "ala-wat-restraint": ("ala-wat", "ala-wat-restraint-500.json",
                          dict(constraints="HBonds", rigid_water=True, cutoff=1.0),
                          dict(method="PME", alpha=ALPHA_24, grid=[24, 24, 24])),
}


def fixture_system(topology, build):
    t = R.Topology.load(topology + ".prmtop")
    c = R.Coordinates.load(topology + ".inpcrd", 0)
    return R.System.build(t, c, **build), c


def group_energy(reference, cls):
    for key, value in reference["energy"]["byForceGroup"].items():
        if key.endswith(":" + cls):
            return value["energy"]
    return None


def rel(a, b, floor=1.0):
    return abs(a - b) / max(abs(b), floor)


def rel_rms(a, b):
    a = np.asarray(a).reshape(-1)
    b = np.asarray(b).reshape(-1)
    return float(np.sqrt(np.mean((a - b) ** 2)) / np.sqrt(np.mean(b ** 2)))
