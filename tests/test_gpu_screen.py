"""The barostat's learned screen on the GPU (1189-atom workload, NPT at 1 bar
and 300 K, a volume move every 2 steps): the network trains on the run's own
exact moves, then turns moves down before their exact evaluation, and the
NPT statistics judge finds the same ensemble as the plain barostat."""
import numpy as np
import pytest

import rainmd_amd as R
from rainmd_amd import npt

pytestmark = pytest.mark.gpu

EQUILIBRATE = 2000
SAMPLES = 1500   # volume samples, one every 10 steps
TRAINING_MOVES = 400


def _run(screen):
    system = R.workload_system("ala-389w")
    sim = R.GpuSimulation(system, temperature=300.0, pressure=1.0, barostat_interval=2, seed=21)
    sim.minimise(max_iterations=100, force_tolerance=1000.0)
    sim.set_state(sim.read_positions(), R.maxwell_boltzmann(system, 300.0, 21), step=0)
    sim.run(EQUILIBRATE)
    if screen:
        sim.configure_screen(training_moves=TRAINING_MOVES, training_steps=3000, learning_rate=2e-3, seed=7)
        # the collection phase: plain moves, kept as training samples
        while sim.screen_stats()["phase"] == 0:
            sim.run(20)
    volumes = []
    for _ in range(SAMPLES):
        sim.run(10)
        volumes.append(sim.diagnostics()["volume"])
    stats = sim.screen_stats() if screen else None
    d = sim.diagnostics()
    sim.release()
    return np.array(volumes), stats, d, float(system.masses.sum())


def test_screen_skips_exact_evaluations_and_keeps_the_ensemble():
    plain, _, plain_diag, mass = _run(False)
    screened, st, diag, _ = _run(True)
    print("screen", st)
    print("barostat plain", plain_diag["barostat_accepted"], "/", plain_diag["barostat_attempts"],
          "screened", diag["barostat_accepted"], "/", diag["barostat_attempts"])
    verdict = npt.judge(plain, screened, mass, 300.0, exact_evaluations=st["exact_evaluations"], moves=st["moves"])
    print("judge", verdict)
    assert st["phase"] == 1 and st["samples"] == 2 * TRAINING_MOVES
    assert st["moves"] > 1000
    # what the network buys: moves turned down with no exact evaluation ...
    assert verdict["skipped_fraction"] > 0.15
    # ... because it predicts the energy change better than the first-order
    # (scaling-derivative) estimate it corrects
    assert st["residual_rms"] < 0.8 * st["estimate_rms"]
    # and the ensemble is the plain barostat's
    assert verdict["density_difference_sigma"] < 4.0
    assert verdict["ks_p"] > 1e-4
    assert 0.5 < verdict["compressibility_ratio"] < 2.0
