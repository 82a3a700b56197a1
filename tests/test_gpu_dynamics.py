"""GPU dynamics: the fused BAOAB step against the FP64 host integrator, and
the simulation surface (constraints, thermostat, barostat, minimiser,
capacity growth, stop) through the C ABI."""
import numpy as np
import pytest

import oracle as O
import rainmd_amd as R
from helpers import ALPHA_24, fixture_system, rel

pytestmark = pytest.mark.gpu


def constraint_error(system, positions):
    x = np.asarray(positions, dtype=np.float64).reshape(-1, 3)
    d = np.linalg.norm(x[system.constraint_a] - x[system.constraint_b], axis=1)
    return np.max(np.abs(d - system.constraint_distance) / system.constraint_distance)


def test_one_step_matches_host_integrator_with_shared_noise_and_forces():
    # tools/gpu/check-integrator.mjs:68-127: one step, supplied noise, shared forces
    system, coords = fixture_system("ala-wat", dict(constraints="HBonds", rigid_water=True, cutoff=1.0))
    sim = R.GpuSimulation(system, alpha=ALPHA_24, grid=[24, 24, 24], tolerance=1e-10, temperature=298.0,
                          friction=1.0, seed=9)
    n = system.atom_count
    x0 = coords.positions.astype(np.float32).astype(np.float64)
    v0 = R.maxwell_boltzmann(system, 298.0, 9).astype(np.float32).astype(np.float64)
    noise = np.random.default_rng(1).standard_normal(3 * n).astype(np.float32)
    # put the start on the constraint manifold first, as createState does
    o = O.Oracle(system)
    x0, _ = o.shake(x0, x0, 1e-12)
    v0, _ = o.rattle(x0, v0, 1e-12)
    x0 = x0.astype(np.float32).astype(np.float64)
    v0 = v0.astype(np.float32).astype(np.float64)
    sim.set_state(x0, v0, step=0, supplied_noise=True)
    sim.set_noise(noise)
    _, f0 = sim.evaluate()
    sim.step_once()
    st = sim.read_state()
    x1, v1 = st["positions"].astype(np.float64), st["velocities"].astype(np.float64)
    xo, vo, rc = o.langevin_step(x0, v0, f0.reshape(-1), timestep=0.002, temperature=298.0, friction=1.0,
                                 tolerance=1e-12, max_iterations=5000, noise=noise.astype(np.float64))
    assert rc == 0
    _, f1 = sim.evaluate()
    vo, rc = o.langevin_finish(xo, vo, f1.reshape(-1), tolerance=1e-12, max_iterations=5000)
    dx = np.abs(x1 - xo).max()
    dv = np.abs(v1 - vo).max()
    print("one step: max |dx| %.3e nm, max |dv| %.3e nm/ps" % (dx, dv))
    assert dx < 1e-5
    assert dv < 1e-3 * np.abs(vo).max()


def test_nvt_run_holds_temperature_and_constraints():
    system, coords = fixture_system("ala-wat", dict(constraints="HBonds", rigid_water=True, cutoff=1.0))
    sim = R.GpuSimulation(system, alpha=ALPHA_24, grid=[24, 24, 24], temperature=298.0, seed=3)
    sim.set_state(coords.positions, R.maxwell_boltzmann(system, 298.0, 3), step=0)
    temps = []
    for _ in range(20):
        sim.run(100)
        temps.append(sim.energy()["temperature"])
    e = sim.energy()
    print("temperatures", np.round(temps, 1))
    assert np.isfinite(e["potential"])
    assert 260 < np.mean(temps[5:]) < 340
    assert constraint_error(system, sim.read_positions()) < 1e-5
    d = sim.diagnostics()
    assert d["rebuilds"] > 1 and d["steps"] == 2000


def test_npt_run_moves_the_volume():
    system = R.workload_system("ala-389w")
    sim = R.GpuSimulation(system, temperature=300.0, pressure=1.0, barostat_interval=25, seed=4)
    sim.minimise(max_iterations=150, force_tolerance=1000.0)
    sim.set_state(sim.read_positions(), R.maxwell_boltzmann(system, 300.0, 4), step=0)
    v0 = sim.energy()["volume"]
    sim.run(1000)
    e = sim.energy()
    d = sim.diagnostics()
    print("volume", v0, "->", e["volume"], "barostat", d["barostat_accepted"], "/", d["barostat_attempts"],
          "T", e["temperature"])
    assert d["barostat_attempts"] >= 39
    assert 0 < d["barostat_accepted"] < d["barostat_attempts"]
    assert e["volume"] != v0
    # a minimised lattice heats as it relaxes; finite and near the bath is the check
    assert np.isfinite(e["potential"]) and 240 < e["temperature"] < 400


def test_a_move_appended_to_its_block_equals_the_stand_alone_move():
    # rmd_run appends the trial of the move due after a block to that block's
    # launch sequence; a run that stops at the block boundary makes the same
    # move on its own at the start of the next run: every move of the two
    # runs, and so the final state, must agree bit for bit
    system = R.workload_system("ala-389w")
    x0 = v0 = None
    finals = []
    for split in (False, True):
        sim = R.GpuSimulation(system, temperature=300.0, pressure=1.0, barostat_interval=25, seed=8)
        if x0 is None:
            sim.minimise(max_iterations=30, force_tolerance=1000.0)
            x0, v0 = sim.read_positions().astype(np.float64), R.maxwell_boltzmann(system, 300.0, 8)
        sim.set_state(x0, v0, step=0)
        if split:
            for _ in range(8):
                sim.run(25, 25)  # each call's move (from the second on) runs on its own
        else:
            sim.run(200, 25)  # moves appended to the blocks before them
        d = sim.diagnostics()
        finals.append((sim.read_state(), d["barostat_attempts"], d["barostat_accepted"], sim.energy()["potential"]))
        sim.release()
    (a, na, ka, ua), (b, nb, kb, ub) = finals
    print("moves", na, ka, "/", nb, kb, "potential", ua, ub)
    assert na == nb >= 7 and ka == kb and 0 < ka < na  # both outcomes taken
    assert np.array_equal(a["positions"], b["positions"])
    assert np.array_equal(a["velocities"], b["velocities"])
    assert np.array_equal(a["box_vectors"], b["box_vectors"])
    assert ua == ub


def test_volume_move_reuses_the_block_energy_and_a_rejection_restores_the_state():
    system = R.workload_system("ala-389w")
    sim = R.GpuSimulation(system, temperature=300.0, pressure=1.0, barostat_interval=1000, seed=5)
    sim.minimise(max_iterations=30, force_tolerance=1000.0)
    sim.set_state(sim.read_positions(), R.maxwell_boltzmann(system, 300.0, 5), step=0)
    sim.run(50)
    seen = {True: 0, False: 0}
    for attempt in range(12):
        before = sim.read_state()
        fresh, _ = sim.evaluate(forces=False)
        move = sim.attempt_volume_change()
        # the old potential came from the last step of the block (or the last
        # move), not from a new evaluation, and agrees with one up to the f32
        # per-lane summation order of a list built at another step
        print(attempt, move["accepted"], move["old_potential"], fresh["potential"])
        assert rel(move["old_potential"], fresh["potential"]) < 1e-7
        if not move["accepted"]:
            after = sim.read_state()
            assert np.array_equal(after["positions"], before["positions"])
            assert np.array_equal(after["box_vectors"], before["box_vectors"])
            again, _ = sim.evaluate(forces=False)
            assert again["potential"] == fresh["potential"]
        seen[bool(move["accepted"])] += 1
        sim.run(5)
    print("moves", seen)
    assert seen[False] > 0


@pytest.mark.parametrize("scale", [1.02, 0.97])
def test_scale_volume_matches_openmm_and_oracle(scale, golden):
    # tools/gpu/check-barostat.mjs: energy at forced scales vs ala-wat-scale-*.json
    ref = golden(f"ala-wat-scale-{scale:.2f}.json")
    system, coords = fixture_system("ala-wat", dict(constraints="HBonds", rigid_water=True, cutoff=1.0))
    sim = R.GpuSimulation(system, alpha=ALPHA_24, grid=[24, 24, 24], dispersion_correction=True)
    u = sim.scale_volume(scale)
    pos = sim.read_positions().astype(np.float64)
    o = O.Oracle(system, box=system.box * scale)
    e, _ = o.potential(pos, alpha=ALPHA_24, grid=[24, 24, 24], f32_membership=True, forces=False)
    disp = o.dispersion(np.prod(np.diag(system.box.reshape(3, 3))) * scale ** 3, 1.0)
    print(scale, "gpu", u, "oracle", e["total"] + disp, "openmm", ref["energy"]["total"])
    assert rel(u, e["total"] + disp) < 1e-6
    assert rel(u, ref["energy"]["total"]) < 5e-5


def test_restraint_matches_openmm(golden):
    ref = golden("ala-wat-restraint-500.json")
    system, coords = fixture_system("ala-wat", dict(constraints="HBonds", rigid_water=True, cutoff=1.0))
    atoms = np.array(ref["restraint"]["atoms"], dtype=np.int32)
    sim = R.GpuSimulation(system, alpha=ALPHA_24, grid=[24, 24, 24], restraint_force=ref["restraint"]["force"],
                          restraint_atoms=atoms, restraint_reference=coords.positions)
    sim.set_state(np.array(ref["positions"]), None, step=0)
    e, _ = sim.evaluate(terms=R._lib.TERM_RESTRAINT, forces=False)
    theirs = ref["energy"]["byForceGroup"]["5:CustomExternalForce"]["energy"]
    print("restraint", e["position_restraint"], theirs)
    assert rel(e["position_restraint"], theirs) < 1e-6


def test_minimiser_lowers_energy_and_force():
    system = R.workload_system("ala-389w")
    sim = R.GpuSimulation(system, temperature=300.0)
    e0, f0 = sim.evaluate()
    result = sim.minimise(max_iterations=60, force_tolerance=100.0)
    print(result)
    assert result["potential"] < e0["potential"]
    assert result["max_force"] < np.linalg.norm(f0, axis=1).max()
    assert constraint_error(system, sim.read_positions()) < 1e-5


def test_neighbour_overflow_regrows_to_identical_forces():
    # tools/gpu/check-nonbonded.mjs: a forced-overflow arm must regrow to identical results
    system, coords = fixture_system("ala-wat", dict(constraints="HBonds", rigid_water=True, cutoff=1.0))
    a = R.GpuSimulation(system, alpha=ALPHA_24, grid=[24, 24, 24])
    b = R.GpuSimulation(system, alpha=ALPHA_24, grid=[24, 24, 24], max_neighbours=64)
    ea, fa = a.evaluate()
    eb, fb = b.evaluate()
    db = b.diagnostics()
    print("capacity", db["neighbour_capacity"], "growths", db["neighbour_growths"])
    assert db["neighbour_growths"] >= 1 and db["neighbour_capacity"] >= db["worst_neighbour_required"]
    assert np.array_equal(fa, fb)
    assert ea["potential"] == eb["potential"]


def test_stop_between_blocks_and_nonfinite_input_are_errors():
    system, coords = fixture_system("ala-wat", dict(constraints="HBonds", rigid_water=True, cutoff=1.0))
    sim = R.GpuSimulation(system, alpha=ALPHA_24, grid=[24, 24, 24])
    sim.request_stop()
    with pytest.raises(R.SimulationStopped):
        sim.run(50)
    bad = coords.positions.copy()
    bad[7] = np.nan
    with pytest.raises(R.NonFiniteStateError):
        sim.set_state(bad)


def _drift_fit(times_ps, totals, atoms):
    """Least-squares slope of the total energy, kJ/mol/ns per atom."""
    t = np.asarray(times_ps) / 1000.0
    e = np.asarray(totals)
    return float(np.polyfit(t, e, 1)[0] / atoms)


def test_conservative_run_drift_is_below_the_bar():
    # tools/gpu/check-integrator.mjs:130-260, the conservative arm: no
    # friction, 20 ps from the solvated fixture, the total energy's
    # least-squares slope under 5 kJ/mol/ns/atom. The direct-space force is
    # the rational fit, not the exact gradient of the erfc energy, so this
    # is also what that approximation costs in conservation.
    system, coords = fixture_system("ala-wat", dict(constraints="HBonds", rigid_water=True, cutoff=1.0))
    sim = R.GpuSimulation(system, alpha=ALPHA_24, grid=[24, 24, 24], temperature=298.0, friction=0.0, seed=7,
                          skin=0.2)
    o = O.Oracle(system)
    x0 = coords.positions.astype(np.float64)
    x0, _ = o.shake(x0, x0, 1e-12)
    v0, _ = o.rattle(x0, R.maxwell_boltzmann(system, 298.0, 3), 1e-12)
    sim.set_state(x0, v0, step=0)
    sim.initialise()
    trace = [sim.energy()]
    for _ in range(50):
        sim.run(200, 20)
        trace.append(sim.energy())
    times = [e["time"] for e in trace]
    totals = [e["total"] for e in trace]
    temps = [e["temperature"] for e in trace]
    n = system.atom_count
    windows = {f"{times[k]:.0f} ps": _drift_fit(times[:k + 1], totals[:k + 1], n) for k in (12, 25, 50)}
    spread = float(np.std(totals))
    print("drift windows (kJ/mol/ns/atom)", windows, "total spread", spread, "mean T", np.mean(temps))
    drift = _drift_fit(times, totals, n)
    assert abs(drift) < 5.0
    assert constraint_error(system, sim.read_positions()) < 2e-5
    d = sim.diagnostics()
    assert not d["shake_hit_iteration_cap"] and not d["rattle_hit_iteration_cap"]


def test_thermostat_holds_the_setpoint_over_200_ps():
    # equipartition: after 10 ps of equilibration, the mean temperature of
    # 200 ps at friction 1/ps sampled every 100 fs is the bath's within 2 K
    # (the reference measured a -0.07 K offset at 2 fs, docs/STATUS.md:48).
    # Measured over 100 ps: 300.27 K with a 1.0 K standard error from 10 ps
    # blocks, so 200 ps puts the bar near three standard errors.
    system = R.workload_system("ala-389w")
    sim = R.GpuSimulation(system, temperature=300.0, friction=1.0, seed=12)
    sim.minimise(max_iterations=300, force_tolerance=1000.0)
    sim.set_state(sim.read_positions(), R.maxwell_boltzmann(system, 300.0, 12), step=0)
    sim.run(5000, 25)
    temps = []
    for _ in range(2000):
        sim.run(50, 25)
        temps.append(sim.energy()["temperature"])
    temps = np.array(temps)
    # the standard error from blocks of 10 ps
    blocks = temps.reshape(20, -1).mean(1)
    error = blocks.std(ddof=1) / np.sqrt(len(blocks))
    print("mean T %.3f K, block standard error %.3f K, sample std %.2f K" % (temps.mean(), error, temps.std()))
    assert abs(temps.mean() - 300.0) < 2.0
    # the instantaneous fluctuation is the canonical sqrt(2/DOF) T
    dof = system.degrees_of_freedom
    assert abs(temps.std() / (300.0 * np.sqrt(2.0 / dof)) - 1) < 0.15
