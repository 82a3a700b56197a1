"""Replica batches on the GPU (rmd_batch_*): uneven replicas of the replica
workload stepped by batched launches. Each replica computes exactly what its
own single-system context computes, so positions, velocities and energies
after a batch run equal those of separate single runs bit for bit (NVT and
NPT, through list growth and with graphs on or off); one replica is also
held to the FP64 oracle."""
import numpy as np
import pytest

import oracle as O
import rainmd_amd as R
from helpers import rel

pytestmark = pytest.mark.gpu


def _systems(picks):
    work = R.replica_workload(32)
    chosen = [work[k] for k in picks]
    print("replicas", [(round(e, 3), 22 + 3 * w, s) for e, w, s in chosen])
    return [R.System.solvated(water_count=w, box_edge=e, seed=s, include_solute=True, cutoff=0.9)
            for e, w, s in chosen], [s for _, _, s in chosen]


# small and mid-sized replicas of the workload (1.1k - 10.6k atoms)
PICKS = [1, 2, 5, 9]
# the shapes the batch and the single runs compared share: the lanes per atom
# and the cell plan's reach (the tuned defaults of a batch member and of a
# single system may differ, and the cell plan orders the direct space's sums)
LANES = 16
SHAPE = dict(lanes_per_atom=LANES, cell_reach=3)


def _start_states(systems, seeds):
    """Minimised positions and Maxwell-Boltzmann velocities for each system,
    from a single context (the minimiser is not what is compared)."""
    out = []
    for sys_, seed in zip(systems, seeds):
        sim = R.GpuSimulation(sys_, temperature=300.0, seed=seed)
        sim.minimise(max_iterations=60, force_tolerance=1000.0)
        out.append((sim.read_positions().astype(np.float64), R.maxwell_boltzmann(sys_, 300.0, seed)))
        sim.release()
    return out


def _run_pair(options, steps, graphs=True, max_neighbours=None):
    systems, seeds = _systems(PICKS)
    starts = _start_states(systems, seeds)
    extra = dict(options)
    if max_neighbours:
        extra["max_neighbours"] = max_neighbours
    batch = R.Batch(systems, seeds=seeds, **SHAPE, **extra)
    batch.set_graphs(graphs)
    for r, (x, v) in enumerate(starts):
        batch.replica(r).set_state(x, v, step=0)
    batch.run(steps)
    singles = []
    for r, (sys_, seed) in enumerate(zip(systems, seeds)):
        sim = R.GpuSimulation(sys_, seed=seed, **SHAPE, **extra)
        # the batch members' PME x lines (a small single system would
        # otherwise fold them into its plane gathers)
        sim.set_pme_pipeline("planes")
        sim.set_graphs(graphs)
        sim.set_state(*starts[r], step=0)
        sim.run(steps)
        singles.append(sim)
    return batch, singles, systems


def _compare(batch, singles):
    for r, sim in enumerate(singles):
        b = batch.replica(r)
        sb, ss = b.read_state(), sim.read_state()
        assert np.array_equal(sb["positions"], ss["positions"]), r
        assert np.array_equal(sb["velocities"], ss["velocities"]), r
        assert np.array_equal(sb["box_vectors"], ss["box_vectors"]), r
        eb, es = b.energy(), sim.energy()
        assert eb["potential"] == es["potential"] and eb["kinetic"] == es["kinetic"], r
        db, ds = b.diagnostics(), sim.diagnostics()
        assert db["barostat_attempts"] == ds["barostat_attempts"], r
        assert db["barostat_accepted"] == ds["barostat_accepted"], r
        print("replica", r, "potential", eb["potential"], "volume", db["volume"],
              "moves", db["barostat_accepted"], "/", db["barostat_attempts"], "rebuilds", db["rebuilds"], ds["rebuilds"])


def test_batch_nvt_equals_single_runs():
    batch, singles, _ = _run_pair(dict(temperature=300.0), 100)
    _compare(batch, singles)
    info = batch.info()
    print(info)
    assert info["replicas"] == len(PICKS) and info["blocks_run"] == 4
    for s in singles:
        s.release()
    batch.release()


def test_batch_npt_equals_single_runs():
    batch, singles, _ = _run_pair(dict(temperature=300.0, pressure=1.0, barostat_interval=10), 120)
    _compare(batch, singles)
    attempts = [batch.replica(r).diagnostics()["barostat_attempts"] for r in range(len(PICKS))]
    assert attempts == [11] * len(PICKS)
    for s in singles:
        s.release()
    batch.release()


def test_batch_without_graphs_and_with_list_growth_equals_single_runs():
    batch, singles, _ = _run_pair(dict(temperature=300.0, pressure=1.0, barostat_interval=25), 50, graphs=False,
                                  max_neighbours=200)
    _compare(batch, singles)
    grown = [batch.replica(r).diagnostics()["neighbour_capacity"] for r in range(len(PICKS))]
    print("capacities", grown)
    assert all(c > 200 for c in grown)
    for s in singles:
        s.release()
    batch.release()


def test_batch_replica_energies_match_the_oracle():
    systems, seeds = _systems([3, 14])
    starts = _start_states(systems, seeds)
    batch = R.Batch(systems, seeds=seeds, temperature=300.0, pressure=1.0)
    for r, (x, v) in enumerate(starts):
        batch.replica(r).set_state(x, v, step=0)
    batch.run(50)
    for r, sys_ in enumerate(systems):
        sim = batch.replica(r)
        st = sim.read_state()
        pos = st["positions"].astype(np.float64)
        e = sim.energy()
        diag = sim.diagnostics()
        oracle = O.Oracle(sys_)
        oracle.set_box(np.array(st["box_vectors"]).reshape(9))
        ref, _ = oracle.potential(pos, alpha=diag["pme_alpha"], grid=list(diag["pme_grid"]), f32_membership=True)
        pairs = {"lennard_jones": ref["lennard_jones"], "coulomb_direct": ref["coulomb_direct_total"],
                 "coulomb_reciprocal": ref["reciprocal"], "bonded": ref["bonded"]}
        for key, value in pairs.items():
            print(r, key, e[key], value, rel(e[key], value))
            assert rel(e[key], value) < 1e-6, (r, key)
    batch.release()


def test_batch_refuses_mixed_cutoffs():
    a = R.System.solvated(water_count=389, box_edge=2.2816, seed=101, cutoff=0.9)
    b = R.System.solvated(water_count=389, box_edge=2.2816, seed=102, cutoff=1.0)
    with pytest.raises(R.RainmdError, match="same Ewald alpha"):
        R.Batch([a, b])
