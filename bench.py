#!/usr/bin/env python3
"""Throughput of the MI355X MD step on the 2854-atom NPT workload.

Workload: ACE-ALA-NME + 944 TIP3P waters (2854 atoms) on a seeded jittered
lattice in a 3.0539 nm cube, rc 0.9 nm, PME 30^3 order 5 (Ewald tolerance
5e-4), Langevin middle integrator at 300 K / 1 ps^-1, 2 fs with SETTLE + H-bond
SHAKE, NPT at 1 bar with the Monte Carlo barostat every 25 steps.

Untimed preparation: the system is minimised to its force tolerance, given
Maxwell-Boltzmann velocities at 300 K and thermalised (every block length the
timed run uses is captured as a graph here), and the barostat's phase is set
so that the timed window's first block starts with a volume move. Then W
warm-up steps, then exactly K steps timed with device synchronisation on both
sides: the window holds a move every 25 steps (one for K = 20), counted in the
JSON line with the window's temperature and density.

One JSON line on stdout: ns/day, the nonbonded kernel's FP32 roofline (its
launch time measured with HIP events on the simulation's stream, 78 FLOP per
unique in-cutoff pair), the PME stages' HBM rates, the energy and force gate
(every term and the forces of the final GPU state against FP64 at the same
f32 positions), the CPU baseline (the FP64 restatement of the reference's
cpu/ path timed on a bounded sample on one host core, on this workload and on
the 1189-atom one), and the replica batch's aggregate against the same
replicas run one system per launch.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "ns/day (single system & batched aggregate) at energy rel.err ≤1e-6 vs FP64"
TIMESTEP_PS = 0.002
FLOP_PER_PAIR = 78
FP32_PEAK_TFLOPS = 157.3
HBM_PEAK_GBPS = 8000.0  # MI355X HBM3E
COPY_PROBE_BYTES = 64 << 20  # the counter-calibration copy: 64 MiB read + 64 MiB written
REFERENCE_NS_PER_DAY = 74.0  # ala-wat-big, Apple M-series WebGPU (docs/STATUS.md:199-203)

# The single-GPU workloads: the 2854-atom NPT system the metric is quoted on
# (default), the 1189-atom NVT system and the 100,200-atom water box; and
# "replicas-N", N uneven replicas stepped as one batch.
WORKLOADS = {
    "ala-944w": dict(what="ACE-ALA-NME + 944 TIP3P", ensemble="NPT 1 bar 300 K, MC barostat every 25 steps",
                     options=dict(pressure=1.0, barostat_interval=25), reference_ns_per_day=REFERENCE_NS_PER_DAY,
                     reference_note="74 ns/day, same composition (ala-wat-big), Apple M-series WebGPU, "
                                    "docs/STATUS.md:199-203"),
    "ala-389w": dict(what="ACE-ALA-NME + 389 TIP3P", ensemble="NVT 300 K", options={}, reference_ns_per_day=154.0,
                     reference_note="154 ns/day, ala-wat (1190 atoms) NVT, Apple M-series WebGPU, docs/STATUS.md:186"),
    "water-33400": dict(what="33400 TIP3P", ensemble="NVT 300 K, skin 0.15 nm",
                        options=dict(skin=0.15, grid=[112, 112, 112])),
}


SCREEN_SEED = 505         # the learned screen's network initialisation
PLAIN_SEED_OFFSET = 7919  # the plain-barostat comparison arm's replica seeds
SAMPLE_STEPS = 250        # steps between the volume samples of the ensemble comparison


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def in_cutoff_pairs(positions, box, cutoff, excl_start, excl_partners):
    """Unique non-excluded pairs with r < rc at these positions (FP64): a
    periodic k-d tree count for rectangular boxes, minus the excluded pairs
    inside the cutoff."""
    from scipy.spatial import cKDTree

    x = positions.reshape(-1, 3).astype(np.float64)
    n = len(x)
    L = np.array([box[0], box[4], box[8]])
    wrapped = x - L * np.floor(x / L)
    wrapped[wrapped >= L] -= L[np.nonzero(wrapped >= L)[1]]
    tree = cKDTree(wrapped, boxsize=L)
    ordered = int(tree.count_neighbors(tree, cutoff))
    unique = (ordered - n) // 2
    a = np.repeat(np.arange(n), np.diff(excl_start))
    b = np.asarray(excl_partners)
    keep = a < b
    d = x[a[keep]] - x[b[keep]]
    d -= L * np.round(d / L)
    excluded_inside = int(np.count_nonzero(np.einsum("ij,ij->i", d, d) <= cutoff * cutoff))
    return unique - excluded_inside


def counted_traffic(workload, lanes):
    """HBM bytes per nonbonded launch from the committed hardware-counter
    summary (FETCH_SIZE and WRITE_SIZE scaled by a plain copy's), when it
    covers this workload and lane count; else None."""
    path = os.path.join(ROOT, "profiles", f"nonbonded_{workload.replace('-', '')}_traffic.md")
    if not os.path.exists(path):
        return None, None
    fields = dict(line.split(":", 1) for line in open(path) if line.startswith(("traffic_bytes_per_launch:",
                                                                               "lanes_per_atom:")))
    if int(fields.get("lanes_per_atom", -1)) != lanes or "traffic_bytes_per_launch" not in fields:
        return None, None
    return float(fields["traffic_bytes_per_launch"]), os.path.relpath(path, ROOT)


def cpu_baseline(system, pme, state, target_seconds, temperature=300.0, seed=202):
    """The FP64 CPU port of the reference's host path (oracle/), timed on one
    core: BAOAB steps with an all-pairs (cell-list for large boxes) direct
    space and a naive-DFT PME, for about target_seconds. Its first
    evaluation also gives the energy and force gate's FP64 side."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    o = O.Oracle(system, box=state["box"])
    cells = system.atom_count > 20000
    pos = state["positions"].astype(np.float64)
    vel = state["velocities"].astype(np.float64)
    t0 = time.perf_counter()
    ref, forces = o.potential(pos, alpha=pme["alpha"], grid=pme["grid"], f32_membership=True, use_cells=cells)
    first = time.perf_counter() - t0
    steps = max(1, int(target_seconds / max(first, 1e-3)))
    gate_forces = forces.copy()
    t0 = time.perf_counter()
    for k in range(steps):
        pos, vel, _ = o.langevin_step(pos, vel, forces, temperature=temperature, seed=seed, step=k)
        _, forces = o.potential(pos, alpha=pme["alpha"], grid=pme["grid"], use_cells=cells)
        vel, _ = o.langevin_finish(pos, vel, forces)
    elapsed = time.perf_counter() - t0
    ms = 1e3 * elapsed / steps
    return ref, gate_forces, {
        "value": TIMESTEP_PS * 1e-3 * 86400.0 / (ms * 1e-3),
        "unit": "ns/day",
        "cores": 1,
        "host_cores": os.cpu_count(),
        "kind": "port",
        "ms_per_step": ms,
        "sample": f"{steps} NVT BAOAB steps of the {system.atom_count}-atom system "
                  f"(FP64 {'cell-list' if cells else 'all-pairs'} direct space, naive-DFT PME, global SHAKE/RATTLE), "
                  "single-threaded C",
    }


def term_errors(gpu, ref):
    """Relative error of every term and the total against the FP64 oracle."""
    pairs = {
        "bonded": (gpu["bonded"], ref["bonded"]),
        "lennard_jones": (gpu["lennard_jones"], ref["lennard_jones"]),
        "coulomb_direct": (gpu["coulomb_direct"], ref["coulomb_direct_total"]),
        "coulomb_reciprocal": (gpu["coulomb_reciprocal"], ref["reciprocal"]),
        "total": (gpu["potential"] - gpu["dispersion"], ref["total"]),
    }
    return {k: abs(a - b) / max(abs(b), 1.0) for k, (a, b) in pairs.items()}


def force_error(gpu_forces, ref_forces):
    a = np.asarray(gpu_forces).reshape(-1)
    b = np.asarray(ref_forces).reshape(-1)
    return float(np.sqrt(np.mean((a - b) ** 2)) / np.sqrt(np.mean(b ** 2)))


def energy_gate(sim, system):
    """Every term and the forces of a GPU state against the FP64 oracle at
    the same f32 positions and box; returns the largest relative term error,
    the terms and the forces' relative RMS error."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    st = sim.read_state()
    diag = sim.diagnostics()
    o = O.Oracle(system, box=np.array(st["box_vectors"]))
    ref, ref_forces = o.potential(st["positions"].astype(np.float64), alpha=diag["pme_alpha"],
                                  grid=list(diag["pme_grid"]), f32_membership=True)
    gpu, forces = sim.evaluate()
    check = term_errors(gpu, ref)
    return max(check.values()), check, force_error(forces, ref_forces)


def barostat_phase_after(since, steps, block, interval):
    """steps_since_barostat after run(steps, block) from `since`, by rmd_run's
    rule: a move is attempted at a block start once `since` reaches the
    interval (context.hip, rmd_run)."""
    while steps > 0:
        if since >= interval:
            since = 0
        n = min(block, steps)
        since += n
        steps -= n
    return since


def prepare(sim, system, args, interval):
    """Untimed preparation: minimise to the force tolerance, Maxwell-Boltzmann
    velocities at 300 K, thermalise; capture the graphs of every block length
    the warm-up and the timed run use; with a barostat, run the few steps that
    put a volume move at the timed window's first block start. Returns the
    minimiser's result and the thermalisation's temperatures."""
    import rainmd_amd as R

    mini = sim.minimise(max_iterations=args.minimise_iterations, force_tolerance=args.force_tolerance)
    log(f"minimised: {mini['iterations']} iterations, potential {mini['potential']:.1f}, "
        f"max force {mini['max_force']:.1f} (tolerance {args.force_tolerance}), converged {bool(mini['converged'])}")
    sim.set_state(sim.read_positions().astype(np.float64), R.maxwell_boltzmann(system, 300.0, 202), step=0)
    temps = []
    chunk = 250
    for done in range(0, args.prep_steps, chunk):
        sim.run(min(chunk, args.prep_steps - done), args.block)
        if done >= args.prep_steps // 2:
            temps.append(sim.energy()["temperature"])
    for length in sorted({args.block, args.steps % args.block, args.warmup % args.block} - {0}):
        sim.run(length, args.block)
    if interval:
        since = sim.diagnostics()["steps_since_barostat"]
        extra = next(a for a in range(0, 2 * interval + 1)
                     if barostat_phase_after(barostat_phase_after(since, a, args.block, interval), args.warmup,
                                             args.block, interval) >= interval)
        if extra:
            sim.run(extra, args.block)
    return mini, temps


def replica_batch(args, count, warmup, steps, single_steps, screen=False):
    """`count` uneven replicas (edges U(2.2, 5.0) nm) in one batch, NPT: the
    aggregate ns/day over `steps` timed batch steps after `warmup`, the
    batched direct space's roofline, the smallest and largest replicas'
    energy and force gates, and the same replicas run one system per launch
    for `single_steps` each (each alone through its own context, graphs and
    concurrent force branches on, at the lanes per atom the autotuner picked
    for a single system of its size). Returns the result and,
    with the screen, what the plain-barostat comparison needs."""
    import rainmd_amd as R

    work = R.replica_workload(count)
    seeds = [s for _, _, s in work]
    t0 = time.perf_counter()
    systems = R.replica_systems(count)
    atoms = [s.atom_count for s in systems]
    log(f"{count} replicas, {sum(atoms)} atoms ({min(atoms)}-{max(atoms)}), built in {time.perf_counter() - t0:.1f} s")
    t0 = time.perf_counter()
    extra = dict(skin=args.replica_skin) if args.replica_skin != 0 else {}
    if args.lanes > 0:
        extra["lanes_per_atom"] = args.lanes
    batch = R.Batch(systems, seeds=seeds, temperature=300.0, pressure=1.0, barostat_interval=25, **extra)
    if args.no_graphs:
        batch.set_graphs(False)
    if args.serial_forces:
        batch.set_concurrent_forces(False)
    log(f"batch created in {time.perf_counter() - t0:.1f} s: {batch.info()}")
    t0 = time.perf_counter()
    starts = []
    for r, system in enumerate(systems):
        sim = batch.replica(r)
        sim.minimise(max_iterations=args.replica_minimise, force_tolerance=1000.0)
        starts.append((sim.read_positions().astype(np.float64), R.maxwell_boltzmann(system, 300.0, seeds[r])))
        sim.set_state(*starts[-1], step=0)
    log(f"replicas minimised in {time.perf_counter() - t0:.1f} s")
    t0 = time.perf_counter()
    batch.run(warmup, args.block)
    if steps % args.block:
        batch.run(steps % args.block, args.block)  # its block graph, outside the clock
    log(f"warm-up {warmup} steps in {time.perf_counter() - t0:.2f} s")
    collection_steps = 0
    if screen:
        # the shared learned screen, trained on the first exact moves of all replicas
        batch.configure_screen(training_moves=args.screen_training_moves, training_steps=4000, learning_rate=1e-3,
                               seed=SCREEN_SEED)
        t0 = time.perf_counter()
        while batch.screen_stats()["phase"] == 0:
            batch.run(args.block, args.block)
            collection_steps += args.block
        log(f"screen trained in {time.perf_counter() - t0:.1f} s after {collection_steps} steps: "
            f"{batch.screen_stats()}")
    # the timed run; with the screen, in chunks with every replica's volume
    # read between them (outside the clock) for the ensemble comparison
    chunk = SAMPLE_STEPS if screen else steps
    volumes = []
    before = [batch.replica(r).diagnostics() for r in range(count)]
    elapsed = 0.0
    for done in range(0, steps, chunk):
        n = min(chunk, steps - done)
        t0 = time.perf_counter()
        batch.run(n, args.block)
        elapsed += time.perf_counter() - t0
        if screen:
            volumes.append([batch.replica(r).diagnostics()["volume"] for r in range(count)])
    after = [batch.replica(r).diagnostics() for r in range(count)]
    ms_per_step = 1e3 * elapsed / steps
    aggregate = count * TIMESTEP_PS * 1e-3 * 86400.0 / (ms_per_step * 1e-3)
    rebuilds = sum(a["rebuilds"] - b["rebuilds"] for a, b in zip(after, before))
    moves = sum(a["barostat_attempts"] - b["barostat_attempts"] for a, b in zip(after, before))
    accepted = sum(a["barostat_accepted"] - b["barostat_accepted"] for a, b in zip(after, before))
    log(f"timed {steps} batch steps: {ms_per_step:.3f} ms/step, aggregate {aggregate:.1f} ns/day; "
        f"{rebuilds} list rebuilds, one per {count * steps / max(1, rebuilds):.1f} replica steps; "
        f"barostat {accepted}/{moves} in the window")
    temps = [batch.replica(r).energy()["temperature"] for r in range(count)]

    # the batched direct-space launch at the final state
    nb_ms = batch.profile_nonbonded(reps=20)
    pairs = 0
    for r, system in enumerate(systems):
        st = batch.replica(r).read_state()
        pairs += in_cutoff_pairs(st["positions"], np.array(st["box_vectors"]), system.cutoff,
                                 system.exclusion_start, system.exclusion_partners)
    achieved = FLOP_PER_PAIR * pairs / (nb_ms * 1e-3) / 1e12
    log(f"batched nonbonded {nb_ms * 1e3:.1f} us over {pairs} pairs -> {achieved:.2f} TFLOP/s")
    small, large = int(np.argmin(atoms)), int(np.argmax(atoms))
    gates = {}
    if not args.no_cpu_baseline:
        for r in (small, large):
            gates[r] = energy_gate(batch.replica(r), systems[r])
            log(f"energy gate replica {r} ({atoms[r]} atoms): max rel {gates[r][0]:.2e}, forces {gates[r][2]:.2e}")

    # one system per launch: every replica alone through its own context, the
    # same kernels and options, at the lanes per atom tuned for a single system
    # of its size
    single_total = 0.0
    for r in range(count if single_steps > 0 else 0):
        sim = batch.replica(r)
        lanes = sim.diagnostics()["lanes_per_atom"]
        sim.set_lanes(R.tuned_shape("single", atoms[r])["lanes_per_atom"])
        # captures its block graphs: with a volume move's trial appended (the
        # first block, a move being due after it) and without (the last)
        sim.run(2 * args.block, args.block)
        if single_steps % args.block:
            sim.run(single_steps % args.block, args.block)
        t0 = time.perf_counter()
        sim.run(single_steps, args.block)
        single_total += time.perf_counter() - t0
        sim.set_lanes(lanes)
    single_ms = 1e3 * single_total / single_steps if single_steps > 0 else None
    single_aggregate = count * TIMESTEP_PS * 1e-3 * 86400.0 / (single_ms * 1e-3) if single_ms else None
    if single_ms:
        log(f"one system per launch: {single_ms:.3f} ms per step of all replicas, aggregate {single_aggregate:.1f} "
            f"ns/day; batch speedup {aggregate / single_aggregate:.2f}x")
    result = {
        "value": aggregate,
        "unit": "ns/day",
        "steps": steps,
        "warmup": warmup,
        "ms_per_step": ms_per_step,
        "config": {
            "workload": f"replicas-{count}: {count} replicas of ACE-ALA-NME + TIP3P at 1.000 g/cm^3, cube edges "
                        f"U(2.2, 5.0) nm (seed 404), {min(atoms)}-{max(atoms)} atoms, {sum(atoms)} in all, NPT 1 bar "
                        f"300 K, MC barostat every 25 steps per replica, 2 fs, rc 0.9 nm, PME grid per replica",
            "replicas": count,
            "atoms": sum(atoms),
            "block": args.block,
            "skin_nm": batch.replica(0).diagnostics()["skin"],
            "lanes_per_atom": batch.replica(0).diagnostics()["lanes_per_atom"],
            "force_branches": "serial" if args.serial_forces else "concurrent",
            "replica_steps_per_rebuild": count * steps / max(1, rebuilds),
        },
        "single_launch_baseline": {
            "value": single_aggregate,
            "unit": "ns/day",
            "ms_per_step": single_ms,
            "sample": f"every replica run alone for {single_steps} steps through its own context "
                      "(the batch's options; graphs, concurrent force branches and the single-system tuned lanes "
                      "per atom for its size)",
        } if single_ms else None,
        "batch_speedup": aggregate / single_aggregate if single_ms else None,
        "barostat_in_window": {"attempts": moves, "accepted": accepted},
        "temperature_mean_K": float(np.mean(temps)),
        "roofline": {
            "kernel": "batched nonbonded (LJ + screened Coulomb real space, every replica in one launch)",
            "bound": "fp32_valu",
            "achieved": achieved,
            "peak": FP32_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / FP32_PEAK_TFLOPS,
            "traffic": None,
            "launch_us": nb_ms * 1e3,
            "pairs": pairs,
            "flop_per_pair": FLOP_PER_PAIR,
            "note": "FP32 vector ALU bound; peak is MI355X's published FP32 figure",
        },
    }
    if gates:
        result["energy_check"] = {"max_rel_err": max(g[0] for g in gates.values()),
                                  "forces_rel_rms": max(g[2] for g in gates.values()),
                                  "replicas": {str(r): dict(g[1], forces_rel_rms=g[2]) for r, g in gates.items()}}
    if screen:
        st = batch.screen_stats()
        st["skipped_fraction"] = 1.0 - st["exact_evaluations"] / max(1, st["moves"])
        st["network_seed"] = SCREEN_SEED
        result["screen"] = st
    batch.release()
    return result, (systems, starts, extra, collection_steps, np.array(volumes))


def bench_replicas(args, count):
    """The replica workload on its own: one JSON line whose value is the
    batch's aggregate ns/day."""
    result, (systems, starts, extra, collection_steps, volumes) = replica_batch(
        args, count, args.warmup, args.steps, args.single_steps, screen=args.screen)
    if args.screen:
        result["screen"]["ensemble"] = plain_barostat_comparison(args, systems, starts, extra, collection_steps,
                                                                 volumes)
    line = {"metric": METRIC, "value": result.pop("value"), "unit": result.pop("unit"), "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup, "ms_per_step": result.pop("ms_per_step"),
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic"}
    result.pop("steps"), result.pop("warmup")
    line.update(result)
    print(json.dumps(line), flush=True)


def plain_barostat_comparison(args, systems, starts, extra, collection_steps, screened):
    """The same replicas from the same minimised states under the plain
    barostat (independent seeds), stepped through the same protocol - warm-up,
    the screen's collection steps, then the timed steps sampled at the same
    points - and the two arms' densities compared: a KS test on the pooled
    densities thinned to one sample per correlation time, the replicas' paired
    mean densities, and the compressibilities pooled over the replicas."""
    import rainmd_amd as R
    from rainmd_amd import npt

    count = len(systems)
    seeds = [s + PLAIN_SEED_OFFSET for _, _, s in R.replica_workload(count)]
    plain = R.Batch(systems, seeds=seeds, temperature=300.0, pressure=1.0, barostat_interval=25, **extra)
    if args.no_graphs:
        plain.set_graphs(False)
    if args.serial_forces:
        plain.set_concurrent_forces(False)
    t0 = time.perf_counter()
    for r, (x, v) in enumerate(starts):
        plain.replica(r).set_state(x, v, step=0)
    plain.run(args.warmup + collection_steps, args.block)
    volumes = []
    for done in range(0, args.steps, SAMPLE_STEPS):
        plain.run(min(SAMPLE_STEPS, args.steps - done), args.block)
        volumes.append([plain.replica(r).diagnostics()["volume"] for r in range(count)])
    moves = sum(plain.replica(r).diagnostics()["barostat_attempts"] for r in range(count))
    plain.release()
    log(f"plain barostat arm: {args.warmup + collection_steps + args.steps} steps in {time.perf_counter() - t0:.1f} s")
    volumes = np.array(volumes)
    masses = np.array([float(s.masses.sum()) for s in systems])
    rho_plain = npt.density(volumes, 1.0) * masses
    rho_screened = npt.density(screened, 1.0) * masses
    taus = [max(npt.autocorrelation_time(rho_plain[:, r]), npt.autocorrelation_time(rho_screened[:, r]))
            for r in range(count)]
    tau = float(np.median(taus))
    d, p = npt.ks_two_sample(npt.thin(rho_plain, tau).ravel(), npt.thin(rho_screened, tau).ravel())
    diff = rho_screened.mean(0) - rho_plain.mean(0)
    sigma = float(abs(diff.mean()) / (diff.std(ddof=1) / np.sqrt(count)))
    kappa = (sum(npt.compressibility(screened[:, r], 300.0) for r in range(count))
             / sum(npt.compressibility(volumes[:, r], 300.0) for r in range(count)))
    out = {"samples_per_replica": int(volumes.shape[0]), "sample_interval_steps": SAMPLE_STEPS,
           "correlation_time_samples": tau, "ks_statistic": d, "ks_p": p,
           "density_plain": float(rho_plain.mean()), "density_screened": float(rho_screened.mean()),
           "paired_density_difference": float(diff.mean()), "paired_difference_sigma": sigma,
           "pooled_compressibility_ratio": float(kappa), "plain_moves": int(moves),
           "plain_seed_offset": PLAIN_SEED_OFFSET}
    log(f"ensemble: {out}")
    return out


def pme_stage_rates(stages_ms, atoms, grid):
    """HBM rates of the PME stages against their algorithmic bytes: spread
    16N + 4K^3 (positions and charges in, the grid out), the transforms and
    convolution 4K^3 + 4K^3 (the grid in and out once, fused in LDS), the
    gather 16N + 4K^3 + 12N."""
    k3 = int(np.prod(grid))
    algorithmic = {"spread": 16 * atoms + 4 * k3, "transform": 8 * k3, "gather": 28 * atoms + 4 * k3}
    out = {}
    for stage, ms in stages_ms.items():
        if ms == 0.0:  # folded into the gathers (force-only chains of small systems)
            out[stage] = {"us": 0.0, "folded_into": "gather"}
            continue
        gbps = algorithmic[stage] / (ms * 1e-3) / 1e9
        out[stage] = {"us": ms * 1e3, "algorithmic_bytes": algorithmic[stage], "GBps": gbps,
                      "frac_of_hbm": gbps / HBM_PEAK_GBPS}
    return out


def c1_cpu_baseline(seconds):
    """The CPU path on the 1189-atom NVT workload (the configuration the
    reference runs on its CPU path): the oracle from the workload's minimised
    lattice with Maxwell-Boltzmann velocities, a bounded sample of its 1000
    steps."""
    import rainmd_amd as R

    system = R.workload_system("ala-389w")
    sim = R.GpuSimulation(system, temperature=300.0, seed=101)
    sim.minimise(max_iterations=500, force_tolerance=1000.0)
    diag = sim.diagnostics()
    state = {"positions": sim.read_positions(), "velocities": R.maxwell_boltzmann(system, 300.0, 101),
             "box": np.array(system.box)}
    sim.release()
    _, _, baseline = cpu_baseline(system, {"alpha": diag["pme_alpha"], "grid": list(diag["pme_grid"])}, state,
                                  seconds, seed=101)
    baseline["sample"] += " (the 1189-atom NVT workload; 1000 steps in its configuration)"
    return baseline


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5000)
    ap.add_argument("--warmup", type=int, default=1000)
    ap.add_argument("--workload", default="ala-944w")
    ap.add_argument("--skin", type=float, default=None,
                    help="nm; 0 derives the reference's skinFor, < 0 takes the autotuner's skin; default: the "
                         "workload's configured skin, else the autotuner's")
    ap.add_argument("--lanes", type=int, default=0)
    ap.add_argument("--block", type=int, default=25)
    ap.add_argument("--prep-steps", type=int, default=5000,
                    help="untimed thermalisation after minimising, before the warm-up")
    ap.add_argument("--minimise-iterations", type=int, default=2000)
    ap.add_argument("--force-tolerance", type=float, default=2000.0,
                    help="kJ/mol/nm, the minimiser's target (the reference's own device minimiser stalls near 2100 on "
                         "the solvated fixture, webgpu/simulation.js:861-870; the thermalisation follows)")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--c1-cpu-seconds", type=float, default=10.0,
                    help="the CPU baseline's sample on the 1189-atom workload (0 skips it)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-graphs", action="store_true", help="launch kernels directly (for profilers)")
    ap.add_argument("--kernel-only", default=None,
                    help="after warm-up, launch only this kernel family (nonbonded, pme, integrate, bonded, rebuild) "
                         "and the calibration copy, for hardware-counter passes; prints their durations and exits")
    ap.add_argument("--reps", type=int, default=50)
    ap.add_argument("--batch-replicas", type=int, default=256,
                    help="the replica batch measured beside the single system (0 skips it)")
    ap.add_argument("--batch-warmup", type=int, default=1000)
    ap.add_argument("--batch-steps", type=int, default=200)
    ap.add_argument("--batch-single-steps", type=int, default=50)
    ap.add_argument("--single-steps", type=int, default=100,
                    help="replica workloads: steps of each replica alone for the one-system-per-launch baseline")
    ap.add_argument("--replica-minimise", type=int, default=150)
    ap.add_argument("--serial-forces", action="store_true",
                    help="run the force branches one after another (for per-kernel profiles)")
    ap.add_argument("--replica-skin", type=float, default=-1.0,
                    help="replica workloads: nm; 0 derives skinFor, < 0 takes the autotuner's skin")
    ap.add_argument("--screen", action="store_true", help="replica workloads: the shared learned barostat screen")
    ap.add_argument("--screen-training-moves", type=int, default=2000)
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build runs one system per MI355X and has no multi-GPU "
              f"path (asked for --gpus {args.gpus}).", file=sys.stderr)
        sys.exit(2)

    if args.workload.startswith("replicas-"):
        bench_replicas(args, int(args.workload.split("-")[1]))
        return

    import rainmd_amd as R

    system = R.workload_system(args.workload)
    spec = WORKLOADS[args.workload]
    opts = dict(temperature=300.0, seed=202, **spec["options"])
    if args.skin is not None:
        opts["skin"] = args.skin
    elif "skin" not in opts:
        opts["skin"] = -1.0
    if args.lanes > 0:
        opts["lanes_per_atom"] = args.lanes
    sim = R.GpuSimulation(system, **opts)
    if args.no_graphs:
        sim.set_graphs(False)
    diag = sim.diagnostics()
    pme = {"alpha": diag["pme_alpha"], "grid": diag["pme_grid"]}
    log(f"workload {args.workload}: {system.atom_count} atoms, grid {pme['grid']}, skin {diag['skin']:.3f}")
    interval = spec["options"].get("barostat_interval", 0) if "pressure" in spec["options"] else 0
    t0 = time.perf_counter()
    mini, prep_temps = prepare(sim, system, args, interval)
    log(f"prepared in {time.perf_counter() - t0:.2f} s: thermalised {args.prep_steps} steps, T over the second half "
        f"{np.mean(prep_temps) if prep_temps else float('nan'):.1f} K; steps since the last move "
        f"{sim.diagnostics()['steps_since_barostat']}")
    t0 = time.perf_counter()
    sim.run(args.warmup, args.block)
    log(f"warm-up {args.warmup} steps in {time.perf_counter() - t0:.2f} s")

    if args.kernel_only:
        ms = sim.profile_kernel(args.kernel_only, reps=args.reps)
        copy_bytes = COPY_PROBE_BYTES
        copy_ms = sim.copy_probe(copy_bytes, reps=args.reps)
        print(json.dumps({"kernel": args.kernel_only, "reps": args.reps, "launch_us": ms * 1e3,
                          "copy_bytes": copy_bytes, "copy_us": copy_ms * 1e3,
                          "copy_GBps": 2 * copy_bytes / (copy_ms * 1e-3) / 1e9}), flush=True)
        sim.release()
        return

    before = sim.diagnostics()
    t0 = time.perf_counter()
    sim.run(args.steps, args.block)
    elapsed = time.perf_counter() - t0
    ms_per_step = 1e3 * elapsed / args.steps
    ns_day = TIMESTEP_PS * 1e-3 * 86400.0 / (ms_per_step * 1e-3)
    energy = sim.energy()
    diag = sim.diagnostics()
    moves = diag["barostat_attempts"] - before["barostat_attempts"]
    accepted = diag["barostat_accepted"] - before["barostat_accepted"]
    density = float(system.masses.sum()) / (energy["volume"] * 602.214076)
    log(f"timed {args.steps} steps: {ms_per_step * 1e3:.1f} us/step, {ns_day:.1f} ns/day; "
        f"T {energy['temperature']:.1f} K, volume {energy['volume']:.4f} nm^3 ({density:.4f} g/cm^3), "
        f"barostat {accepted}/{moves} in the window, rebuilds {diag['rebuilds'] - before['rebuilds']} in the window")

    # Roofline of the dominant kernel at the final state.
    state = sim.read_state()
    state["box"] = np.array(energy["box"])
    nb_ms = sim.profile_kernel("nonbonded", reps=200)
    pairs = in_cutoff_pairs(state["positions"], state["box"], system.cutoff, system.exclusion_start,
                            system.exclusion_partners)
    achieved = FLOP_PER_PAIR * pairs / (nb_ms * 1e-3) / 1e12
    traffic, traffic_source = counted_traffic(args.workload, diag["lanes_per_atom"])
    kernels = {k: sim.profile_kernel(k, reps=100) for k in ("pme", "integrate", "bonded", "rebuild")}
    stages = pme_stage_rates(sim.profile_pme_stages(reps=100), system.atom_count, pme["grid"])
    log(f"nonbonded {nb_ms * 1e3:.2f} us/launch over {pairs} pairs -> {achieved:.3f} TFLOP/s; "
        + ", ".join(f"{k} {v * 1e3:.2f} us" for k, v in kernels.items())
        + "; PME stages " + ", ".join(f"{k} {v['us']:.1f} us ({v['GBps']:.0f} GB/s)" if "GBps" in v
                                      else f"{k} folded into the {v['folded_into']}" for k, v in stages.items()))

    result = {
        "metric": METRIC,
        "value": ns_day,
        "unit": "ns/day",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": ns_day / spec["reference_ns_per_day"] if spec.get("reference_ns_per_day") else None,
        "baseline_reference": spec.get("reference_note"),
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"{args.workload}: {spec['what']}, {system.atom_count} atoms, "
                        f"{system.box[0]:.4f} nm cube, {spec['ensemble']}, 2 fs, rc {system.cutoff} nm, "
                        f"PME {pme['grid'][0]}^3 order 5",
            "atoms": system.atom_count,
            "skin_nm": diag["skin"],
            "lanes_per_atom": diag["lanes_per_atom"],
            "cells": list(diag["cell_counts"]),
            "cell_reach": diag["cell_reach"],
            "block": args.block,
            "graphs": not args.no_graphs,
        },
        "window": {
            "barostat_attempts": moves,
            "barostat_accepted": accepted,
            "rebuilds": diag["rebuilds"] - before["rebuilds"],
            "temperature_end_K": energy["temperature"],
            "temperature_mean_K": float(np.mean(prep_temps)) if prep_temps else None,
            "temperature_mean_note": f"sampled every 250 steps over the last {args.prep_steps // 2} steps of the "
                                     "untimed thermalisation before the window",
            "volume_nm3": energy["volume"],
            "density_g_cm3": density,
        },
        "preparation": {
            "minimiser_iterations": mini["iterations"],
            "minimiser_max_force": mini["max_force"],
            "minimiser_converged": bool(mini["converged"]),
            "force_tolerance": args.force_tolerance,
            "thermalisation_steps": args.prep_steps,
        },
        "roofline": {
            "kernel": "nonbonded (LJ + screened Coulomb real space)",
            "bound": "fp32_valu",
            "achieved": achieved,
            "peak": FP32_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / FP32_PEAK_TFLOPS,
            "traffic": traffic,
            "traffic_source": traffic_source,
            "launch_us": nb_ms * 1e3,
            "pairs": pairs,
            "flop_per_pair": FLOP_PER_PAIR,
            "note": "FP32 vector ALU bound; peak is MI355X's published FP32 figure",
        },
        "kernels_us": {k: v * 1e3 for k, v in kernels.items()},
        "pme_stages": stages,
    }
    if not args.no_cpu_baseline:
        gpu_terms, gpu_forces = sim.evaluate()
        ref, ref_forces, baseline = cpu_baseline(system, pme, state, args.cpu_seconds)
        check = term_errors(gpu_terms, ref)
        result["energy_check"] = {"max_rel_err": max(check.values()), "terms": check,
                                  "forces_rel_rms": force_error(gpu_forces, ref_forces)}
        result["cpu_baseline"] = baseline
        log(f"energy check vs FP64: max rel {result['energy_check']['max_rel_err']:.2e}, forces rel rms "
            f"{result['energy_check']['forces_rel_rms']:.2e}; cpu baseline {baseline['ms_per_step']:.1f} ms/step")
    sim.release()
    if not args.no_cpu_baseline and args.c1_cpu_seconds > 0 and args.workload == "ala-944w":
        result["cpu_baseline_1189_atoms"] = c1_cpu_baseline(args.c1_cpu_seconds)
        log(f"cpu baseline, 1189 atoms: {result['cpu_baseline_1189_atoms']['ms_per_step']:.1f} ms/step")
    if args.batch_replicas > 0 and args.workload == "ala-944w":
        batched, _ = replica_batch(args, args.batch_replicas, args.batch_warmup, args.batch_steps,
                                   args.batch_single_steps)
        result["batched"] = batched
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
