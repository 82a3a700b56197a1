#!/usr/bin/env python3
"""Throughput of the MI355X MD step on the 2854-atom NPT workload.

Workload: ACE-ALA-NME + 944 TIP3P waters (2854 atoms) on a seeded jittered
lattice in a 3.0539 nm cube, rc 0.9 nm, PME 30^3 order 5 (Ewald tolerance
5e-4), Langevin middle integrator at 300 K / 1 ps^-1, 2 fs with SETTLE + H-bond
SHAKE, NPT at 1 bar with the Monte Carlo barostat every 25 steps. The system
is minimised, given Maxwell-Boltzmann velocities and run W warm-up steps;
then exactly K steps are timed (barostat moves included) with device
synchronisation on both sides.

One JSON line on stdout: ns/day, the nonbonded kernel's FP32 roofline (its
launch time measured with HIP events on the simulation's stream, 78 FLOP per
unique in-cutoff pair), and the CPU baseline: the FP64 restatement of the
reference's cpu/ path timed on a bounded sample on one host core, whose first
evaluation is also the energy gate (every term of the final GPU state against
FP64 at the same f32 positions).
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
    if not os.path.exists(path) or lanes != 64:
        return None, None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            return float(line.split(":")[1]), os.path.relpath(path, ROOT)
    return None, None


def cpu_baseline(system, pme, state, target_seconds):
    """The FP64 CPU port of the reference's host path (oracle/), timed on one
    core: BAOAB steps with an all-pairs direct space and a naive-DFT PME. Its
    first evaluation also checks the GPU's energies term by term."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    o = O.Oracle(system, box=state["box"])
    pos = state["positions"].astype(np.float64)
    vel = state["velocities"].astype(np.float64)
    t0 = time.perf_counter()
    ref, forces = o.potential(pos, alpha=pme["alpha"], grid=pme["grid"], f32_membership=True)
    first = time.perf_counter() - t0
    steps = max(1, int(target_seconds / max(first, 1e-3)))
    t0 = time.perf_counter()
    for k in range(steps):
        pos, vel, _ = o.langevin_step(pos, vel, forces, temperature=300.0, seed=202, step=k)
        _, forces = o.potential(pos, alpha=pme["alpha"], grid=pme["grid"])
        vel, _ = o.langevin_finish(pos, vel, forces)
    elapsed = time.perf_counter() - t0
    ms = 1e3 * elapsed / steps
    return ref, {
        "value": TIMESTEP_PS * 1e-3 * 86400.0 / (ms * 1e-3),
        "unit": "ns/day",
        "cores": 1,
        "kind": "port",
        "ms_per_step": ms,
        "sample": f"{steps} NVT BAOAB steps of the same {system.atom_count}-atom system from the final GPU state "
                  "(FP64 all-pairs direct space, naive-DFT PME, global SHAKE/RATTLE), single-threaded C",
    }


def energy_gate(sim, system):
    """Every term of a GPU state against the FP64 oracle at the same f32
    positions and box; returns the largest relative error and the terms."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    st = sim.read_state()
    diag = sim.diagnostics()
    o = O.Oracle(system, box=np.array(st["box_vectors"]))
    ref, _ = o.potential(st["positions"].astype(np.float64), alpha=diag["pme_alpha"], grid=list(diag["pme_grid"]),
                         f32_membership=True, forces=False)
    gpu, _ = sim.evaluate(forces=False)
    pairs = {
        "bonded": (gpu["bonded"], ref["bonded"]),
        "lennard_jones": (gpu["lennard_jones"], ref["lennard_jones"]),
        "coulomb_direct": (gpu["coulomb_direct"], ref["coulomb_direct_total"]),
        "coulomb_reciprocal": (gpu["coulomb_reciprocal"], ref["reciprocal"]),
        "total": (gpu["potential"] - gpu["dispersion"], ref["total"]),
    }
    check = {k: abs(a - b) / max(abs(b), 1.0) for k, (a, b) in pairs.items()}
    return max(check.values()), check


def bench_replicas(args, count):
    """The replica workload: `count` uneven replicas (edges U(2.2, 5.0) nm) in
    one batch, NPT. The aggregate is count x 2 fs per batch step; the
    baseline is the same replicas run one system per launch (each alone
    through its own context, graphs and concurrent force branches on)."""
    import rainmd_amd as R

    work = R.replica_workload(count)
    seeds = [s for _, _, s in work]
    t0 = time.perf_counter()
    systems = R.replica_systems(count)
    atoms = [s.atom_count for s in systems]
    log(f"{count} replicas, {sum(atoms)} atoms ({min(atoms)}-{max(atoms)}), built in {time.perf_counter() - t0:.1f} s")
    t0 = time.perf_counter()
    extra = dict(skin=args.replica_skin) if args.replica_skin > 0 else {}
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
    batch.run(args.warmup, args.block)
    log(f"warm-up {args.warmup} steps in {time.perf_counter() - t0:.2f} s")
    collection_steps = 0
    if args.screen:
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
    chunk = SAMPLE_STEPS if args.screen else args.steps
    volumes = []
    rebuilds = sum(batch.replica(r).diagnostics()["rebuilds"] for r in range(count))
    elapsed = 0.0
    for done in range(0, args.steps, chunk):
        steps = min(chunk, args.steps - done)
        t0 = time.perf_counter()
        batch.run(steps, args.block)
        elapsed += time.perf_counter() - t0
        if args.screen:
            volumes.append([batch.replica(r).diagnostics()["volume"] for r in range(count)])
    ms_per_step = 1e3 * elapsed / args.steps
    aggregate = count * TIMESTEP_PS * 1e-3 * 86400.0 / (ms_per_step * 1e-3)
    rebuilds = sum(batch.replica(r).diagnostics()["rebuilds"] for r in range(count)) - rebuilds
    log(f"timed {args.steps} batch steps: {ms_per_step:.3f} ms/step, aggregate {aggregate:.1f} ns/day; "
        f"{rebuilds} list rebuilds, one per {count * args.steps / max(1, rebuilds):.1f} replica steps")

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
            log(f"energy gate replica {r} ({atoms[r]} atoms): max rel {gates[r][0]:.2e}")
    barostat = [batch.replica(r).diagnostics() for r in range(count)]
    moves = sum(d["barostat_attempts"] for d in barostat)
    accepted = sum(d["barostat_accepted"] for d in barostat)

    # one system per launch: every replica alone through its own context, the
    # same kernels and options, at the single system's 64 lanes per atom
    single_total = 0.0
    for r in range(count if args.single_steps > 0 else 0):
        sim = batch.replica(r)
        lanes = sim.diagnostics()["lanes_per_atom"]
        sim.set_lanes(64)
        sim.run(args.block, args.block)  # captures its block graph
        t0 = time.perf_counter()
        sim.run(args.single_steps, args.block)
        single_total += time.perf_counter() - t0
        sim.set_lanes(lanes)
    single_ms = 1e3 * single_total / args.single_steps if args.single_steps > 0 else None
    single_aggregate = count * TIMESTEP_PS * 1e-3 * 86400.0 / (single_ms * 1e-3) if single_ms else None
    if single_ms:
        log(f"one system per launch: {single_ms:.3f} ms per step of all replicas, aggregate {single_aggregate:.1f} "
            f"ns/day; batch speedup {aggregate / single_aggregate:.2f}x")
    result = {
        "metric": METRIC,
        "value": aggregate,
        "unit": "ns/day",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
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
            "replica_steps_per_rebuild": count * args.steps / max(1, rebuilds),
        },
        "single_launch_baseline": {
            "value": single_aggregate,
            "unit": "ns/day",
            "ms_per_step": single_ms,
            "sample": f"every replica run alone for {args.single_steps} steps through its own context "
                      "(the batch's options; graphs, concurrent force branches and 64 lanes per atom)",
        } if single_ms else None,
        "batch_speedup": aggregate / single_aggregate if single_ms else None,
        "barostat": {"attempts": moves, "accepted": accepted},
        "roofline": {
            "kernel": "batched nonbonded (LJ + screened Coulomb real space, every replica in one launch)",
            "bound": "mfma",
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
                                  "replicas": {str(r): g[1] for r, g in gates.items()}}
    if args.screen:
        st = batch.screen_stats()
        st["skipped_fraction"] = 1.0 - st["exact_evaluations"] / max(1, st["moves"])
        st["network_seed"] = SCREEN_SEED
        result["screen"] = st
    batch.release()
    if args.screen:
        result["screen"]["ensemble"] = plain_barostat_comparison(args, systems, starts, extra, collection_steps,
                                                                 np.array(volumes))
    print(json.dumps(result), flush=True)


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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5000)
    ap.add_argument("--warmup", type=int, default=1000)
    ap.add_argument("--workload", default="ala-944w")
    ap.add_argument("--skin", type=float, default=0.0, help="nm; 0 derives the reference's skinFor")
    ap.add_argument("--lanes", type=int, default=0)
    ap.add_argument("--block", type=int, default=25)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-graphs", action="store_true", help="launch kernels directly (for profilers)")
    ap.add_argument("--kernel-only", default=None,
                    help="after warm-up, launch only this kernel family (nonbonded, pme, integrate, bonded, rebuild) "
                         "and the calibration copy, for hardware-counter passes; prints their durations and exits")
    ap.add_argument("--reps", type=int, default=50)
    ap.add_argument("--single-steps", type=int, default=100,
                    help="replica workloads: steps of each replica alone for the one-system-per-launch baseline")
    ap.add_argument("--replica-minimise", type=int, default=150)
    ap.add_argument("--serial-forces", action="store_true",
                    help="run the force branches one after another (for per-kernel profiles)")
    ap.add_argument("--replica-skin", type=float, default=0.0, help="replica workloads: nm; 0 derives skinFor")
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
    if args.skin > 0:
        opts["skin"] = args.skin
    if args.lanes > 0:
        opts["lanes_per_atom"] = args.lanes
    sim = R.GpuSimulation(system, **opts)
    if args.no_graphs:
        sim.set_graphs(False)
    diag = sim.diagnostics()
    pme = {"alpha": diag["pme_alpha"], "grid": diag["pme_grid"]}
    log(f"workload {args.workload}: {system.atom_count} atoms, grid {pme['grid']}, skin {diag['skin']:.3f}")
    mini = sim.minimise(max_iterations=100, force_tolerance=1000.0)
    log(f"minimised: {mini['iterations']} iterations, potential {mini['potential']:.1f}, "
        f"max force {mini['max_force']:.1f}")
    positions = sim.read_positions().astype(np.float64)
    sim.set_state(positions, R.maxwell_boltzmann(system, 300.0, 202), step=0)
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

    t0 = time.perf_counter()
    sim.run(args.steps, args.block)
    elapsed = time.perf_counter() - t0
    ms_per_step = 1e3 * elapsed / args.steps
    ns_day = TIMESTEP_PS * 1e-3 * 86400.0 / (ms_per_step * 1e-3)
    energy = sim.energy()
    diag = sim.diagnostics()
    log(f"timed {args.steps} steps: {ms_per_step * 1e3:.1f} us/step, {ns_day:.1f} ns/day; "
        f"T {energy['temperature']:.1f} K, volume {energy['volume']:.4f} nm^3, "
        f"barostat {diag['barostat_accepted']}/{diag['barostat_attempts']}, rebuilds {diag['rebuilds']}")

    # Roofline of the dominant kernel at the final state.
    state = sim.read_state()
    state["box"] = np.array(energy["box"])
    nb_ms = sim.profile_kernel("nonbonded", reps=200)
    pairs = in_cutoff_pairs(state["positions"], state["box"], system.cutoff, system.exclusion_start,
                            system.exclusion_partners)
    achieved = FLOP_PER_PAIR * pairs / (nb_ms * 1e-3) / 1e12
    traffic, traffic_source = counted_traffic(args.workload, diag["lanes_per_atom"])
    kernels = {k: sim.profile_kernel(k, reps=100) for k in ("pme", "integrate", "bonded", "rebuild")}
    log(f"nonbonded {nb_ms * 1e3:.2f} us/launch over {pairs} pairs -> {achieved:.3f} TFLOP/s; "
        + ", ".join(f"{k} {v * 1e3:.2f} us" for k, v in kernels.items()))

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
            "block": args.block,
            "graphs": not args.no_graphs,
        },
        "roofline": {
            "kernel": "nonbonded (LJ + screened Coulomb real space)",
            "bound": "mfma",
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
    }
    if not args.no_cpu_baseline:
        gpu_terms, _ = sim.evaluate(forces=False)
        ref, baseline = cpu_baseline(system, pme, state, args.cpu_seconds)
        pairs_ = {
            "bonded": (gpu_terms["bonded"], ref["bonded"]),
            "lennard_jones": (gpu_terms["lennard_jones"], ref["lennard_jones"]),
            "coulomb_direct": (gpu_terms["coulomb_direct"], ref["coulomb_direct_total"]),
            "coulomb_reciprocal": (gpu_terms["coulomb_reciprocal"], ref["reciprocal"]),
            "total": (gpu_terms["potential"] - gpu_terms["dispersion"], ref["total"]),
        }
        check = {k: abs(a - b) / max(abs(b), 1.0) for k, (a, b) in pairs_.items()}
        result["energy_check"] = {"max_rel_err": max(check.values()), "terms": check}
        baseline["cores"] = 1
        result["cpu_baseline"] = baseline
        log(f"energy check vs FP64: max rel {result['energy_check']['max_rel_err']:.2e}; "
            f"cpu baseline {baseline['ms_per_step']:.1f} ms/step")
    sim.release()
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
