"""The autotuner for the MD step's kernel shapes.

It measures, on the GPU it runs on, the step time of representative
workloads under candidate shapes and keeps the fastest:

- lanes: the direct space's tile, lanes per atom (the width of the lane
  group that walks one atom's neighbour row);
- reach and search: the neighbour search's blocks, the widest cell stencil
  reach the plan may take (cells of (rc + skin)/reach) and whether a wave
  searches for one atom or for one cell;
- skin: the Verlet skin, which trades list rebuilds against list length
  (the list is rebuilt on a device-side displacement test, so the skin is a
  speed choice, not a correctness one);
- pme_bins: whether a single system's PME planes read per-plane atom bins or
  scan every atom.

Workload classes: single systems of 1,189, 2,854, 12,472 and 100,200 atoms
(the table's rows start at 0, 2,000, 6,000 and 40,000 atoms) and the uneven
256-replica batch, each in its configuration's ensemble (NVT, or NPT at 1 bar
with a volume move every 25 steps). Each class is prepared once (minimised,
Maxwell-Boltzmann velocities at 300 K, 5,000 steps), and every trial restarts
from that
state: a fresh context with the trial's options, a warm-up, then the best of
two timed rounds (tools/gpu/bench-throughput.mjs:69-89 takes the minimum of
rounds the same way). The search is coordinate descent over the parameters
in a fixed order, two passes, from the library's current table.

The reference has no tuner: its skin and rebuild interval were swept by hand
(tools/gpu/bench-skin.mjs, docs/STATUS.md:1104-1116).

    python -m rainmd_amd.autotune            # measure and print
    python -m rainmd_amd.autotune --write    # and write tuning.json, csrc/tuned.h

Then rebuild the library (rainmd_amd.build) so the new table takes effect.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
TABLE_JSON = os.path.join(HERE, "tuning.json")
TABLE_HEADER = os.path.join(HERE, "csrc", "tuned.h")

# search_mode 0 and pme_bins 0 are the library's rules (by cell occupancy, by
# atom count): in a batch the rule decides per replica
SINGLE = {"lanes": [16, 32, 64], "search_mode": [0, 1, 2], "cell_reach": [1, 2, 3], "pme_bins": [0, 1, 2],
          "skin": [0.0, 0.1, 0.15, 0.2, 0.3, 0.4]}
BATCH = {"lanes": [8, 16, 32], "search_mode": [0, 1, 2], "cell_reach": [2, 3], "skin": [0.1, 0.12, 0.15, 0.2]}
ORDER = ["lanes", "search_mode", "cell_reach", "pme_bins", "skin"]

# (label, kind, minAtoms, how to build the workload, ensemble): each class runs
# in its configuration's ensemble (NPT: 1 bar, a volume move every 25 steps)
NPT = dict(pressure=1.0, barostat_interval=25)
CLASSES = [
    ("ala-389w", 0, 0, ("workload", "ala-389w"), {}),
    ("ala-944w", 0, 2000, ("workload", "ala-944w"), NPT),
    ("replica-5nm", 0, 6000, ("largest-replica", None), NPT),
    ("water-33400", 0, 40000, ("workload", "water-33400"), {}),
    ("replicas-256", 1, 0, ("batch", 256), NPT),
]
PREPARE_STEPS = 5000  # thermalisation before the trials, at 300 K


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def _options(params):
    # a table value of 0 is the library's rule, which the options ask for as 3
    o = dict(lanes_per_atom=params["lanes"], search_mode=params["search_mode"] or 3, cell_reach=params["cell_reach"])
    if "pme_bins" in params:
        o["pme_bins"] = params["pme_bins"] or 3
    # skin 0 is the reference's skinFor
    if params["skin"] > 0:
        o["skin"] = params["skin"]
    return o


def _room(system):
    """The largest skin the box leaves beside the cutoff (skinFor's cap)."""
    box = np.asarray(system.box).reshape(9)
    return 0.9 * (0.5 * min(box[0], box[4], box[8]) - system.cutoff)


class SingleWorkload:
    def __init__(self, system, steps, ensemble):
        import rainmd_amd as R

        self.R, self.steps, self.ensemble = R, steps, ensemble
        sim = R.GpuSimulation(system, temperature=300.0, seed=77, **ensemble)
        sim.minimise(max_iterations=500, force_tolerance=2000.0)
        sim.set_state(sim.read_positions().astype(np.float64), R.maxwell_boltzmann(system, 300.0, 77), step=0)
        sim.run(PREPARE_STEPS, 25)
        st = sim.read_state()
        # trials start in the box the thermalisation left
        self.system = system.rebox(st["box_vectors"]) if ensemble else system
        self.x, self.v = st["positions"].astype(np.float64), st["velocities"].astype(np.float64)
        sim.release()

    def allowed(self, name, value):
        return not (name == "skin" and value > _room(self.system))

    def measure(self, params, rounds=2):
        sim = self.R.GpuSimulation(self.system, temperature=300.0, seed=78, **self.ensemble, **_options(params))
        sim.set_state(self.x, self.v, step=0)
        sim.run(100, 25)
        sim.run(self.steps % 25 or 25, 25)
        best = float("inf")
        for _ in range(rounds):
            t0 = time.perf_counter()
            sim.run(self.steps, 25)
            sim.diagnostics()  # synchronises the stream
            best = min(best, (time.perf_counter() - t0) / self.steps * 1e3)
        sim.release()
        return best


class BatchWorkload:
    def __init__(self, count, steps, ensemble):
        import rainmd_amd as R

        self.R, self.steps, self.ensemble = R, steps, ensemble
        work = R.replica_workload(count)
        self.seeds = [s for _, _, s in work]
        systems = R.replica_systems(count)
        batch = R.Batch(systems, seeds=self.seeds, temperature=300.0, **ensemble)
        self.states = []
        for r, system in enumerate(systems):
            sim = batch.replica(r)
            sim.minimise(max_iterations=150, force_tolerance=2000.0)
            sim.set_state(sim.read_positions().astype(np.float64), R.maxwell_boltzmann(system, 300.0, self.seeds[r]),
                          step=0)
        # minimised lattices release heat for the first few picoseconds; the
        # rebuild rate (the skin's cost) is measured at the bath temperature
        batch.run(PREPARE_STEPS, 25)
        self.systems = []
        for r in range(count):
            st = batch.replica(r).read_state()
            self.states.append((st["positions"].astype(np.float64), st["velocities"].astype(np.float64)))
            self.systems.append(systems[r].rebox(st["box_vectors"]) if ensemble else systems[r])
        batch.release()

    def allowed(self, name, value):
        return True

    def measure(self, params, rounds=2):
        batch = self.R.Batch(self.systems, seeds=self.seeds, temperature=300.0, **self.ensemble, **_options(params))
        for r, (x, v) in enumerate(self.states):
            batch.replica(r).set_state(x, v, step=0)
        batch.run(100, 25)
        best = float("inf")
        for _ in range(rounds):
            t0 = time.perf_counter()
            batch.run(self.steps, 25)
            batch.info()
            best = min(best, (time.perf_counter() - t0) / self.steps * 1e3)
        batch.release()
        return best


def descend(workload, space, start, passes=2):
    """Coordinate descent over `space` in ORDER from `start`; returns the best
    parameters, their time and every trial."""
    trials = []
    cache = {}

    def timed(params):
        key = json.dumps(params, sort_keys=True)
        if key not in cache:
            cache[key] = workload.measure(params)
            trials.append({"params": dict(params), "ms_per_step": cache[key]})
            log(f"  {params} -> {cache[key]:.4f} ms/step")
        return cache[key]

    best = dict(start)
    best_ms = timed(best)
    for _ in range(passes):
        for name in ORDER:
            if name not in space:
                continue
            for value in space[name]:
                if value == best[name] or not workload.allowed(name, value):
                    continue
                trial = dict(best, **{name: value})
                ms = timed(trial)
                if ms < best_ms:
                    best, best_ms = trial, ms
    return best, best_ms, trials


def current_table():
    """The rows of the compiled-in table (csrc/tuned.h)."""
    rows = []
    with open(TABLE_HEADER) as f:
        text = f.read()
    body = text[text.index("TUNED_SHAPES[] = {") + len("TUNED_SHAPES[] = {"):text.index("};", text.index("TUNED_SHAPES"))]
    for line in body.strip().splitlines():
        vals = line.strip().strip(",").strip("{}").split(",")
        kind, min_atoms, lanes, reach, search = (int(v) for v in vals[:5])
        rows.append(dict(kind=kind, minAtoms=min_atoms, lanes=lanes, cell_reach=reach, search_mode=search,
                         skin=float(vals[5]), pme_bins=int(vals[6])))
    return rows


def header(rows):
    lines = [
        "// Kernel shapes picked by the autotuner (rainmd_amd/autotune.py), which",
        "// writes this file from its measurements (rainmd_amd/tuning.json).",
        "// Regenerate with: python -m rainmd_amd.autotune --write",
        "#pragma once",
        "",
        "namespace rmd {",
        "",
        "// One row per workload class: single systems (kind 0) by atom count, replica",
        "// batches (kind 1). lanes: direct-space lanes per atom; reach: the widest cell",
        "// stencil reach the plan may take; searchMode: 0 by cell occupancy, 1 one wave",
        "// per atom, 2 one wave per cell; skin: nm when the caller asks for the tuned",
        "// skin (0: the reference's skinFor); pmeBins: 0 by atom count, 1 atom bins,",
        "// 2 every plane scans every atom.",
        "struct TunedShape {",
        "  int kind, minAtoms, lanes, reach, searchMode;",
        "  double skin;",
        "  int pmeBins;",
        "};",
        "",
        "constexpr TunedShape TUNED_SHAPES[] = {",
    ]
    for r in rows:
        lines.append(f"    {{{r['kind']}, {r['minAtoms']}, {r['lanes']}, {r['cell_reach']}, {r['search_mode']}, "
                     f"{r['skin']!r}, {r['pme_bins']}}},")
    lines += ["};", "", "}  // namespace rmd", ""]
    return "\n".join(lines)


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--write", action="store_true", help="write tuning.json and csrc/tuned.h")
    ap.add_argument("--classes", default=",".join(c[0] for c in CLASSES))
    ap.add_argument("--passes", type=int, default=2)
    args = ap.parse_args()
    import rainmd_amd as R

    wanted = args.classes.split(",")
    table = current_table()
    out = {"device": R.device_info().get("name"), "method": "coordinate descent, best of 2 timed rounds per trial",
           "classes": []}
    rows = []
    for label, kind, min_atoms, (how, what), ensemble in CLASSES:
        row = next((r for r in table if r["kind"] == kind and r["minAtoms"] == min_atoms),
                   next(r for r in table if r["kind"] == kind))
        start = dict(lanes=row["lanes"], search_mode=row["search_mode"], cell_reach=row["cell_reach"],
                     skin=row["skin"])
        if kind == 0:
            start["pme_bins"] = row["pme_bins"]
        if label not in wanted:
            rows.append(dict(row, kind=kind, minAtoms=min_atoms))
            continue
        t0 = time.perf_counter()
        if how == "batch":
            workload, space, atoms = BatchWorkload(what, 300, ensemble), BATCH, None
        else:
            if how == "workload":
                system = R.workload_system(what)
            else:
                edge, waters, seed = max(R.replica_workload(256))
                system = R.System.solvated(water_count=waters, box_edge=edge, seed=seed, include_solute=True,
                                           cutoff=0.9)
            steps = 1000 if system.atom_count < 20000 else 300
            workload, space, atoms = SingleWorkload(system, steps, ensemble), SINGLE, system.atom_count
        log(f"{label}: prepared in {time.perf_counter() - t0:.1f} s; from {start}")
        best, best_ms, trials = descend(workload, space, start, args.passes)
        log(f"{label}: best {best} at {best_ms:.4f} ms/step ({len(trials)} trials)")
        out["classes"].append({"label": label, "kind": kind, "minAtoms": min_atoms, "atoms": atoms, "best": best,
                               "best_ms_per_step": best_ms, "start": start, "trials": trials})
        rows.append(dict(kind=kind, minAtoms=min_atoms, lanes=best["lanes"], cell_reach=best["cell_reach"],
                         search_mode=best["search_mode"], skin=float(best["skin"]),
                         pme_bins=best.get("pme_bins", 0)))
    print(json.dumps(out, indent=1))
    if args.write:
        # classes not measured this time keep their earlier record
        if os.path.exists(TABLE_JSON):
            with open(TABLE_JSON) as f:
                earlier = json.load(f)
            done = {c["label"] for c in out["classes"]}
            out["classes"] = [c for c in earlier.get("classes", []) if c["label"] not in done] + out["classes"]
            order = [c[0] for c in CLASSES]
            out["classes"].sort(key=lambda c: order.index(c["label"]) if c["label"] in order else len(order))
        with open(TABLE_JSON, "w") as f:
            json.dump(out, f, indent=1)
        with open(TABLE_HEADER, "w") as f:
            f.write(header(rows))
        log(f"wrote {TABLE_JSON} and {TABLE_HEADER}; rebuild the library for the table to take effect")


if __name__ == "__main__":
    main()
