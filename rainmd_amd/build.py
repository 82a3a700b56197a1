"""Build the gfx950 shared library librainmd_amd.so in-tree, and the Node
N-API addon over it.

Every source under csrc/ is compiled by hipcc for gfx950 (host C++ files go
through the same driver so one toolchain links the result), in parallel, and
linked into rainmd_amd/librainmd_amd.so next to this file. The addon
(node/addon.cc) is host C++ against Node's node_api.h, built with g++ into
rainmd_amd/node/rainmd.node, linked to the library by an $ORIGIN rpath.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIBRARY = os.path.join(HERE, "librainmd_amd.so")
ADDON_SOURCE = os.path.join(HERE, "node", "addon.cc")
ADDON = os.path.join(HERE, "node", "rainmd.node")
NODE_INCLUDE = os.environ.get("NODE_INCLUDE", "/usr/include/node")
SOURCES = [
    "host_amber.cpp",
    "host_system.cpp",
    "neighbours.hip",
    "nonbonded.hip",
    "bonded.hip",
    "pme.hip",
    "integrate.hip",
    "surrogate.hip",
    "context.hip",
]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
COMMON = ["-O3", "-std=c++17", "-fPIC", "-Wall"]
# hipcc's own -O3 forms packed FP32 operations from ordinary float code; they
# are switched off for every kernel in this library.
DEVICE = [f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def _compile(source):
    src = os.path.join(CSRC, source)
    obj = os.path.join(BUILD, source + ".o")
    cmd = [HIPCC] + COMMON + DEVICE + ["-c", src, "-o", obj]
    proc = subprocess.run(cmd, capture_output=True, text=True)
    noise = "is not a recognized feature for this target"
    errs = "\n".join(line for line in proc.stderr.splitlines() if noise not in line)
    if proc.returncode != 0:
        raise RuntimeError(f"hipcc failed on {source}:\n{errs}")
    return obj, errs


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    workers = min(len(SOURCES), int(os.environ.get("MAX_JOBS", os.cpu_count() or 4)))
    with ThreadPoolExecutor(max_workers=max(1, workers)) as pool:
        results = list(pool.map(_compile, SOURCES))
    if verbose:
        for obj, errs in results:
            if errs.strip():
                print(errs, file=sys.stderr)
    objects = [obj for obj, _ in results]
    tmp = LIBRARY + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objects
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        raise RuntimeError("link failed:\n" + proc.stderr)
    os.replace(tmp, LIBRARY)
    build_addon()
    return LIBRARY


def build_addon():
    """The N-API addon; skipped (returns None) where Node's headers are absent."""
    if not os.path.exists(os.path.join(NODE_INCLUDE, "node_api.h")):
        return None
    tmp = ADDON + ".tmp"
    cmd = ["g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-Wall", "-DNODE_GYP_MODULE_NAME=rainmd",
           f"-I{NODE_INCLUDE}", ADDON_SOURCE, "-o", tmp, f"-L{HERE}", "-lrainmd_amd",
           "-Wl,-rpath,$ORIGIN/.."]
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        raise RuntimeError("addon build failed:\n" + proc.stderr)
    os.replace(tmp, ADDON)
    return ADDON


if __name__ == "__main__":
    print(build(verbose=True))
