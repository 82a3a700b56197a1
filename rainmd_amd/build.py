"""Build the gfx950 shared library librainmd_amd.so in-tree.

Every source under csrc/ is compiled by hipcc for gfx950 (host C++ files go
through the same driver so one toolchain links the result), in parallel, and
linked into rainmd_amd/librainmd_amd.so next to this file.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIBRARY = os.path.join(HERE, "librainmd_amd.so")
SOURCES = [
    "host_amber.cpp",
    "host_system.cpp",
    "neighbours.hip",
    "nonbonded.hip",
    "bonded.hip",
    "pme.hip",
    "integrate.hip",
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
    return LIBRARY


if __name__ == "__main__":
    print(build(verbose=True))
