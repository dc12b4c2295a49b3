"""Builds libflagsparse_hip.so (gfx950) from flagsparse_amd/csrc with hipcc,
and the measurement library libflagsparse_bench.so from csrc/bench.

Each .hip translation unit is compiled to an object under _build/ in
parallel, then linked into flagsparse_amd/libflagsparse_hip.so (the operator
library: csrc/*.hip) and flagsparse_amd/libflagsparse_bench.so (the probes
bench.py times: csrc/bench/*.hip).  The libraries link libamdhip64.so.7;
imported after torch they bind to the HIP runtime torch already loaded (same
SONAME), so torch's stream handles are valid in them.
"""

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(PKG_DIR, "csrc")
REPO = os.path.dirname(PKG_DIR)
BUILD_DIR = os.path.join(PKG_DIR, "_build")
LIB_NAME = "libflagsparse_hip.so"
LIB_PATH = os.path.join(PKG_DIR, LIB_NAME)
BENCH_CSRC = os.path.join(CSRC, "bench")
BENCH_LIB_PATH = os.path.join(PKG_DIR, "libflagsparse_bench.so")

ARCH = os.environ.get("FLAGSPARSE_AMD_ARCH", "gfx950")

# Packed-FP32 instruction selection is switched off for every HIP object (the
# kernels are written for scalar FP32 FMA); code object v5 is accepted by the
# HIP 7.0 runtime that torch ships.
HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-mcode-object-version=5",
    "-Xclang",
    "-target-feature",
    "-Xclang",
    "-packed-fp32-ops",
    "-I",
    os.path.join(REPO, "include"),
]


def _hipcc():
    for cand in (os.environ.get("HIPCC"), shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found: the flagsparse_amd HIP library cannot be built")


def sources(directory=CSRC):
    return sorted(
        os.path.join(directory, f) for f in os.listdir(directory) if f.endswith(".hip")
    )


def _object(src):
    sub = "bench_" if os.path.dirname(src) == BENCH_CSRC else ""
    return os.path.join(BUILD_DIR, sub + os.path.basename(src)[:-4] + ".o")


def _compile(hipcc, src):
    obj = _object(src)
    cmd = [hipcc, *HIP_FLAGS, "-c", src, "-o", obj]
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        raise RuntimeError(
            f"hipcc failed for {src}:\n{' '.join(cmd)}\n{proc.stdout}\n{proc.stderr}"
        )
    return obj


def _link(hipcc, objs, path):
    tmp = path + ".tmp"
    cmd = [hipcc, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp]
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        raise RuntimeError(f"link failed:\n{' '.join(cmd)}\n{proc.stdout}\n{proc.stderr}")
    os.replace(tmp, path)


def build(verbose=True, jobs=None):
    """Compile every csrc/*.hip and csrc/bench/*.hip for gfx950 and link the
    operator and measurement libraries."""
    hipcc = _hipcc()
    os.makedirs(BUILD_DIR, exist_ok=True)
    srcs = sources()
    bench_srcs = sources(BENCH_CSRC)
    jobs = jobs or min(len(srcs) + len(bench_srcs), max(1, (os.cpu_count() or 4)))
    if verbose:
        print(f"[flagsparse_amd] compiling {len(srcs) + len(bench_srcs)} HIP sources for {ARCH}", flush=True)
    with ThreadPoolExecutor(max_workers=jobs) as pool:
        objs = list(pool.map(lambda s: _compile(hipcc, s), srcs + bench_srcs))
    _link(hipcc, objs[: len(srcs)], LIB_PATH)
    _link(hipcc, objs[len(srcs):], BENCH_LIB_PATH)
    if verbose:
        print(f"[flagsparse_amd] built {LIB_PATH} and {BENCH_LIB_PATH}", flush=True)
    return LIB_PATH


if __name__ == "__main__":
    build()
    sys.exit(0)
