"""Builds libsqaod_rocm.so (HIP kernels + C ABI) for gfx950, in-tree.

Usage: python -m sqaod_amd.build   (or sqaod_amd.build.build() from Python)
Outputs sqaod_amd/_lib/libsqaod_rocm.so; objects go to sqaod_amd/_build/.
"""
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
INCLUDE = os.path.join(os.path.dirname(HERE), "include")
OBJDIR = os.path.join(HERE, "_build")
LIBDIR = os.path.join(HERE, "_lib")
LIB = os.path.join(LIBDIR, "libsqaod_rocm.so")
ARCH = os.environ.get("SQAOD_ROCM_ARCH", "gfx950")

SOURCES = ["rng.hip", "dense_sweep.hip", "dense_naive.hip", "dense_field.hip", "dense_field_mfma.hip", "energy.hip", "energy_planes.hip", "dense_bf.hip", "bf_noalloc.hip", "bipartite.hip", "bipartite_bf.hip", "diag.hip", "dense_annealer.cpp", "dense_bf_searcher.cpp", "bipartite_annealer.cpp", "formulas.cpp", "bipartite_bf_searcher.cpp", "dense_batch.cpp", "serve.cpp", "capi.cpp"]

# -ffp-contract=off: the annealers' float arithmetic must round exactly like
# the CPU annealer (no fused multiply-adds); the packed-FP32 feature is off.
COMMON_FLAGS = [
    "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off", f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]


def _hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def _compile(src):
    obj = os.path.join(OBJDIR, os.path.splitext(src)[0] + ".o")
    cmd = [_hipcc(), *COMMON_FLAGS, f"-I{INCLUDE}", "-x", "hip", "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    if r.stderr.strip():
        sys.stderr.write(r.stderr)
    return obj


def build(verbose=True):
    os.makedirs(OBJDIR, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    with ThreadPoolExecutor(max_workers=min(6, os.cpu_count() or 1)) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [_hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    build()
