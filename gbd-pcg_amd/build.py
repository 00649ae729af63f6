"""Builds libgbdpcg.so (HIP kernels for gfx950 + the C ABI) in-tree with hipcc.

Usage: python gbd-pcg_amd/build.py   (always recompiles)
"""
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OUT_DIR = os.path.join(HERE, "gbdpcg")
LIB = os.path.join(OUT_DIR, "libgbdpcg.so")
BUILD_DIR = os.path.join(HERE, "_build")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
# Packed FP32 (v_pk_*_f32) disabled for all device code.
DEVICE_FLAGS = ["--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC", "-Wall",
                "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
HOST_FLAGS = ["--offload-arch=" + ARCH, "-O2", "-std=c++17", "-fPIC", "-Wall"]

HIP_SOURCES = ["pcg_kernel.hip", "pcg_reg_kernel.hip", "pcg_reg_ref_kernel.hip", "pcg_rowcast_kernel.hip",
               "pcg_pair_kernel_f32.hip", "pcg_pair_kernel_f64.hip",
               "aux_kernels.hip", "noalloc.hip"]
CPP_SOURCES = ["api.cpp", "batch.cpp", "problem.cpp", "tune.cpp"]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    noise = "is not a recognized feature for this target"
    err = "\n".join(l for l in r.stderr.splitlines() if noise not in l)
    if r.returncode != 0:
        raise RuntimeError("compile failed: " + " ".join(cmd) + "\n" + err)
    if verbose and err.strip():
        print(err, flush=True)


def build(verbose: bool = False) -> str:
    os.makedirs(BUILD_DIR, exist_ok=True)
    jobs = []
    objs = []
    for s in HIP_SOURCES:
        o = os.path.join(BUILD_DIR, s + ".o")
        objs.append(o)
        jobs.append([HIPCC, *DEVICE_FLAGS, "-c", os.path.join(CSRC, s), "-o", o])
    for s in CPP_SOURCES:
        o = os.path.join(BUILD_DIR, s + ".o")
        objs.append(o)
        jobs.append([HIPCC, *HOST_FLAGS, "-c", os.path.join(CSRC, s), "-o", o])
    with ThreadPoolExecutor(max_workers=min(8, len(jobs))) as ex:
        list(ex.map(lambda c: _run(c, verbose), jobs))
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp, *objs], verbose)
    os.replace(tmp, LIB)
    # the C++ drop-in example: host-only C++ against include/ and the library
    ex = os.path.join(OUT_DIR, "pcg_solve_example")
    _run(["g++", "-O2", "-std=c++17", "-Wall", "-I" + os.path.join(ROOT, "include"),
          os.path.join(HERE, "examples", "pcg_solve.cpp"), "-o", ex + ".tmp", "-L" + OUT_DIR, "-lgbdpcg",
          "-Wl,-rpath,$ORIGIN"], verbose)
    os.replace(ex + ".tmp", ex)
    return LIB


def clean():
    shutil.rmtree(BUILD_DIR, ignore_errors=True)


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
