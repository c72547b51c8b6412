"""Build the gfx950 library libsdot.so in-tree (semidiscreteot_amd/lib/).

HIP sources go through hipcc for gfx950 only; the host-only C++ goes through
g++ against the HIP runtime headers.  Packed-FP32 instruction selection is
disabled for all device code (-Xclang -target-feature -Xclang -packed-fp32-ops).
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIBDIR = os.path.join(HERE, "lib")
LIB = os.path.join(LIBDIR, "libsdot.so")
RATES = os.path.join(LIBDIR, "sdot_rates")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")

HIP_SOURCES = ["sdot_kernels.hip", "sdot_serve_kernels.hip", "sdot_refine.hip", "sdot_points.hip",
               "sdot_layout.hip", "sdot_lbfgs.hip"]
CXX_SOURCES = ["sdot_api.cpp", "sdot_mesh.cpp", "sdot_prm.cpp", "sdot_serve.cpp"]

HIPCC_FLAGS = ["--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17",
               "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
               "-Wall", "-Wno-unused-result"]
CXX_FLAGS = ["-O2", "-fPIC", "-std=c++17", "-D__HIP_PLATFORM_AMD__",
             f"-I{ROCM}/include", "-Wall", "-Wextra", "-Wno-unused-parameter"]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build(verbose: bool = False, extra_hip_flags=None) -> str:
    os.makedirs(BUILD, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    hipcc = os.path.join(ROCM, "bin", "hipcc")
    jobs = []
    for src in HIP_SOURCES:
        obj = os.path.join(BUILD, src + ".o")
        jobs.append(([hipcc, *HIPCC_FLAGS, *(extra_hip_flags or []), "-c",
                      os.path.join(CSRC, src), "-o", obj], obj))
    for src in CXX_SOURCES:
        obj = os.path.join(BUILD, src + ".o")
        jobs.append((["g++", *CXX_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj], obj))
    # the translation units compile independently: one compiler process each
    workers = max(1, min(len(jobs), int(os.environ.get("MAX_JOBS", os.cpu_count() or 1))))
    with ThreadPoolExecutor(max_workers=workers) as pool:
        for f in [pool.submit(_run, cmd, verbose) for cmd, _ in jobs]:
            f.result()
    # the instruction-rate microbenchmark bench.py runs for the measured peaks
    rates = RATES
    jobs_rates = [hipcc, *HIPCC_FLAGS, os.path.join(CSRC, "sdot_rates.hip"), "-o", rates + ".tmp"]
    objs = [obj for _, obj in jobs]
    tmp = LIB + ".tmp"
    _run([hipcc, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs,
          f"-L{ROCM}/lib", "-lamdhip64", f"-Wl,-rpath,{ROCM}/lib"], verbose)
    os.replace(tmp, LIB)
    _run(jobs_rates, verbose)
    os.replace(rates + ".tmp", rates)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
