"""Build the HIP backend in place: embloks_amd/libembloks_hip.so for gfx950.

All device code is compiled with hipcc for gfx950 only, with the packed-FP32
instructions turned off (-Xclang -target-feature -Xclang -packed-fp32-ops); the
host-only FP64 preparation (setup.cpp) is compiled with the host C++ compiler.
"""
from __future__ import annotations

import concurrent.futures
import os
import pathlib
import shutil
import subprocess

HERE = pathlib.Path(__file__).resolve().parent
CSRC = HERE / "csrc"
OUT = HERE / "libembloks_hip.so"
BUILD = HERE.parent / "_build" / "embloks_amd"

HIP_SOURCES = ["kernels.hip", "gemm.hip", "runtime.hip", "server.hip", "poisson.hip", "tables.hip"]
# The table build sums in the reference's order without fused multiply-adds.
EXTRA_FLAGS = {"tables.hip": ["-ffp-contract=off"]}
CXX_SOURCES = ["setup.cpp"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
CXX_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-Wall"]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"build failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return r


def build(verbose: bool = False) -> pathlib.Path:
    """Compile every source and link the shared library (always rebuilds)."""
    BUILD.mkdir(parents=True, exist_ok=True)
    jobs = []
    for src in HIP_SOURCES:
        jobs.append([HIPCC, *HIP_FLAGS, *EXTRA_FLAGS.get(src, []), "-c", str(CSRC / src), "-o", str(BUILD / (src + ".o"))])
    cxx = shutil.which("g++") or "c++"
    for src in CXX_SOURCES:
        jobs.append([cxx, *CXX_FLAGS, "-c", str(CSRC / src), "-o", str(BUILD / (src + ".o"))])
    with concurrent.futures.ThreadPoolExecutor(max_workers=min(8, len(jobs))) as pool:
        for r in pool.map(_run, jobs):
            if verbose and r.stderr.strip():
                print(r.stderr)
    objs = [str(BUILD / (s + ".o")) for s in HIP_SOURCES + CXX_SOURCES]
    tmp = OUT.with_suffix(".so.tmp")
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", str(tmp), *objs])
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))
