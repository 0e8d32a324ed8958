"""Builds libvertexsurge.so in-tree for gfx950.

HIP sources are compiled with hipcc (one process per file, in parallel), the
host half (generators, samplers) with g++ -fopenmp, and everything is linked
into ``vertexsurge_amd/libvertexsurge.so``.  The C++ operator shims and their
test driver are built into ``vertexsurge_amd/_build``.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
BUILD = PKG / "_build"
LIB = PKG / "libvertexsurge.so"
INCLUDE = ROOT / "include"

HIP_SOURCES = ["vs_graph.hip", "vs_vexpand.hip", "vs_mintersect.hip", "vs_mlp.hip",
               "vs_batch.hip", "vs_tables.hip", "vs_flat.hip"]
HOST_SOURCES = ["vs_host.cpp", "vs_store.cpp"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
GXX = os.environ.get("CXX_HOST", "g++")
ARCH = "gfx950"

HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17",
    # packed FP32 forms are kept out of generated code on this platform
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result", f"-I{INCLUDE}",
]
HOST_FLAGS = ["-O3", "-fPIC", "-std=c++17", "-fopenmp", "-Wall", f"-I{INCLUDE}"]


def _run(cmd: list[str]) -> None:
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if proc.returncode != 0:
        sys.stderr.write(proc.stdout)
        raise RuntimeError(f"build step failed ({proc.returncode}): {' '.join(cmd)}")


def _compile_hip(src: str) -> Path:
    obj = BUILD / (src + ".o")
    _run([HIPCC, *HIP_FLAGS, "-c", str(CSRC / src), "-o", str(obj)])
    return obj


def _compile_host(src: str) -> Path:
    obj = BUILD / (src + ".o")
    _run([GXX, *HOST_FLAGS, "-c", str(CSRC / src), "-o", str(obj)])
    return obj


def build(verbose: bool = True) -> Path:
    """Compiles every source (always) and links the shared library."""
    if shutil.which(HIPCC) is None and not Path(HIPCC).exists():
        raise RuntimeError(f"hipcc not found at {HIPCC}")
    BUILD.mkdir(exist_ok=True)
    jobs = max(1, min(8, os.cpu_count() or 1))
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        futs = [ex.submit(_compile_hip, s) for s in HIP_SOURCES if (CSRC / s).exists()]
        futs += [ex.submit(_compile_host, s) for s in HOST_SOURCES]
        objs = [f.result() for f in futs]
    tmp = LIB.with_suffix(".so.tmp")
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp),
          *map(str, objs), "-lgomp", "-Wl,-rpath,/opt/rocm/lib"])
    os.replace(tmp, LIB)
    # C++ operator shims test driver (links the library just built)
    _run([GXX, "-O2", "-std=c++17", "-Wall", "-o", str(BUILD / "test_operators"),
          str(ROOT / "tests" / "cpp" / "test_operators.cpp"), f"-L{PKG}", "-lvertexsurge",
          "-Wl,-rpath,$ORIGIN/.."])
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    build()
