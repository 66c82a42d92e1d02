"""Compiles the gfx950 library ``librhizome_sdf.so`` in-tree with hipcc.

Usage: ``python rhizome_v24.15_amd/build.py`` (or ``__graft_entry__.build()``).
Each source compiles to an object under ``_build/`` in parallel, then one
link step writes the shared library next to this file.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "librhizome_sdf.so")

SOURCES = [
    "k_grid.hip",
    "k_splat.hip",
    "k_parity.hip",
    "k_fsm.hip",
    "k_winding.hip",
    "k_mesh.hip",
    "k_learned.hip",
    "k_mlp.hip",
    "k_cluster.hip",
    "k_mc.hip",
    "k_topology.hip",
    "noalloc.hip",
    "sdf_pipeline.cpp",
]

ARCH = os.environ.get("RHZ_OFFLOAD_ARCH", "gfx950")

# -ffp-contract=off: the kernels compared bit-for-bit with the CPU oracle
# (band distances, scanline hits, Godunov) must round every product;
# kernels that want FMAs write __builtin_fmaf.  Packed FP32 ops are disabled.
COMMON = [
    "-O3",
    "-std=c++17",
    "-fPIC",
    f"--offload-arch={ARCH}",
    "-ffp-contract=off",
    "-fno-fast-math",
    "-fno-gpu-flush-denormals-to-zero",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
    "-Wno-unused-variable",
    "-Wno-unused-but-set-variable",
    f"-I{os.path.join(REPO, 'include')}",
    f"-I{CSRC}",
]


def hipcc() -> str:
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found: the gfx950 library cannot be built")


# Per-source additions.  k_fsm.hip: the sweeps never see a NaN (every value is
# a distance or the 1e30 sentinel), so fminf / fmaxf need no NaN quieting.
EXTRA = {"k_fsm.hip": ["-fno-honor-nans"]}


def _compile(src: str) -> str:
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    cmd = [hipcc(), *COMMON, *EXTRA.get(src, []), "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    # hipcc's host pass warns that it does not know -packed-fp32-ops; drop that line
    noise = ("packed-fp32-ops", "not a recognized feature", "warning generated", "warnings generated")
    out = "\n".join(l for l in (r.stderr or "").splitlines() if not any(n in l for n in noise)).strip()
    if out:
        print(out, file=sys.stderr)
    return obj


def build(verbose: bool = True) -> str:
    os.makedirs(BUILD, exist_ok=True)
    srcs = [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    jobs = min(len(srcs), int(os.environ.get("MAX_JOBS", "8")) or 1)
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-pthread", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    build_dropin_example()
    if verbose:
        print(f"built {LIB}")
    return LIB


DROPIN_SRC = os.path.join(REPO, "tests", "cxx", "dropin_main.cpp")
DROPIN_BIN = os.path.join(REPO, "tests", "cxx", "dropin_main")


def build_dropin_example() -> str | None:
    """Host-only C++ program over include/rhizome/cuda_rasterizer.hpp (g++,
    linked against the library by a relative rpath so the tree can move)."""
    if not os.path.exists(DROPIN_SRC):
        return None
    rel = os.path.relpath(HERE, os.path.dirname(DROPIN_BIN))
    cmd = ["g++", "-std=c++17", "-O2", "-Wall", f"-I{os.path.join(REPO, 'include')}", DROPIN_SRC,
           f"-L{HERE}", "-lrhizome_sdf", f"-Wl,-rpath,$ORIGIN/{rel}", "-o", DROPIN_BIN]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"drop-in example build failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return DROPIN_BIN


if __name__ == "__main__":
    build()
