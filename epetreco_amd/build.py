"""Build libepet.so (HIP, gfx950) in-tree with hipcc.

    python -m epetreco_amd.build            # compile every source, link
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libepet.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("EPET_OFFLOAD_ARCH", "gfx950")

COMMON = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    # packed FP32 instructions stay off (compiler-generated v_pk_* FP32 forms)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]
# Bit-exact generation: no FMA contraction, IEEE division and square root.
EXACT = ["-ffp-contract=off", "-fhip-fp32-correctly-rounded-divide-sqrt"]

SOURCES = {
    "core.hip": EXACT,
    "sm_gen.hip": EXACT,
    "sparse.hip": [],
    "spmv.hip": [],
    "mlem.hip": [],
    "mframe.hip": [],
    "lds.hip": [],
    "blk.hip": [],
}


def _compile(src: str, extra: list[str]) -> str:
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC] + COMMON + extra + ["-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


DRIVERS = ["reco", "backprojection", "epet_h5"]
HDF5_ROOT = os.environ.get("HDF5_ROOT", "/opt/conda")
BIN = os.path.join(HERE, "bin")


def build_drivers(verbose: bool = False) -> list[str]:
    """The HDF5 command-line drivers (reco, backprojection, epet_h5), built with
    g++ against libepet.so and the HDF5 C library under HDF5_ROOT.  The HDF5
    library is linked by path and the system library directory comes first in
    the run path, so that the C++ runtime the HIP runtime needs is the
    system's.  Skipped, with a message, when the HDF5 headers are absent."""
    h5 = os.path.join(HDF5_ROOT, "include", "hdf5.h")
    if not os.path.exists(h5):
        print(f"drivers skipped: {h5} not found (set HDF5_ROOT)")
        return []
    os.makedirs(BIN, exist_ok=True)
    inc = os.path.join(os.path.dirname(HERE), "include")
    h5lib = os.path.join(HDF5_ROOT, "lib")
    out = []
    for name in DRIVERS:
        exe = os.path.join(BIN, name)
        cmd = ["g++", "-std=c++17", "-O2", "-I", inc, "-I", os.path.join(HDF5_ROOT, "include"),
               os.path.join(HERE, "drivers", name + ".cpp"), "-o", exe,
               "-L", HERE, "-l:libepet.so", os.path.join(h5lib, "libhdf5.so"),
               "-Wl,-rpath-link," + h5lib,
               "-Wl,-rpath,$ORIGIN/..:/usr/lib/x86_64-linux-gnu:" + h5lib]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"g++ failed for {name}:\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
        out.append(exe)
    if verbose:
        print("built drivers:", " ".join(out))
    return out


def build(verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as ex:
        objs = list(ex.map(lambda kv: _compile(*kv), SOURCES.items()))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    build_drivers(verbose=True)
    sys.exit(0)
