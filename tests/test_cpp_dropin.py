"""The C++ drop-in surface (include/epetreco/) compiled with g++ against
libepet.so: host-only checks here, device checks (workqueue cursors, the
chunked reco.cu loop) on the GPU, its image compared with the FP64 oracle run
on the same ECSR chunks."""
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT

SRC = os.path.join(ROOT, "tests", "cpp", "test_dropin.cpp")
OUT = os.path.join(ROOT, "tests", "cpp", "_build")


def _build():
    from epetreco_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        from epetreco_amd import build
        build.build()
    os.makedirs(OUT, exist_ok=True)
    exe = os.path.join(OUT, "test_dropin")
    libdir = os.path.dirname(_lib.LIB_PATH)
    cmd = ["g++", "-std=c++17", "-O1", "-ffp-contract=off", "-I", os.path.join(ROOT, "include"),
           SRC, "-o", exe, "-L", libdir, "-l:libepet.so", f"-Wl,-rpath,{libdir}"]
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    return exe


def test_dropin_host_only():
    exe = _build()
    r = subprocess.run([exe], capture_output=True, text=True, timeout=60)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr


@pytest.mark.gpu
def test_dropin_device_and_reco_loop(tmp_path, oracle):
    exe = _build()
    r = subprocess.run([exe, "--gpu"], capture_output=True, text=True, timeout=300, cwd=tmp_path)
    print(r.stdout, flush=True)
    assert r.returncode == 0, r.stdout + r.stderr
    x = np.fromfile(tmp_path / "dropin_mlem_x.bin", dtype=np.float32)
    # the same loop on the oracle: matrix of every candidate (reference
    # structure) of the listed channels, 3 FP64 iterations
    from epetreco_amd import geometry as G
    s, g, _ = G.workload("small")
    lst = (13 * np.arange(600) + 5).astype(np.int32)
    y = (1.0 + (np.arange(600) % 7)).astype(np.float32)
    rp, col, val = oracle.sm_rows(s, g, lst, 8, 0, keep_zeros=True)
    x64 = oracle.mlem_f64(600, g.nvoxels, val, rp, col, y, np.ones(g.nvoxels, np.float32),
                          np.full(g.nvoxels, 1.0 / g.nvoxels), 3)
    assert np.linalg.norm(x - x64) / np.linalg.norm(x64) < 1e-4
