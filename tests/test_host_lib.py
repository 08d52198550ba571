"""CPU checks of the C-ABI boundary (no GPU needed): libepet.so loads, exports
every function include/epet.h declares, binds them in the Python surface, and
fails loudly (no CPU fallback) when no device is present."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "epet.h")


def declared_functions():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(epet_[a-zA-Z0-9_]+)\s*\(", src)))


def test_library_builds_and_exports_every_declared_symbol():
    from epetreco_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        from epetreco_amd import build
        build.build()
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = set(re.findall(r"\bT (epet_\w+)", out))
    missing = [f for f in declared_functions() if f not in exported]
    assert not missing, f"declared but not exported: {missing}"
    assert set(declared_functions()) <= set(_lib.SIGNATURES), "Python bindings incomplete"


def test_host_only_chunk_helpers_match_oracle(oracle):
    from epetreco_amd import _lib
    L, O = _lib.lib(), oracle.lib()
    for args in [(0, 1000, 300), (3, 1000, 300), (4, 1000, 300), (2, 900, 300), (0, 0, 5)]:
        assert L.epet_nInChunk(*args) == O.eo_nInChunk(*args)
    assert L.epet_chunkPtr(7, 954) == 7 * 954
    # LIMNNZ chunks of the reference at 52^3 (tests/reco.cu:26-30, :206)
    limm = (512 * 1024 * 1024 // 4) // 140608
    assert L.epet_nChunks(5140980 * 140608, limm * 140608) == O.eo_nChunks(5140980 * 140608,
                                                                         limm * 140608)


def test_errors_are_loud_without_a_device():
    import torch
    if torch.cuda.is_available():
        pytest.skip("device present")
    from epetreco_amd import _lib
    L = _lib.lib()
    h = ctypes.c_void_p()
    rc = L.epet_ctx_create(ctypes.byref(h), 0, None)
    assert rc != _lib.EPET_OK
    assert L.epet_last_error().decode()   # message with file(line)
    from epetreco_amd.api import Context
    with pytest.raises(RuntimeError):     # EpetError or torch's own error: never a fallback
        Context(0)


def test_null_arguments_rejected_before_any_device_call():
    from epetreco_amd import _lib
    L = _lib.lib()
    assert L.epet_set_geometry(None, None, None) == 1       # EPET_ERR_ARG
    assert L.epet_sum(None, None, 0, None) == 1
    assert "null" in L.epet_last_error().decode()


def test_geometry_struct_layout_matches_header():
    from epetreco_amd.geometry import CGrid, CSetup
    # DefaultMeasurementSetup<float>: 11 4-byte members; DefaultVoxelGrid<float>: 9
    assert ctypes.sizeof(CSetup) == 44
    assert ctypes.sizeof(CGrid) == 36
    assert [f[0] for f in CSetup._fields_] == ["pos0x", "pos1x", "na", "da", "n0z", "n0y", "n1z",
                                               "n1y", "segx", "segy", "segz"]


def test_phantoms_are_seeded_and_in_range():
    from epetreco_amd import geometry as G
    from epetreco_amd import phantom as PH
    g = G.grid_52_1mm()
    a, b = PH.phantom(g, "full", 0), PH.phantom(g, "full", 0)
    assert np.array_equal(a, b) and a.max() > 0.05 and a.min() >= 0
    c = g.centres()
    outside = c[:, 0] ** 2 + c[:, 2] ** 2 > 0.024 ** 2
    assert np.all(a[outside] <= 1.1 * a.max())       # background only inside the cylinder
    imgs, totals = PH.dynamic_series(G.grid_16_3mm(), "small", 32, 0)
    assert imgs.shape == (32, 4096) and np.all(totals >= 1e5) and np.all(totals <= 1e7)
    assert np.all(np.diff(imgs.sum(1)) <= 0)          # activities decay with t


def test_bench_refuses_more_than_one_gpu():
    """bench.py measures one GPU: --gpus 2 exits with status 2 and a JSON line
    that says so, before touching a device."""
    import json
    import sys
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    line = json.loads(r.stdout.strip().splitlines()[-1])
    assert line["n_gpus"] == 2 and "one GPU" in line["error"]
