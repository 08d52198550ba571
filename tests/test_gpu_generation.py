"""GPU parity of ray generation, workqueue and system-matrix generation.

Bar: bit-exact (integer structure and float values) against the CPU oracle,
which is itself pinned to the reference's host-compiled workqueue/geometry
(tests/golden) -- given identical ray endpoints.
"""
import hashlib
import os

import numpy as np
import pytest
import torch

from conftest import GOLDEN

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


def _sha(cnl, vxl):
    a = np.stack([np.asarray(cnl, np.int32), np.asarray(vxl, np.int32)], axis=1)
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


@pytest.mark.parametrize("wl", ["small", "full"])
def test_rays_bit_exact(gpu_ctx_factory, oracle, wl):
    ctx = gpu_ctx_factory(wl, seed=0x1234abcd9876)
    s = ctx.setup
    rng = np.random.default_rng(1)
    cnl = rng.integers(0, s.nchannels, 64).astype(np.int32)
    _say("rays", wl, ctx.nrays)
    got = ctx.rays(cnl).cpu().numpy()
    for i, c in enumerate(cnl):
        ref = oracle.rays(s, int(c), ctx.nrays, ctx.seed)
        np.testing.assert_array_equal(got[i], ref)


def test_workqueue_small_matches_reference(gpu_ctx_factory):
    ctx = gpu_ctx_factory("small")
    s, g = ctx.setup, ctx.grid
    z = np.load(os.path.join(GOLDEN, "small_geometry.npz"))
    lst = np.arange(s.nchannels, dtype=np.int32)
    _say("workqueue small full")
    cnl, vxl, li, vi = ctx.getWorkqueue(lst)
    cnl, vxl = cnl.cpu().numpy(), vxl.cpu().numpy()
    assert len(cnl) == int(z["wq_count"])
    assert _sha(cnl, vxl) == str(z["wq_sha256"])
    assert [li, vi] == list(z["wq_end_cursor"])
    # capped calls resume from the cursor (test_getWorkqueue.cu:186-198)
    for row, sha in zip(z["caps"], z["caps_sha256"]):
        l0, v0, cap, n, l1, v1 = (int(t) for t in row)
        c2, w2, l2, u2 = ctx.getWorkqueue(lst, l0, v0, cap)
        assert (len(c2), l2, u2) == (n, l1, v1)
        assert _sha(c2.cpu().numpy(), w2.cpu().numpy()) == str(sha)


def test_workqueue_full_sampled_rows(gpu_ctx_factory):
    ctx = gpu_ctx_factory("full")
    z = np.load(os.path.join(GOLDEN, "full_geometry.npz"))
    for k, c in enumerate(z["sample_cnl"]):
        _say("workqueue full channel", int(c))
        _, vxl, li, vi = ctx.getWorkqueue(np.array([c], np.int32))
        ref = z["sample_vxl"][z["sample_ptr"][k]:z["sample_ptr"][k + 1]]
        np.testing.assert_array_equal(vxl.cpu().numpy(), ref)
        assert (li, vi) == (1, 0)


def _rows_equal(rp, col, val, orp, ocol, oval):
    np.testing.assert_array_equal(rp, orp)
    np.testing.assert_array_equal(col, ocol)
    # bit-exact values (== treats +0/-0 alike; values here are >= 0)
    assert np.array_equal(val, oval), (
        f"{np.count_nonzero(val != oval)} values differ, max rel "
        f"{np.max(np.abs(val - oval) / np.maximum(np.abs(oval), 1e-30))}")


def test_sm_build_small_all_channels(gpu_ctx_factory, oracle):
    """Fused nonzero build == oracle candidates with value != 0, all 11,250 rows."""
    ctx = gpu_ctx_factory("small", seed=0)
    s, g = ctx.setup, ctx.grid
    lst = np.arange(s.nchannels, dtype=np.int32)
    _say("sm_build small, all channels")
    A = ctx.sm_build(lst)
    rp, col, val, _ = A.to_torch()
    orp, ocol, oval = oracle.sm_rows(s, g, lst, ctx.nrays, ctx.seed, keep_zeros=False)
    assert A.nnz == len(ocol)
    _rows_equal(rp.cpu().numpy(), col.cpu().numpy(), val.cpu().numpy(), orp, ocol, oval)


def test_sm_build_full_sampled_rows(gpu_ctx_factory, oracle):
    ctx = gpu_ctx_factory("full", seed=0)
    s, g = ctx.setup, ctx.grid
    rng = np.random.default_rng(7)
    lst = np.sort(rng.choice(s.nchannels, 24, replace=False)).astype(np.int32)
    _say("sm_build full, 24 sampled rows")
    A = ctx.sm_build(lst)
    rp, col, val, _ = A.to_torch()
    orp, ocol, oval = oracle.sm_rows(s, g, lst, ctx.nrays, ctx.seed, keep_zeros=False)
    _rows_equal(rp.cpu().numpy(), col.cpu().numpy(), val.cpu().numpy(), orp, ocol, oval)


def test_get_system_matrix_keep_zeros(gpu_ctx_factory, oracle):
    """Reference structure: every candidate, zeros included (getSystemMatrixDeviceOnly.cu:391-439)."""
    ctx = gpu_ctx_factory("small", seed=3)
    s, g = ctx.setup, ctx.grid
    rng = np.random.default_rng(11)
    lst = np.sort(rng.choice(s.nchannels, 200, replace=False)).astype(np.int32)
    _say("getSystemMatrix keep_zeros")
    sme, vxl, cnl, n = ctx.getSystemMatrix(lst, capacity=len(lst) * g.nvoxels, keep_zeros=True)
    orp, ocol, oval = oracle.sm_rows(s, g, lst, ctx.nrays, ctx.seed, keep_zeros=True)
    assert n == len(ocol)
    np.testing.assert_array_equal(vxl.cpu().numpy(), ocol)
    np.testing.assert_array_equal(cnl.cpu().numpy(), np.repeat(lst, np.diff(orp)))
    assert np.array_equal(sme.cpu().numpy(), oval)
    # the workqueue of the same list is the same candidate structure
    c2, v2, _, _ = ctx.getWorkqueue(lst)
    np.testing.assert_array_equal(v2.cpu().numpy(), ocol)


def test_sm_from_workqueue(gpu_ctx_factory, oracle):
    ctx = gpu_ctx_factory("small", seed=5)
    s, g = ctx.setup, ctx.grid
    lst = np.arange(0, s.nchannels, 97, dtype=np.int32)
    cnl, vxl, _, _ = ctx.getWorkqueue(lst)
    _say("getSystemMatrixFromWorkqueue", cnl.numel())
    val = ctx.getSystemMatrixFromWorkqueue(cnl, vxl, 8).cpu().numpy()
    ref = oracle.sm_from_workqueue(s, g, cnl.cpu().numpy(), vxl.cpu().numpy(), 8, ctx.seed)
    assert np.array_equal(val, ref)


def test_system_matrix_calculation_chain(gpu_ctx_factory, oracle):
    """systemMatrixCalculation (wrappers.hpp:172-247): ECSR of listed rows."""
    ctx = gpu_ctx_factory("small", seed=9)
    s, g = ctx.setup, ctx.grid
    rng = np.random.default_rng(13)
    lst = np.sort(rng.choice(s.nchannels, 120, replace=False)).astype(np.int32)
    _say("systemMatrixCalculation")
    ecsr, vxl, val, n = ctx.systemMatrixCalculation(lst, s.nchannels,
                                                    capacity=len(lst) * g.nvoxels)
    orp, ocol, oval = oracle.sm_rows(s, g, lst, ctx.nrays, ctx.seed, keep_zeros=True)
    np.testing.assert_array_equal(ecsr.cpu().numpy().astype(np.int64), orp)
    np.testing.assert_array_equal(vxl.cpu().numpy(), ocol)
    assert np.array_equal(val.cpu().numpy(), oval)
