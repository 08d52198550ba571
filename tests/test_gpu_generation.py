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


@pytest.mark.timeout(600)
@pytest.mark.parametrize("wl,nrows", [("full", 1000), ("large", 200)])
def test_sm_build_sampled_rows(gpu_ctx_factory, oracle, wl, nrows):
    """Seeded channel samples of the 52^3 / 1 mm (64 rays) and the 104^3 /
    0.5 mm (32 rays) workloads, bit-exact against the oracle's full voxel scan
    (every voxel tested with the reference predicate, every candidate's value
    evaluated): the ray traversal's voxel skipping must lose no nonzero."""
    ctx = gpu_ctx_factory(wl, seed=0)
    s, g = ctx.setup, ctx.grid
    rng = np.random.default_rng(7)
    lst = np.sort(rng.choice(s.nchannels, nrows, replace=False)).astype(np.int32)
    _say(f"sm_build {wl}, {nrows} sampled rows")
    A = ctx.sm_build(lst)
    rp, col, val, _ = A.to_torch()
    orp, ocol, oval = oracle.sm_rows(s, g, lst, ctx.nrays, ctx.seed, keep_zeros=False)
    _say(f"  {A.nnz} nonzeros")
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


def test_get_chords_matches_oracle(gpu_ctx_factory, oracle):
    """Device Siddon chords (getChords, ChordsCalc_lowlevel.hpp:173-357) are
    bit-exact with the oracle's restatement, quirks included, on: detector rays
    of 200 channels, 300 random rays between points outside the grid,
    axis-parallel rays, and rays with an endpoint inside the grid (no chords);
    a short capacity gives -1."""
    import ctypes
    import torch
    ctx = gpu_ctx_factory("small", seed=0)
    g = ctx.grid
    rng = np.random.default_rng(9)
    cnl = torch.as_tensor(rng.choice(ctx.setup.nchannels, 200, replace=False).astype(np.int32),
                          device="cuda")
    det = ctx.rays(cnl).reshape(-1, 6).cpu().numpy()[::4]          # a quarter of the rays
    c = np.array([g.ox + g.nx * g.dx / 2, g.oy + g.ny * g.dy / 2, g.oz + g.nz * g.dz / 2])
    u = rng.normal(size=(300, 2, 3))
    u /= np.linalg.norm(u, axis=2, keepdims=True)
    rnd = (c + 0.06 * u).reshape(300, 6)
    axis = []
    for d in range(3):
        p = c + rng.uniform(-0.02, 0.02, 3)
        a, b = p.copy(), p.copy()
        a[d], b[d] = c[d] - 0.1, c[d] + 0.1
        axis.append(np.concatenate([a, b]))
        axis.append(np.concatenate([b, a]))
    inside = np.concatenate([c + 0.001, c + 0.2])[None, :]
    rays = np.concatenate([det, rnd, np.array(axis), inside]).astype(np.float32)
    n = len(rays)
    cap = 64
    ch, vid, cnt = ctx.getChords(torch.as_tensor(rays, device="cuda"), cap)
    ch, vid, cnt = ch.cpu().numpy(), vid.cpu().numpy(), cnt.cpu().numpy()
    L = oracle.lib()
    O = np.array([g.ox, g.oy, g.oz], np.float32)
    D = np.array([g.dx, g.dy, g.dz], np.float32)
    N = np.array([g.nx, g.ny, g.nz], np.int32)
    nonzero = 0
    for r in range(n):
        och = np.zeros(cap, np.float32)
        oid = np.zeros(3 * cap, np.int32)
        k = L.eo_get_chords(oracle.ptr(och, ctypes.c_float), oracle.ptr(oid, ctypes.c_int), cap,
                            oracle.ptr(rays[r], ctypes.c_float), oracle.ptr(O, ctypes.c_float),
                            oracle.ptr(D, ctypes.c_float), oracle.ptr(N, ctypes.c_int))
        assert cnt[r] == k, r
        if k > 0:
            nonzero += 1
            assert np.array_equal(ch[r, :k], och[:k]), r
            assert np.array_equal(vid[r, :k].reshape(-1), oid[:3 * k]), r
    assert nonzero > 200 and cnt[-1] == 0
    _, _, cnt5 = ctx.getChords(torch.as_tensor(rays, device="cuda"), 5)
    cnt5 = cnt5.cpu().numpy()
    assert np.all((cnt5 == -1) == (cnt > 5)) and np.all(cnt5[cnt <= 5] == cnt[cnt <= 5])
