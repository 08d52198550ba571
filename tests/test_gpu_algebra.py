"""GPU parity of the sparse formats, SpMV/SpMM and the MLEM operations."""
import json
import os

import numpy as np
import pytest
import torch

from conftest import GOLDEN

pytestmark = pytest.mark.gpu

KATS = json.load(open(os.path.join(GOLDEN, "ref_kats.json")))


def _say(*a):
    print(*a, flush=True)


def _t(a, dtype):
    return torch.as_tensor(np.asarray(a), dtype=dtype, device="cuda")


def test_coo_sort_kat(gpu_ctx_factory):
    ctx = gpu_ctx_factory()
    k = KATS["cooSort"]
    row, col, val = _t(k["row"], torch.int32), _t(k["col"], torch.int32), _t(k["val"], torch.float32)
    ctx.cooSort(val, row, col)
    assert row.tolist() == k["expect_row"]
    assert col.tolist() == k["expect_col"]
    assert val.tolist() == k["expect_val"]


@pytest.mark.parametrize("n", [1, 2, 2047, 2048, 2049, 300001])
def test_coo_sort_random_stable(gpu_ctx_factory, n):
    ctx = gpu_ctx_factory()
    rng = np.random.default_rng(n)
    row = rng.integers(-5, 3000, n).astype(np.int32)
    col = rng.integers(-2, 70000, n).astype(np.int32)
    val = np.arange(n, dtype=np.float32)   # original position: checks stability
    tr, tc, tv = _t(row, torch.int32), _t(col, torch.int32), _t(val, torch.float32)
    _say("cooSort", n)
    ctx.cooSort(tv, tr, tc)
    order = np.lexsort((np.arange(n), col, row))
    np.testing.assert_array_equal(tr.cpu().numpy(), row[order])
    np.testing.assert_array_equal(tc.cpu().numpy(), col[order])
    np.testing.assert_array_equal(tv.cpu().numpy(), val[order])


def test_coo2csr_kat(gpu_ctx_factory):
    ctx = gpu_ctx_factory()
    k = KATS["coo2csr"]
    out = ctx.convertCoo2Csr(_t(k["row"], torch.int32), k["m"])
    assert out.tolist() == k["expect"]


def test_csr2ecsr_kat(gpu_ctx_factory):
    ctx = gpu_ctx_factory()
    k = KATS["csr2ecsr"]
    out = ctx.convertCsr2Ecsr(_t(k["yRowId"], torch.int32), _t(k["csr"], torch.int32))
    assert out.tolist() == k["expect"]


def test_mlem_operations_kat(gpu_ctx_factory):
    ctx = gpu_ctx_factory()
    k = KATS["mlemOperations"]
    A, B, C = (_t(k[n], torch.float32) for n in "ABC")
    D = torch.zeros_like(A)
    ctx.divides(D, A, C)
    assert D.tolist() == k["expect_divides"]
    ctx.dividesMultiplies(D, A, B, C)
    assert D.tolist() == k["expect_dividesMultiplies"]
    norm = ctx.sum(D)
    assert norm == k["expect_sum"]
    ctx.scales(D, 1.0 / norm)
    np.testing.assert_allclose(D.cpu().numpy(), k["expect_scaled"], rtol=1e-6)


def _random_csr(rng, m, n, avg):
    lens = rng.poisson(avg, m)
    rp = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    col = np.concatenate([np.sort(rng.choice(n, l, replace=False)) for l in lens]).astype(np.int32)
    val = rng.random(rp[-1]).astype(np.float32)
    return rp, col, val


@pytest.mark.parametrize("avg", [3, 30, 300, 1500])
def test_spmv_against_fp64(gpu_ctx_factory, oracle, avg):
    from epetreco_amd.api import Matrix, OP_N, OP_T
    ctx = gpu_ctx_factory()
    rng = np.random.default_rng(avg)
    m, n = 3000, 5000
    rp, col, val = _random_csr(rng, m, n, avg)
    A = Matrix.from_csr(ctx, torch.as_tensor(rp), torch.as_tensor(col), torch.as_tensor(val), n)
    x = rng.random(n).astype(np.float32)
    e = rng.random(m).astype(np.float32)
    _say("spmv", avg, A.nnz)
    y = A.spmv(OP_N, _t(x, torch.float32)).cpu().numpy()
    yr = oracle.csrmv(0, m, n, 1.0, val, rp, col, x, 0.0, np.zeros(m, np.float32))
    np.testing.assert_allclose(y, yr, rtol=2e-5, atol=1e-6)
    c = A.spmv(OP_T, _t(e, torch.float32), alpha=2.0).cpu().numpy()
    cr = oracle.csrmv(1, m, n, 2.0, val, rp, col, e, 0.0, np.zeros(n, np.float32))
    np.testing.assert_allclose(c, cr, rtol=2e-5, atol=1e-6)
    # int32 CSRmv drop-in, beta = 1 accumulation (tests/reco.cu:247-249)
    y0 = rng.random(n).astype(np.float32)
    yt = _t(y0, torch.float32)
    ctx.csrmv(OP_T, m, n, 1.0, _t(val, torch.float32), _t(rp.astype(np.int32), torch.int32),
              _t(col, torch.int32), _t(e, torch.float32), 1.0, yt)
    cr2 = oracle.csrmv(1, m, n, 1.0, val, rp, col, e, 1.0, y0)
    np.testing.assert_allclose(yt.cpu().numpy(), cr2, rtol=2e-5, atol=1e-6)


@pytest.mark.parametrize("F", [1, 3, 8, 32])
def test_spmm_against_fp64(gpu_ctx_factory, oracle, F):
    from epetreco_amd.api import Matrix, OP_N, OP_T
    ctx = gpu_ctx_factory()
    rng = np.random.default_rng(F)
    m, n = 2000, 4000
    rp, col, val = _random_csr(rng, m, n, 200)
    A = Matrix.from_csr(ctx, torch.as_tensor(rp), torch.as_tensor(col), torch.as_tensor(val), n)
    X = rng.random((n, F)).astype(np.float32)
    E = rng.random((m, F)).astype(np.float32)
    Y = A.spmm(OP_N, _t(X, torch.float32)).cpu().numpy()
    C = A.spmm(OP_T, _t(E, torch.float32)).cpu().numpy()
    for f in range(F):
        yr = oracle.csrmv(0, m, n, 1.0, val, rp, col, X[:, f].copy(), 0.0, np.zeros(m, np.float32))
        cr = oracle.csrmv(1, m, n, 1.0, val, rp, col, E[:, f].copy(), 0.0, np.zeros(n, np.float32))
        np.testing.assert_allclose(Y[:, f], yr, rtol=2e-5, atol=1e-6)
        np.testing.assert_allclose(C[:, f], cr, rtol=2e-5, atol=1e-6)


@pytest.mark.parametrize("n,avg", [(40000, 900), (70001, 200), (16384, 50)])
def test_spmv_multi_tile_back_projection(gpu_ctx_factory, oracle, n, avg):
    """Back projection across several 16384-voxel LDS tiles, and the vector forward."""
    from epetreco_amd.api import Matrix, OP_N, OP_T
    ctx = gpu_ctx_factory()
    rng = np.random.default_rng(n + avg)
    m = 2500
    rp, col, val = _random_csr(rng, m, n, avg)
    A = Matrix.from_csr(ctx, torch.as_tensor(rp), torch.as_tensor(col), torch.as_tensor(val), n)
    x = rng.random(n).astype(np.float32)
    e = rng.random(m).astype(np.float32)
    e[::7] = 0.0                       # rows skipped by the back projection
    y0 = rng.random(n).astype(np.float32)
    _say("multi-tile", n, avg, A.nnz)
    y = A.spmv(OP_N, _t(x, torch.float32)).cpu().numpy()
    yr = oracle.csrmv(0, m, n, 1.0, val, rp, col, x, 0.0, np.zeros(m, np.float32))
    np.testing.assert_allclose(y, yr, rtol=2e-5, atol=1e-6)
    for beta in (0.0, 1.0, 0.5):
        c = A.spmv(OP_T, _t(e, torch.float32), _t(y0, torch.float32).clone(), alpha=1.5,
                   beta=beta).cpu().numpy()
        cr = oracle.csrmv(1, m, n, 1.5, val, rp, col, e, beta, y0)
        np.testing.assert_allclose(c, cr, rtol=2e-5, atol=1e-6)


@pytest.mark.parametrize("m,n,avg", [(60000, 3000, 100), (50001, 777, 40), (9000, 1000, 70)])
def test_blocked_transposed_back_projection(gpu_ctx_factory, oracle, monkeypatch, m, n, avg):
    """Back projection through the transposed copy with long transposed rows
    (>= 512 entries: the column-blocked kernel), with ragged block boundaries
    (m not a multiple of 8) and empty rows, vs the FP64 oracle and the plain
    row kernel, for beta 0, 1 and 0.5."""
    from epetreco_amd.api import Matrix, OP_T
    ctx = gpu_ctx_factory()
    rng = np.random.default_rng(m + n)
    rp, col, val = _random_csr(rng, m, n, avg)
    A = Matrix.from_csr(ctx, torch.as_tensor(rp), torch.as_tensor(col), torch.as_tensor(val), n)
    A.attach_transpose()
    _say("blocked back projection", m, n, A.nnz, "avg transposed row", A.nnz / n)
    e = rng.random(m).astype(np.float32)
    e[::5] = 0.0
    y0 = rng.random(n).astype(np.float32)
    for beta in (0.0, 1.0, 0.5):
        c = A.spmv(OP_T, _t(e, torch.float32), _t(y0, torch.float32).clone(), alpha=1.5,
                   beta=beta).cpu().numpy()
        cr = oracle.csrmv(1, m, n, 1.5, val, rp, col, e, beta, y0)
        np.testing.assert_allclose(c, cr, rtol=2e-5, atol=1e-6)
    from helpers import env_ctx
    ctx_rows = env_ctx(monkeypatch, EPET_BP="rows")
    c_rows = A.spmv(OP_T, _t(e, torch.float32), ctx=ctx_rows).cpu().numpy()
    c_blk = A.spmv(OP_T, _t(e, torch.float32)).cpu().numpy()
    np.testing.assert_allclose(c_blk, c_rows, rtol=2e-6, atol=1e-7)


def test_unsorted_columns_fall_back(gpu_ctx_factory, oracle):
    """Columns not ascending within rows: still correct (atomic back projection)."""
    from epetreco_amd.api import Matrix, OP_T
    ctx = gpu_ctx_factory()
    rng = np.random.default_rng(1)
    m, n = 500, 30000
    rp, col, val = _random_csr(rng, m, n, 100)
    for i in range(m):
        seg = col[rp[i]:rp[i + 1]]
        rng.shuffle(seg)
    A = Matrix.from_csr(ctx, torch.as_tensor(rp), torch.as_tensor(col), torch.as_tensor(val), n)
    e = rng.random(m).astype(np.float32)
    c = A.spmv(OP_T, _t(e, torch.float32)).cpu().numpy()
    cr = oracle.csrmv(1, m, n, 1.0, val, rp, col, e, 0.0, np.zeros(n, np.float32))
    np.testing.assert_allclose(c, cr, rtol=2e-5, atol=1e-6)


def test_invalid_csr_rejected(gpu_ctx_factory):
    from epetreco_amd.api import EpetError, Matrix
    ctx = gpu_ctx_factory()
    rp = torch.tensor([0, 2, 3], dtype=torch.int64)
    col = torch.tensor([0, 5, 1], dtype=torch.int32)     # 5 >= n
    with pytest.raises(EpetError):
        Matrix.from_csr(ctx, rp, col, torch.ones(3), 4)


def test_error_paths_are_loud(gpu_ctx_factory):
    """Invalid calls return error codes with a message (raised as EpetError),
    never a silent result: too many frames, a null image, a negative chord
    capacity, tuning without the transposed copy."""
    import ctypes
    from epetreco_amd._lib import EpetError, lib
    from epetreco_amd.api import Matrix
    ctx = gpu_ctx_factory("small", seed=0)
    rng = np.random.default_rng(2)
    rp, col, val = _random_csr(rng, 200, 300, 20)
    A = Matrix.from_csr(ctx, torch.as_tensor(rp), torch.as_tensor(col), torch.as_tensor(val), 300)
    y = torch.ones(200, 65, device="cuda")
    s = torch.ones(300, device="cuda")
    X = torch.full((300, 65), 1.0 / 300, device="cuda")
    with pytest.raises(EpetError, match="frames"):
        ctx.mlem(A, y, s, X, 1)
    y1 = _t(np.ones(200), torch.float32)
    rc = lib().epet_mlem(ctx.handle, A.handle, 1, y1.data_ptr(), s.data_ptr(), None, 1, 0, 0.0,
                         None, None)
    assert rc != 0 and "null" in lib().epet_last_error().decode()
    with pytest.raises(EpetError, match="negative"):
        ctx.getChords(torch.zeros(1, 6, device="cuda"), -1)
    with pytest.raises(EpetError, match="transpose"):
        ctx.mlem_tune(A, 8)
    # host-side argument checks: dtype, frame layout and sizes against the matrix
    x1 = torch.full((300,), 1.0 / 300, device="cuda")
    with pytest.raises(EpetError, match="float32"):
        ctx.mlem(A, y1, s.double(), x1, 1)
    with pytest.raises(EpetError, match="frames"):
        ctx.mlem(A, y1, s, torch.full((300, 2), 1.0 / 300, device="cuda"), 1)
    with pytest.raises(EpetError, match="elements"):
        ctx.mlem(A, y1[:150].contiguous(), s, x1, 1)
    with pytest.raises(EpetError, match="elements"):
        A.spmv(0, torch.ones(200, device="cuda"))
    with pytest.raises(EpetError, match="elements"):
        A.spmv(1, torch.ones(300, device="cuda"))
    with pytest.raises(EpetError, match="elements"):
        A.spmm(0, torch.ones(200, 4, device="cuda"))
    with pytest.raises(EpetError, match="frames"):
        A.spmm(0, torch.ones(300, 4, device="cuda"), torch.zeros(200, 3, device="cuda"))
