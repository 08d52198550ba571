"""Differentiable projector workload: small-scanner geometry (11,250 channels)
with a 52^3 grid of 1 mm voxels, loss = sum(Ax - y log(Ax + 1e-8)) at a random
positive image x ~ U(0.5, 1.5).

Checks (tolerances written here):
  * autograd gradient vs FP64 central differences (h = 1e-3) on 1,000 seeded
    voxel indices: relative error <= 1e-3 (FP32 forward of ~1e3-term rows);
  * gradient vs the reference back projection A^T (1 - y/(Ax+1e-8)) computed by
    the oracle's CSRmv-T semantics in FP64: relative L2 <= 1e-5;
  * 50 steps of projected gradient descent decrease the loss, and MLEM reaches
    the same likelihood (at its optimal image scale) within 50 iterations.
"""
import numpy as np
import pytest
import scipy.sparse as sp
import torch

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def problem(gpu_ctx_factory):
    from epetreco_amd import geometry as G
    from epetreco_amd import phantom as PH
    ctx = gpu_ctx_factory("projector", seed=0)
    s, g = ctx.setup, ctx.grid
    _say("projector workload: building", s.nchannels, "x", g.nvoxels)
    A = ctx.sm_build(None, m=s.nchannels)
    xt = torch.as_tensor(PH.phantom(g, "projector", 0).astype(np.float32), device="cuda")
    from epetreco_amd.api import OP_N
    ybar = A.spmv(OP_N, xt).double()
    ybar = (ybar * (1e5 / ybar.sum())).cpu().numpy()
    y = torch.as_tensor(PH.poisson_counts(ybar, 3), device="cuda")
    rp, col, val, _ = A.to_torch()
    M = sp.csr_matrix((val.cpu().numpy().astype(np.float64), col.cpu().numpy(), rp.cpu().numpy()),
                      shape=(A.m, A.n))
    return ctx, A, y, M


def test_gradient_vs_central_differences(problem):
    from epetreco_amd.projector import poisson_nll
    ctx, A, y, M = problem
    rng = np.random.default_rng(0)
    x = torch.as_tensor(rng.uniform(0.5, 1.5, A.n).astype(np.float32), device="cuda")
    xg = x.clone().requires_grad_(True)
    loss = poisson_nll(A, xg, y)
    (grad,) = torch.autograd.grad(loss, xg)
    grad = grad.cpu().numpy().astype(np.float64)
    # FP64 central differences: only the rows of column j change
    x64 = x.cpu().numpy().astype(np.float64)
    y64 = y.cpu().numpy().astype(np.float64)
    yh = M @ x64
    Mc = M.tocsc()
    h = 1e-3
    idx = rng.choice(A.n, 1000, replace=False)
    f = lambda v, yy: v - yy * np.log(v + 1e-8)
    errs = []
    for j in idx:
        rows = Mc.indices[Mc.indptr[j]:Mc.indptr[j + 1]]
        a = Mc.data[Mc.indptr[j]:Mc.indptr[j + 1]]
        if len(rows) == 0:
            assert abs(grad[j]) < 1e-6
            continue
        base = yh[rows]
        d = (f(base + h * a, y64[rows]).sum() - f(base - h * a, y64[rows]).sum()) / (2 * h)
        errs.append(abs(grad[j] - d) / max(abs(d), 1e-3 * np.abs(grad).max()))
    _say("max rel err of autograd vs FP64 central differences", max(errs))
    assert max(errs) < 1e-3


def test_gradient_vs_reference_back_projection(problem, oracle):
    from epetreco_amd.projector import poisson_nll
    ctx, A, y, M = problem
    rng = np.random.default_rng(1)
    x = torch.as_tensor(rng.uniform(0.5, 1.5, A.n).astype(np.float32), device="cuda")
    xg = x.clone().requires_grad_(True)
    (grad,) = torch.autograd.grad(poisson_nll(A, xg, y), xg)
    yh = M @ x.cpu().numpy().astype(np.float64)
    w = (1.0 - y.cpu().numpy() / (yh + 1e-8)).astype(np.float32)
    rp = M.indptr.astype(np.int64)
    ref = oracle.csrmv(1, A.m, A.n, 1.0, M.data.astype(np.float32), rp, M.indices.astype(np.int32),
                       w, 0.0, np.zeros(A.n, np.float32))
    err = np.linalg.norm(grad.cpu().numpy() - ref) / np.linalg.norm(ref)
    assert err < 1e-5


def test_projected_gradient_descent_vs_mlem(problem):
    from epetreco_amd.projector import poisson_nll, projected_gradient_descent
    ctx, A, y, M = problem
    x0 = torch.full((A.n,), float(y.sum()) / float(M.sum()), device="cuda")
    x, losses = projected_gradient_descent(A, y, x0, 50)
    assert all(b <= a + 1e-6 * abs(a) for a, b in zip(losses, losses[1:]))
    target = losses[-1]
    s = A.colsum_f64().float()
    xm = torch.full((A.n,), 1.0 / A.n, device="cuda")
    reached = None
    for k in range(1, 51):
        ctx.mlem(A, y, s, xm, 1)
        scale = float(y.sum()) / float(A.spmv(0, xm).sum())
        Lk = float(poisson_nll(A, xm * scale, y))
        if Lk <= target:
            reached = k
            break
    _say("PGD loss after 50 steps", target, "; MLEM reaches it after", reached, "iterations")
    assert reached is not None
