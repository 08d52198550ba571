"""GPU MLEM parity on the small workload (11,250 channels, 16^3, 8 rays).

Bar (tolerances written here): after 10 iterations the GPU image is within
1e-4 relative L2 of the FP64 oracle and of the FP32 restatement of the
reference loop (tests/reco.cu:209-300) run on the same matrix and data.
"""
import numpy as np
import pytest
import torch

from epetreco_amd import phantom

pytestmark = pytest.mark.gpu

REL_L2 = 1e-4


def _say(*a):
    print(*a, flush=True)


def rel_l2(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / np.linalg.norm(b))


@pytest.fixture(scope="module")
def small_problem(gpu_ctx_factory):
    ctx = gpu_ctx_factory("small", seed=0)
    s, g = ctx.setup, ctx.grid
    _say("building small system matrix")
    A_full = ctx.sm_build(None, m=s.nchannels)
    prob = phantom.make_problem(ctx, A_full, "small", seed=0)
    return ctx, prob


@pytest.mark.parametrize("use_graph", [False, True])
def test_mlem_small_vs_oracles(small_problem, oracle, use_graph):
    ctx, prob = small_problem
    A = prob.A
    rp, col, val, _ = A.to_torch()
    rp, col, val = rp.cpu().numpy(), col.cpu().numpy(), val.cpu().numpy()
    y = prob.y.cpu().numpy()
    s = prob.s.cpu().numpy()
    V = A.n
    x0 = np.full(V, 1.0 / V, dtype=np.float32)
    x = torch.as_tensor(x0, device="cuda").clone()
    _say("mlem small graph=", use_graph)
    ll, it = ctx.mlem(A, prob.y, prob.s, x, 10, use_graph=use_graph)
    xg = x.cpu().numpy()
    x64 = oracle.mlem_f64(A.m, V, val, rp, col, y, s, x0.astype(np.float64), 10)
    x32 = oracle.mlem_f32(A.m, V, val, rp, col, y, s, x0, 10)
    assert rel_l2(xg, x64) < REL_L2
    assert rel_l2(xg, x32) < REL_L2
    assert abs(xg.sum() - 1.0) < 1e-5
    assert (xg >= 0).all()
    # Poisson log-likelihood is nondecreasing under MLEM
    assert np.all(np.diff(ll[:, 0]) >= -1e-9 * np.abs(ll[:-1, 0]))


def test_mlem_identity_one_step(gpu_ctx_factory):
    """A = I: one iteration from a uniform image gives y / sum(y)."""
    from epetreco_amd.api import Matrix
    ctx = gpu_ctx_factory()
    n = 1000
    rp = torch.arange(n + 1, dtype=torch.int64)
    col = torch.arange(n, dtype=torch.int32)
    val = torch.ones(n)
    A = Matrix.from_csr(ctx, rp, col, val, n)
    y = torch.rand(n, device="cuda") + 0.5
    s = torch.ones(n, device="cuda")
    x = torch.full((n,), 1.0 / n, device="cuda")
    ctx.mlem(A, y, s, x, 1)
    np.testing.assert_allclose(x.cpu().numpy(), (y / y.sum()).cpu().numpy(), rtol=1e-5)


def test_mlem_batched_matches_single(small_problem):
    """F frames in one launch == F single-frame runs (up to FP32 order)."""
    ctx, prob = small_problem
    A = prob.A
    F = 5
    g = torch.Generator(device="cpu").manual_seed(3)
    Y = torch.poisson(prob.y.cpu()[:, None].expand(-1, F) * torch.rand(F, generator=g) * 2,
                      generator=g).to("cuda").contiguous()
    X = torch.full((A.n, F), 1.0 / A.n, device="cuda")
    ctx.mlem(A, Y, prob.s, X, 6)
    for f in range(F):
        x = torch.full((A.n,), 1.0 / A.n, device="cuda")
        ctx.mlem(A, Y[:, f].contiguous(), prob.s, x, 6)
        assert rel_l2(X[:, f].cpu().numpy(), x.cpu().numpy()) < 1e-5
