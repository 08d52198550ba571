"""The allocation-free single-frame MLEM entry point (epet_mlem_noalloc)
against epet_mlem on the small workload (11,250 channels, 16^3, 8 rays).

Bar (tolerances written here), three calls in a row sharing the GPU's fixed
working array:
1. the measured-channel matrix, after epet_mlem has built its product plans
   (both calls then run the same kernels): images equal within rel 1e-6 per
   voxel, log-likelihoods within rel 1e-7 (the FP64 image sums are atomics,
   so 1 / sum x can round to a neighbouring FP32 value, which moves the FP32
   forward products by an ulp);
2. a smaller input (the first half of the rows, a fresh matrix with no plans:
   the plan-free kernels run): within 1e-5 relative L2 of epet_mlem on the
   same matrix, and within 1e-4 of the FP64 oracle;
3. a still smaller one (A = I, n = 1000): one iteration gives y / sum(y).
Sizes beyond the fixed array fail with EPET_ERR_CAPACITY and leave x alone.
"""
import numpy as np
import pytest
import torch

from epetreco_amd import phantom

pytestmark = pytest.mark.gpu

NITER = 10


def _say(*a):
    print(*a, flush=True)


def rel_l2(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / np.linalg.norm(b))


@pytest.fixture(scope="module")
def small(gpu_ctx_factory):
    ctx = gpu_ctx_factory("small", seed=0)
    A_full = ctx.sm_build(None, m=ctx.setup.nchannels)
    return ctx, phantom.make_problem(ctx, A_full, "small", seed=0)


def test_noalloc_matches_mlem_three_calls(small, oracle):
    from epetreco_amd.api import Matrix
    ctx, prob = small
    A = prob.A
    V = A.n
    x0 = torch.full((V,), 1.0 / V, device="cuda")

    _say("call 1: m =", A.m, "n =", V)
    xr = x0.clone()
    llr, _ = ctx.mlem(A, prob.y, prob.s, xr, NITER)
    xn = x0.clone()
    lln = ctx.mlem_noalloc(A, prob.y, prob.s, xn, NITER)
    np.testing.assert_allclose(xn.cpu().numpy(), xr.cpu().numpy(), rtol=1e-6, atol=1e-12)
    np.testing.assert_allclose(lln, llr[:, 0], rtol=1e-7)

    rows = torch.arange(A.m // 2, dtype=torch.int32)
    B = A.select_rows(rows)
    yb = prob.y[: A.m // 2].contiguous()
    sb = B.colsum_f64().float()            # the half system's own sensitivity
    _say("call 2: m =", B.m, "n =", V, "(no plans)")
    xn = x0.clone()
    lln = ctx.mlem_noalloc(B, yb, sb, xn, NITER)
    xr = x0.clone()
    llr, _ = ctx.mlem(B, yb, sb, xr, NITER)
    assert rel_l2(xn.cpu().numpy(), xr.cpu().numpy().astype(np.float64)) < 1e-5
    np.testing.assert_allclose(lln, llr[:, 0], rtol=1e-6)
    rp, col, val, _ = B.to_torch()
    x64 = oracle.mlem_f64(B.m, V, val.cpu().numpy(), rp.cpu().numpy(), col.cpu().numpy(),
                          yb.cpu().numpy(), sb.cpu().numpy(),
                          np.full(V, 1.0 / V), NITER)
    assert rel_l2(xn.cpu().numpy(), x64) < 1e-4

    n = 1000
    _say("call 3: identity, n =", n)
    I = Matrix.from_csr(ctx, torch.arange(n + 1, dtype=torch.int64),
                        torch.arange(n, dtype=torch.int32), torch.ones(n), n)
    y = torch.rand(n, device="cuda", generator=torch.Generator(device="cuda").manual_seed(5)) + 0.5
    x = torch.full((n,), 1.0 / n, device="cuda")
    ctx.mlem_noalloc(I, y, torch.ones(n, device="cuda"), x, 1)
    np.testing.assert_allclose(x.cpu().numpy(), (y / y.sum()).cpu().numpy(), rtol=1e-5)


def test_noalloc_rejects_sizes_beyond_the_array(small):
    from epetreco_amd.api import EpetError
    ctx, prob = small
    A = prob.A
    x = torch.full((A.n,), 1.0 / A.n, device="cuda")
    before = x.clone()
    niter = (64 << 20) // 8   # the log-likelihoods alone fill the array
    with pytest.raises(EpetError, match="EPET_ERR_CAPACITY.*limit is 67108864"):
        ctx.mlem_noalloc(A, prob.y, prob.s, x, niter)
    assert torch.equal(x, before)
