"""Batched multi-frame MLEM parity.

Per-frame stop (small workload: 11,250 channels, 16^3, 8 rays; 32 frames of
the dynamic series, totals log-uniform in [1e5, 1e7]): frames stop on their own
when the relative log-likelihood change is <= 1e-6, or after 50 iterations,
in one captured graph per iteration.  Each frame must equal the single-frame
run of its own iteration count (relative L2 <= 1e-5) and the FP64 oracle at
that count (relative L2 <= 1e-4).

Reduced-precision storage (104^3 grid of 0.5 mm voxels, 32 rays per channel,
13x13-segment heads): every 8th channel of angles 0 and 90 deg (7,142
channels, so that the FP64 oracle runs in seconds), 8 frames with totals
log-uniform in [1e6, 1e7], 30 iterations, values stored BF16 and FP16 with a
per-row FP32 scale.  Two figures per frame:
  * kernel error: GPU image vs the FP64 oracle run on the same stored
    operator (decoded values times row scale): <= 1e-4, as for FP32 storage;
  * storage error: GPU image vs the FP64 oracle on the FP32 matrix.  It is
    what the value format costs this operator, so it must agree with the
    oracle's own format cost (FP64 MLEM on the stored operator vs on the FP32
    matrix) to within 10% + 1e-4.  MLEM amplifies value rounding by a factor
    that shrinks with the entries per voxel column: this subset has ~18, so
    BF16 costs ~1.5e-2 here.  The absolute bar (1e-3 against FP64) is held at
    the full 52^3 operator (~2e4 entries per column) in test_gpu_full.py,
    where FP16 storage measures <= 1.1e-4 and BF16 up to 1.2e-3 after 20
    iterations.
"""
import numpy as np
import pytest
import torch

from epetreco_amd import phantom

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _frames(ctx, A, name, F, seed, counts):
    from epetreco_amd.api import OP_N
    imgs, totals = phantom.dynamic_series(ctx.grid, name, F, seed=seed, counts=counts)
    X = torch.as_tensor(imgs.T.astype(np.float32), device="cuda").contiguous()
    Yb = A.spmm(OP_N, X).double()
    Yb = Yb * torch.as_tensor(totals / Yb.sum(0).cpu().numpy(), device="cuda")
    rng = np.random.default_rng(seed + 5)
    return torch.as_tensor(rng.poisson(Yb.cpu().numpy()).astype(np.float32), device="cuda")


def test_per_frame_stop(gpu_ctx_factory, oracle):
    ctx = gpu_ctx_factory("small", seed=0)
    s = ctx.setup
    A = ctx.sm_build(None, m=s.nchannels).attach_transpose()
    sens = A.colsum_f64().float()
    F = 32
    Y = _frames(ctx, A, "small", F, 0, (1e5, 1e7))
    V = A.n
    X = torch.full((V, F), 1.0 / V, device="cuda")
    _say("per-frame stop, 32 frames")
    ll, iters = ctx.mlem(A, Y, sens, X, 50, use_graph=True, tol=1e-6)
    _say("iterations per frame", iters.tolist())
    assert iters.min() >= 2 and iters.max() <= 50
    assert iters.min() < 50            # the stop rule fires for some frame
    rp, col, val, _ = A.to_torch()
    rp, col, val = rp.cpu().numpy(), col.cpu().numpy(), val.cpu().numpy()
    Xh = X.cpu().numpy()
    x64 = oracle.mlem_f64_frames(A.m, V, val, rp, col, Y.cpu().numpy(), sens.cpu().numpy(),
                                 np.full((V, F), 1.0 / V), iters)
    for f in range(F):
        xs = torch.full((V,), 1.0 / V, device="cuda")
        ctx.mlem(A, Y[:, f].contiguous(), sens, xs, int(iters[f]), use_graph=True)
        assert rel_l2(Xh[:, f], xs.cpu().numpy().astype(np.float64)) < 1e-5, f
        assert rel_l2(Xh[:, f], x64[:, f]) < 1e-4, f
    # the rule itself: a stopped frame's last change is within tolerance, and
    # the change before it was not
    for f in range(F):
        k = int(iters[f])
        if k < 50:
            a, b = ll[k - 1, f], ll[k - 2, f]
            assert abs(a - b) <= 1e-6 * abs(b)
            if k >= 3:
                assert abs(b - ll[k - 3, f]) > 1e-6 * abs(ll[k - 3, f])


@pytest.fixture(scope="module")
def large_problem(gpu_ctx_factory):
    ctx = gpu_ctx_factory("large", seed=0)
    s = ctx.setup
    per_angle = s.nchannels // s.na
    a90 = int(round(90.0 / s.da))
    lst = np.concatenate([np.arange(0, per_angle, 8),
                          a90 * per_angle + np.arange(0, per_angle, 8)]).astype(np.int32)
    _say("104^3 workload: building", len(lst), "channels x", ctx.grid.nvoxels, "voxels")
    A = ctx.sm_build(lst)
    _say("  nonzeros", A.nnz)
    F = 8
    Y = _frames(ctx, A, "large", F, 3, (1e6, 1e7))
    keep = np.nonzero(Y.sum(1).cpu().numpy() > 0)[0].astype(np.int32)
    sens = A.colsum_f64().float()
    A = A.select_rows(keep)
    Y = Y[torch.as_tensor(keep, device="cuda").long()].contiguous()
    rp, col, val, _ = A.to_torch()
    rp, col, val = rp.cpu().numpy(), col.cpu().numpy(), val.cpu().numpy()
    V = A.n
    _say("  FP64 oracle, 8 frames x 30 iterations over", A.nnz, "nonzeros")
    x64 = oracle_mlem(A.m, V, val, rp, col, Y.cpu().numpy(), sens.cpu().numpy(),
                      np.full((V, F), 1.0 / V), 30)
    return ctx, A, Y, sens, x64


def oracle_mlem(*a):
    from oracle import oracle as O
    return O.mlem_f64_frames(*a)


@pytest.mark.parametrize("storage", ["f32", "bf16", "f16"])
def test_reduced_precision_storage_104(large_problem, storage):
    from epetreco_amd.api import VAL_BF16, VAL_F16
    ctx, A, Y, sens, x64 = large_problem
    V, F = A.n, Y.shape[1]
    M = A if storage == "f32" else A.convert({"bf16": VAL_BF16, "f16": VAL_F16}[storage])
    M.attach_transpose()
    X = torch.full((V, F), 1.0 / V, device="cuda")
    ctx.mlem(M, Y, sens, X, 30, use_graph=True)
    Xh = X.cpu().numpy()
    e_gpu = np.array([rel_l2(Xh[:, f], x64[:, f]) for f in range(F)])
    _say(f"{storage} storage, per-frame rel-L2 vs FP64 oracle (FP32 matrix):",
         " ".join(f"{e:.2e}" for e in e_gpu))
    if storage == "f32":
        assert e_gpu.max() < 1e-4
        return
    from helpers import decoded
    rp, col, vq = decoded(M)
    xq = oracle_mlem(A.m, V, vq, rp, col, Y.cpu().numpy(), sens.cpu().numpy(),
                     np.full((V, F), 1.0 / V), 30)
    e_kernel = np.array([rel_l2(Xh[:, f], xq[:, f]) for f in range(F)])
    e_fmt = np.array([rel_l2(xq[:, f], x64[:, f]) for f in range(F)])
    _say(f"{storage}: kernel error (vs FP64 on the stored operator):",
         " ".join(f"{e:.2e}" for e in e_kernel))
    _say(f"{storage}: format cost in the oracle:", " ".join(f"{e:.2e}" for e in e_fmt))
    assert e_kernel.max() < 1e-4
    assert np.all(np.abs(e_gpu - e_fmt) <= 0.1 * e_fmt + 1e-4)
