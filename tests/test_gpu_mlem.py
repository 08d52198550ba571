"""GPU MLEM parity on the small workload (11,250 channels, 16^3, 8 rays).

Bar (tolerances written here): after 10 iterations the GPU image is within
1e-4 relative L2 of the FP64 oracle and of the FP32 restatement of the
reference loop (tests/reco.cu:209-300) run on the same matrix and data.
"""
import numpy as np
import pytest
import torch

from epetreco_amd import phantom
from helpers import env_ctx

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


def test_sensitivity_small_vs_oracle(gpu_ctx_factory, oracle):
    """In-tree sensitivity (the FP64 column sums of the all-channel matrix)
    against the oracle's FP64 A^T 1 over its own all-channel matrix of the
    small workload (tests/backprojection.cu:115-161, tests/reco.cu:150-183)."""
    ctx = gpu_ctx_factory("small", seed=0)
    s, g = ctx.setup, ctx.grid
    A = ctx.sm_build(None, m=s.nchannels)
    s_dev = A.colsum_f64().cpu().numpy()
    orp, ocol, oval = oracle.sm_rows(s, g, np.arange(s.nchannels, dtype=np.int32),
                                     ctx.nrays, ctx.seed)
    s_ref = oracle.colsum_f64(g.nvoxels, oval, orp, ocol)
    np.testing.assert_allclose(s_dev, s_ref, rtol=1e-12, atol=0)
    assert np.array_equal(s_dev == 0, s_ref == 0)


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


def test_transpose_path_matches_atomic_path(small_problem):
    """One batched iteration through the attached transpose (gather back projection)
    equals the row-walk atomic back projection."""
    ctx, prob = small_problem
    from epetreco_amd.api import Matrix
    rp, col, val, _ = prob.A.to_torch()
    A = Matrix.from_csr(ctx, rp, col, val, prob.A.n).attach_transpose()
    B = Matrix.from_csr(ctx, rp, col, val, prob.A.n)
    Y = torch.rand(A.m, 4, device="cuda") * 10
    X1 = torch.full((A.n, 4), 1.0 / A.n, device="cuda")
    X2 = X1.clone()
    ctx.mlem(A, Y, prob.s, X1, 2)
    ctx.mlem(B, Y, prob.s, X2, 2)
    assert rel_l2(X1.cpu().numpy(), X2.cpu().numpy()) < 1e-5


@pytest.mark.parametrize("storage", ["f32", "bf16"])
def test_mlem_batched_transpose_path(small_problem, oracle, storage):
    """Multi-frame MLEM through the transposed gather path vs the FP64 oracle
    per frame, run on the operator the device stores (FP32, or BF16 values
    times the row scale): images within 1e-4 rel-L2 after 10 iterations, and
    the device log-likelihood of the last iteration equals the oracle's
    log-likelihood of the FP64 image it was evaluated at (rel 1e-6)."""
    from epetreco_amd.api import VAL_BF16
    from helpers import decoded
    ctx, prob = small_problem
    A = prob.A if storage == "f32" else prob.A.convert(VAL_BF16)
    A.attach_transpose()
    rp, col, val = decoded(A)
    F = 8
    g = torch.Generator(device="cpu").manual_seed(5)
    Y = torch.poisson(prob.y.cpu()[:, None].expand(-1, F) * (0.2 + torch.rand(F, generator=g)),
                      generator=g).contiguous()
    X = torch.full((A.n, F), 1.0 / A.n, device="cuda")
    _say("batched transpose path", storage)
    ll, _ = ctx.mlem(A, Y.to("cuda"), prob.s, X, 10, use_graph=True)
    s = prob.s.cpu().numpy()
    X0 = np.full((A.n, F), 1.0 / A.n)
    x9 = oracle.mlem_f64_frames(A.m, A.n, val, rp, col, Y.numpy(), s, X0, 9)
    x10 = oracle.mlem_f64_frames(A.m, A.n, val, rp, col, Y.numpy(), s, X0, 10)
    for f in range(F):
        assert rel_l2(X[:, f].cpu().numpy(), x10[:, f]) < REL_L2
        l9 = oracle.loglik(val, rp, col, Y[:, f].numpy().copy(), x9[:, f])
        assert abs(ll[9, f] - l9) <= 1e-6 * abs(l9)


def test_graph_replays_match_eager(small_problem):
    """Every replay of the captured iteration equals the eager iteration: the
    per-iteration log-likelihoods (one per replay) and the final image."""
    ctx, prob = small_problem
    A = prob.A
    xg = torch.full((A.n,), 1.0 / A.n, device="cuda")
    xe = xg.clone()
    llg, _ = ctx.mlem(A, prob.y, prob.s, xg, 12, use_graph=True)
    torch.cuda.synchronize()
    _ = (xg * 2).sum().item()          # eager work between the runs
    lle, _ = ctx.mlem(A, prob.y, prob.s, xe, 12, use_graph=False)
    # LDS float atomics make the summation order vary between runs
    np.testing.assert_allclose(llg[:, 0], lle[:, 0], rtol=1e-9)
    assert rel_l2(xg.cpu().numpy(), xe.cpu().numpy().astype(np.float64)) < 1e-5


def test_multiframe_autotuner(small_problem, monkeypatch):
    """The tuner times every valid shape of the path epet_mlem runs and keeps a
    valid one: (rows per block, frames per lane) for the blocked products,
    (rows per wavefront, frame tile) for the gather kernels; results with the
    tuned kernels match the default ones."""
    ctx, prob = small_problem
    from epetreco_amd.api import Matrix
    rp, col, val, _ = prob.A.to_torch()
    A = Matrix.from_csr(ctx, rp, col, val, prob.A.n).attach_transpose()
    (rf, Pf, rbk, Pb), (mf, mb) = ctx.mlem_tune(A, 16)
    _say("blocked: fwd rb,P =", rf, Pf, f"{mf:.3f} ms; bwd rb,P =", rbk, Pb, f"{mb:.3f} ms")
    assert rf in (128, 256, 512, 1024) and Pf in (4, 8)
    assert rbk in (128, 256, 512, 1024) and Pb in (4, 8)
    cfg = ctx.batched_config(A, 16)
    assert cfg["forward"]["rows_per_block"] == rf and cfg["back_projection"]["frames_per_lane"] == Pb
    ctx_g = env_ctx(monkeypatch, EPET_BATCHED="gather")
    G = Matrix.from_csr(ctx, rp, col, val, prob.A.n).attach_transpose()
    (Rf, Pgf, Rb, Pgb), _ = ctx_g.mlem_tune(G, 16)
    assert Rf in (1, 2, 4, 8) and Pgf in (4, 8, 16, 32) and Pgf <= 64 // Rf
    assert Rb in (1, 2, 4, 8) and Pgb in (4, 8, 16, 32) and Pgb <= 64 // Rb
    Y = torch.rand(A.m, 16, device="cuda") * 5
    X1 = torch.full((A.n, 16), 1.0 / A.n, device="cuda")
    X2 = X1.clone()
    ctx.mlem(A, Y, prob.s, X1, 3)                  # tuned configuration
    B = Matrix.from_csr(ctx, rp, col, val, prob.A.n)   # atomic back projection
    ctx.mlem(B, Y, prob.s, X2, 3)
    assert rel_l2(X1.cpu().numpy(), X2.cpu().numpy()) < 1e-5


@pytest.mark.parametrize("storage", ["f32", "bf16"])
def test_single_frame_transpose_path(small_problem, oracle, storage, monkeypatch):
    """Single-frame MLEM and SpMV^T through the attached transpose (row scale
    folded into the ratio for BF16) match the LDS-tiled back projection, and
    the FP64 oracle on the stored operator."""
    from epetreco_amd.api import Matrix, OP_T, VAL_BF16
    ctx, prob = small_problem
    rp, col, val, _ = prob.A.to_torch()
    A = Matrix.from_csr(ctx, rp, col, val, prob.A.n)
    if storage == "bf16":
        A = A.convert(VAL_BF16)
    A.attach_transpose()
    V = A.n
    e = torch.rand(A.m, device="cuda")
    c_tr = A.spmv(OP_T, e)
    ctx_t = env_ctx(monkeypatch, EPET_BP="tiled")
    c_tiled = A.spmv(OP_T, e, ctx=ctx_t)
    x_t = torch.full((V,), 1.0 / V, device="cuda")
    ctx_t.mlem(A, prob.y, prob.s, x_t, 10, use_graph=True)
    assert rel_l2(c_tr.cpu().numpy(), c_tiled.cpu().numpy().astype(np.float64)) < 1e-6
    x = torch.full((V,), 1.0 / V, device="cuda")
    ctx.mlem(A, prob.y, prob.s, x, 10, use_graph=True)
    assert rel_l2(x.cpu().numpy(), x_t.cpu().numpy().astype(np.float64)) < 1e-5
    from helpers import decoded
    rq, cq, vq = decoded(A)
    x64 = oracle.mlem_f64(A.m, V, vq, rq, cq, prob.y.cpu().numpy(), prob.s.cpu().numpy(),
                          np.full(V, 1.0 / V), 10)
    assert rel_l2(x.cpu().numpy(), x64) < REL_L2


@pytest.mark.parametrize("storage", ["f32", "bf16"])
def test_lds_tiled_products_and_mlem(gpu_ctx_factory, oracle, storage, monkeypatch):
    """LDS-tiled row products on rows long enough to use them (avg 300 entries)
    over 70,001 columns (three 32768-entry tiles, the last ragged) and over the
    transposed copy: forward and back projection vs the FP64 oracle on the
    stored operator, and 8 MLEM iterations (graph) vs the same iterations with
    the LDS path off (rel-L2 <= 1e-5) and vs the FP64 oracle (<= 1e-4)."""
    from epetreco_amd.api import Matrix, OP_N, OP_T, VAL_BF16
    from helpers import decoded
    ctx = gpu_ctx_factory()
    rng = np.random.default_rng(11)
    m, n, avg = 6000, 70001, 300
    lens = rng.poisson(avg, m)
    rp = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    col = np.concatenate([np.sort(rng.choice(n, l, replace=False)) for l in lens]).astype(np.int32)
    val = rng.random(rp[-1]).astype(np.float32)
    A = Matrix.from_csr(ctx, torch.as_tensor(rp), torch.as_tensor(col), torch.as_tensor(val), n)
    if storage == "bf16":
        A = A.convert(VAL_BF16)
    A.attach_transpose()
    rq, cq, vq = decoded(A)
    x = rng.random(n).astype(np.float32)
    e = rng.random(m).astype(np.float32)
    y = A.spmv(OP_N, torch.as_tensor(x, device="cuda")).cpu().numpy()
    np.testing.assert_allclose(y, oracle.csrmv(0, m, n, 1.0, vq, rq, cq, x, 0.0, np.zeros(m, np.float32)),
                               rtol=2e-5, atol=1e-6)
    if storage == "f32":
        c = A.spmv(OP_T, torch.as_tensor(e, device="cuda")).cpu().numpy()
        np.testing.assert_allclose(c, oracle.csrmv(1, m, n, 1.0, vq, rq, cq, e, 0.0,
                                                   np.zeros(n, np.float32)), rtol=2e-5, atol=1e-6)
    # the tile-major LDS product against the row kernel over the original arrays
    ctx_rows = env_ctx(monkeypatch, EPET_LDS="0")
    y_rows = A.spmv(OP_N, torch.as_tensor(x, device="cuda"), ctx=ctx_rows).cpu().numpy()
    np.testing.assert_allclose(y, y_rows, rtol=1e-6, atol=1e-7)
    counts = rng.poisson(30.0, m).astype(np.float32)
    s = torch.as_tensor(np.bincount(cq, weights=vq, minlength=n).astype(np.float32), device="cuda")
    yy = torch.as_tensor(counts, device="cuda")
    x1 = torch.full((n,), 1.0 / n, device="cuda")
    ctx.mlem(A, yy, s, x1, 8, use_graph=True)
    x2 = torch.full((n,), 1.0 / n, device="cuda")
    ctx_rows.mlem(A, yy, s, x2, 8, use_graph=True)
    assert rel_l2(x1.cpu().numpy(), x2.cpu().numpy().astype(np.float64)) < 1e-5
    x64 = oracle.mlem_f64(m, n, vq, rq, cq, counts, s.cpu().numpy(), np.full(n, 1.0 / n), 8)
    assert rel_l2(x1.cpu().numpy(), x64) < REL_L2


@pytest.mark.parametrize("F", [2, 5, 13, 16, 32, 64])
def test_multiframe_blocked_path(gpu_ctx_factory, oracle, monkeypatch, F):
    """Multi-frame MLEM through the blocked products (padded frame counts 8,
    16, 32 and 64; several column tiles each way; the back projection split
    over workgroups) on long rows over 70,001 columns, FP32
    values: images after 6 iterations vs the gather kernels (rel-L2 <= 1e-5)
    and vs the multi-frame FP64 oracle (<= 1e-4), and the per-frame
    log-likelihood of the last iteration vs the gather kernels."""
    from epetreco_amd.api import Matrix
    ctx = gpu_ctx_factory()
    rng = np.random.default_rng(F)
    m, n, avg = 5000, 70001, 250
    lens = rng.poisson(avg, m)
    rp = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    col = np.concatenate([np.sort(rng.choice(n, l, replace=False)) for l in lens]).astype(np.int32)
    val = rng.random(rp[-1]).astype(np.float32)
    A = Matrix.from_csr(ctx, torch.as_tensor(rp), torch.as_tensor(col), torch.as_tensor(val), n)
    A.attach_transpose()
    Y = rng.poisson(20.0, (m, F)).astype(np.float32)
    s = torch.as_tensor(np.bincount(col, weights=val, minlength=n).astype(np.float32), device="cuda")
    Yd = torch.as_tensor(Y, device="cuda")
    X1 = torch.full((n, F), 1.0 / n, device="cuda")
    _say("blocked multi-frame MLEM, F =", F)
    ll1, _ = ctx.mlem(A, Yd, s, X1, 6, use_graph=True)
    cfg = ctx.batched_config(A, F)
    _say("  ", cfg)
    assert cfg["path"].startswith("blocked")
    ctx_g = env_ctx(monkeypatch, EPET_BATCHED="gather")
    X2 = torch.full((n, F), 1.0 / n, device="cuda")
    ll2, _ = ctx_g.mlem(A, Yd, s, X2, 6, use_graph=True)
    assert ctx_g.batched_config(A, F)["path"].startswith("gather")
    X64 = oracle.mlem_f64_frames(m, n, val, rp, col, Y, s.cpu().numpy(), np.full((n, F), 1.0 / n), 6)
    for f in range(F):
        assert rel_l2(X1[:, f].cpu().numpy(), X2[:, f].cpu().numpy().astype(np.float64)) < 1e-5
        assert rel_l2(X1[:, f].cpu().numpy(), X64[:, f]) < REL_L2
    np.testing.assert_allclose(ll1[-1], ll2[-1], rtol=1e-7)


@pytest.mark.parametrize("storage", ["f32", "f16"])
def test_multiframe_blocked_uneven_rows(gpu_ctx_factory, oracle, monkeypatch, storage):
    """Blocked products on rows of very uneven length: most rows hold a few
    entries, every 97th row tens of thousands, so that single rows span many
    slots' shares of a (row block, tile) step, slot runs start at the nearest
    row boundary on either side, and some slots have nothing to do.  16
    frames, 5 MLEM iterations: images vs the gather kernels (rel-L2 <= 1e-5)
    and vs the multi-frame FP64 oracle on the stored values (<= 1e-4)."""
    from epetreco_amd.api import Matrix, VAL_F16
    ctx = gpu_ctx_factory()
    rng = np.random.default_rng(97)
    m, n, F = 4000, 40000, 16
    lens = rng.integers(1, 12, m)
    lens[::97] = rng.integers(8000, 30000, len(lens[::97]))
    rp = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    col = np.concatenate([np.sort(rng.choice(n, l, replace=False)) for l in lens]).astype(np.int32)
    val = (0.05 + rng.random(rp[-1])).astype(np.float32)
    A = Matrix.from_csr(ctx, torch.as_tensor(rp), torch.as_tensor(col), torch.as_tensor(val), n)
    if storage == "f16":
        A = A.convert(VAL_F16)
    A.attach_transpose()
    from helpers import decoded
    rq, cq, vq = decoded(A)
    Y = rng.poisson(5.0, (m, F)).astype(np.float32)
    s = torch.as_tensor(np.maximum(np.bincount(cq, weights=vq, minlength=n), 1e-3).astype(np.float32),
                        device="cuda")
    Yd = torch.as_tensor(Y, device="cuda")
    X1 = torch.full((n, F), 1.0 / n, device="cuda")
    ctx.mlem(A, Yd, s, X1, 5, use_graph=True)
    cfg = ctx.batched_config(A, F)
    _say("  uneven rows", storage, cfg)
    assert cfg["path"].startswith("blocked")
    ctx_g = env_ctx(monkeypatch, EPET_BATCHED="gather")
    X2 = torch.full((n, F), 1.0 / n, device="cuda")
    ctx_g.mlem(A, Yd, s, X2, 5, use_graph=True)
    X64 = oracle.mlem_f64_frames(m, n, vq, rq, cq, Y, s.cpu().numpy(), np.full((n, F), 1.0 / n), 5)
    for f in range(F):
        assert rel_l2(X1[:, f].cpu().numpy(), X2[:, f].cpu().numpy().astype(np.float64)) < 1e-5
        assert rel_l2(X1[:, f].cpu().numpy(), X64[:, f]) < REL_L2


def test_voxel_renumbered_tiles(gpu_ctx_factory, monkeypatch):
    """Single-frame tile-major copies in y-major voxel order (EPET_LDS_MAP=1,
    for matrices over the context's grid): the projector workload's forward
    product and 4 MLEM iterations match the copies in the matrix's own column
    order and the kernels without LDS copies."""
    from epetreco_amd.api import Matrix, OP_N
    monkeypatch.setenv("EPET_LDS_MAP", "1")
    ctx = gpu_ctx_factory("projector", seed=0)
    monkeypatch.delenv("EPET_LDS_MAP")
    A = ctx.sm_build(None, m=ctx.setup.nchannels)
    rng = np.random.default_rng(3)
    x = torch.as_tensor(rng.random(A.n).astype(np.float32), device="cuda")
    y_map = A.spmv(OP_N, x).cpu().numpy()
    ctx_rows = env_ctx(monkeypatch, EPET_LDS="0")
    y_rows = A.spmv(OP_N, x, ctx=ctx_rows).cpu().numpy()
    np.testing.assert_allclose(y_map, y_rows, rtol=1e-6, atol=1e-7)
    Y = torch.as_tensor(rng.poisson(5.0, A.m).astype(np.float32), device="cuda")
    s = A.colsum_f64().float()
    x1 = torch.full((A.n,), 1.0 / A.n, device="cuda")
    ctx.mlem(A, Y, s, x1, 4, use_graph=True)
    rp, col, val, _ = A.to_torch()
    B = Matrix.from_csr(ctx, rp, col, val, A.n)
    ctx_plain = env_ctx(monkeypatch, EPET_LDS_MAP="0")
    x2 = torch.full((A.n,), 1.0 / A.n, device="cuda")
    ctx_plain.mlem(B, Y, s, x2, 4, use_graph=True)
    assert rel_l2(x1.cpu().numpy(), x2.cpu().numpy().astype(np.float64)) < 1e-5
