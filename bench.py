"""Benchmark of the MLEM hot path on one MI355X.

Workload (default): two-head scanner with 13x13 segments per head and 180
angles of 2 deg (5,140,980 channels), 52^3 grid of 1 mm voxels, 64 Monte-Carlo
rays per channel, FP32 values in ECSR, phantom of 8 ellipsoids plus a 0.05
background cylinder, Poisson data with 1e7 counts.  A step is one plain FP32
MLEM iteration (forward projection + ratio, back projection, update,
normalisation) over the measured channels with the system matrix resident in
HBM.  The 32-frame batched BF16-stored MLEM of the same matrix is timed beside
it.  Synthetic data, seeded.

    python bench.py [--steps K] [--warmup W] [--gpus 1]
"""
from __future__ import annotations

import argparse
import gc
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBPS = 8000.0


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def event_time_ms(fn, reps):
    import torch
    st = torch.cuda.Event(enable_timing=True)
    en = torch.cuda.Event(enable_timing=True)
    st.record()
    for _ in range(reps):
        fn()
    en.record()
    torch.cuda.synchronize()
    return st.elapsed_time(en) / reps


def host_info():
    """Host cores and CPU model of the machine the CPU legs run on."""
    from oracle import oracle as O
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    aff = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    return {"nproc": os.cpu_count(), "affinity": aff, "threads": int(O.lib().eo_max_threads()),
            "cpu_model": model}


def cpu_baseline(A, y_dev, s_dev, host, target_s=15.0):
    """The oracle's FP32 MLEM iteration (OpenMP over every thread the process
    may use) on a leading row sample of A, scaled to the full row count by
    nonzeros."""
    import ctypes
    import numpy as np
    from oracle import oracle as O
    rp, col, val, _ = A.to_torch()
    rp_h = rp.cpu().numpy()
    # sample: the first rows holding ~2% of the nonzeros
    target_nnz = max(1, int(0.02 * A.nnz))
    msamp = int(np.searchsorted(rp_h, target_nnz))
    msamp = max(1, min(msamp, A.m))
    nnz_s = int(rp_h[msamp])
    colh = col[:nnz_s].cpu().numpy()
    valh = val[:nnz_s].cpu().numpy()
    del rp, col, val
    rph = rp_h[:msamp + 1].copy()
    yh = y_dev[:msamp].cpu().numpy()
    sh = s_dev.cpu().numpy()
    sh = (sh / sh.sum()).astype(np.float32)
    x = np.full(A.n, 1.0 / A.n, dtype=np.float32)
    yhat = np.zeros(msamp, dtype=np.float32)
    c = np.zeros(A.n, dtype=np.float32)
    L = O.lib()
    nth = host["threads"]
    args = (msamp, A.n, O.ptr(valh, ctypes.c_float), O.ptr(rph, ctypes.c_int64),
            O.ptr(colh, ctypes.c_int), O.ptr(yh, ctypes.c_float), O.ptr(sh, ctypes.c_float),
            O.ptr(x, ctypes.c_float), O.ptr(yhat, ctypes.c_float), O.ptr(c, ctypes.c_float), nth)
    L.eo_mlem_iter_omp(*args)   # warm-up
    times = []
    t_all = time.time()
    while len(times) < 5 and (time.time() - t_all) < target_s:
        t0 = time.perf_counter()
        L.eo_mlem_iter_omp(*args)
        times.append(time.perf_counter() - t0)
    t_it = float(np.median(times)) * (A.nnz / nnz_s)
    return {"value": 1.0 / t_it, "unit": "frame-iter/s", "cores": nth, "kind": "port",
            "host": host,
            "sample": f"oracle FP32 MLEM iteration (OpenMP, {nth} threads) on the first "
                      f"{msamp} of {A.m} measured rows ({nnz_s} of {A.nnz} nonzeros), "
                      f"median of {len(times)}, scaled by nonzeros"}


def cpu_generation(setup, grid, nrays, host, frac=5e-4, seed=0):
    """The oracle's system-matrix generation (the reference's full candidate
    scan plus the Monte-Carlo slab values of every candidate) on a seeded
    channel sample, extrapolated linearly to every channel."""
    import numpy as np
    from oracle import oracle as O
    rng = np.random.default_rng(seed + 99)
    ns = max(1, int(round(frac * setup.nchannels)))
    lst = np.sort(rng.choice(setup.nchannels, ns, replace=False)).astype(np.int32)
    t0 = time.perf_counter()
    rp, _, _ = O.sm_rows(setup, grid, lst, nrays, 0)
    t = time.perf_counter() - t0
    return {"seconds_all_channels": t * setup.nchannels / ns, "sample_channels": ns,
            "sample_seconds": t, "sample_nonzeros": int(rp[-1]), "threads": host["threads"],
            "sample": f"{ns} seeded channels of {setup.nchannels} ({100.0 * ns / setup.nchannels:.3g}%), "
                      f"oracle full voxel scan + {nrays}-ray values, OpenMP {host['threads']} "
                      f"threads, extrapolated linearly"}


def spmv_bytes(nnz, m, V):
    """Algorithmic bytes of one FP32-ECSR SpMV: values and columns once, int64
    row pointers, the row-side vector once and the voxel-side vector once."""
    return nnz * 8 + (m + 1) * 8 + m * 4 + V * 4


def batched_bytes(nnz, m, V, F, sval, scaled):
    """Algorithmic bytes of one F-frame MLEM iteration with sval-byte stored
    values: both products stream the values and int32 columns once, the int64
    row pointers and (reduced precision) the row scales; per frame the data,
    ratio and projections on the row side (12 B per row) and the image,
    sensitivity, back projection and update on the voxel side (24 B per
    voxel)."""
    return 2 * (nnz * (sval + 4) + (m * 4 if scaled else 0) + (m + 1) * 8) + F * (12 * m + 24 * V)


def batched_traffic(workload, storage):
    """HBM bytes per batched iteration of the blocked products (both
    directions, the split reduction and the gather) from the committed counter
    summary, or None."""
    path = os.path.join(ROOT, "profiles", f"traffic_batched_{workload}.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        t = json.load(f)
    k = t.get("storage", {}).get(storage)
    if k is None:
        return None
    return k["fetch_bytes"] * t["copy_scale_fetch"] + k["write_bytes"] * t["copy_scale_write"]


def measured_traffic(workload, kernel):
    """HBM bytes per launch of `kernel` from the committed counter summary
    (FETCH_SIZE and WRITE_SIZE passes, each scaled by the same counter's reading
    for a device copy of known size), or None when no summary is present."""
    path = os.path.join(ROOT, "profiles", f"traffic_{workload}.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        t = json.load(f)
    k = t.get("kernels", {}).get(kernel)
    if k is None:
        return None
    return k["fetch_bytes"] * t["copy_scale_fetch"] + k["write_bytes"] * t["copy_scale_write"]


def series_104(frames=8, niter=30, chunk=400_000):
    """Dynamic series on the 104^3 grid of 0.5 mm voxels, all 5,140,980 channels,
    32 rays per channel: `frames` frames of the seeded phantom with totals
    log-uniform in [1e6, 1e7], `niter` MLEM iterations per frame, batched with
    FP16 and with BF16 values (per-row FP32 scale), against FP32 storage (one
    single-frame MLEM per frame) and against the FP64 oracle on the host (all
    frames, run concurrently on the host cores).  The operator is built in
    channel chunks for the data and the sensitivity, then over the measured
    channels once per run, so that each storage format has the device to
    itself; the oracle works on a host copy.  Prints one JSON line."""
    import threading

    import numpy as np
    import torch
    from epetreco_amd import geometry as G
    from epetreco_amd import phantom as PH
    from epetreco_amd.api import Context, OP_N, VAL_BF16, VAL_F16
    from oracle import oracle as O

    setup, grid, nrays = G.workload("large")
    ctx = Context(0)
    ctx.set_geometry(setup, grid, nrays=nrays, seed=0)
    V, N, F = grid.nvoxels, setup.nchannels, frames
    host = host_info()
    imgs, totals = PH.dynamic_series(grid, "large", F, seed=0, counts=(1e6, 1e7))
    Xt = torch.as_tensor(imgs.T.astype(np.float32), device="cuda").contiguous()
    # data and sensitivity over all channels, chunk by chunk
    say(f"104^3 series: {N} channels x {V} voxels, {nrays} rays, {F} frames")
    t0 = time.time()
    ybar = np.zeros((N, F))
    sens = torch.zeros(V, dtype=torch.float64, device="cuda")
    nnz_all = 0
    for c0 in range(0, N, chunk):
        ids = np.arange(c0, min(N, c0 + chunk), dtype=np.int32)
        Ac = ctx.sm_build(ids)
        nnz_all += Ac.nnz
        ybar[c0:c0 + len(ids)] = Ac.spmm(OP_N, Xt).double().cpu().numpy()
        sens += Ac.colsum_f64()
        Ac.close()
        del Ac
        say(f"    channels {c0}..{min(N, c0 + chunk)}: {nnz_all} nonzeros so far")
    torch.cuda.synchronize()
    gen_all_s = time.time() - t0
    say(f"  all channels: {nnz_all} nonzeros, built in chunks in {gen_all_s:.1f} s")
    ybar *= totals / ybar.sum(0)
    counts = np.random.default_rng(5).poisson(ybar).astype(np.float32)
    rows = np.nonzero(counts.sum(1) > 0)[0].astype(np.int32)
    Y = np.ascontiguousarray(counts[rows])
    del ybar, counts
    s = sens.float()
    t0 = time.time()
    A = ctx.sm_build(rows)
    torch.cuda.synchronize()
    gen_s = time.time() - t0
    m, nnz = A.m, A.nnz
    say(f"  measured channels {m}, nonzeros {nnz}, built in {gen_s:.1f} s")
    # host copy of the FP32 operator for the oracle
    rp_t, col_t, val_t, _ = A.to_torch()
    rp_h = rp_t.cpu()
    del rp_t
    col_h = col_t.cpu()
    del col_t
    val_h = val_t.cpu()
    del val_t
    torch.cuda.empty_cache()
    s_h = s.cpu().numpy()
    x0 = np.full((V, F), 1.0 / V)
    oracle_out = {}

    def run_oracle():
        # one iteration per call (an iteration depends on the image only)
        t = time.time()
        Xo = x0
        for it in range(niter):
            Xo = O.mlem_f64_frames_par(m, V, val_h.numpy(), rp_h.numpy(), col_h.numpy(), Y, s_h,
                                       Xo, 1)
            say(f"    FP64 oracle iteration {it + 1}/{niter}: {time.time() - t:.0f} s")
        oracle_out["X"] = Xo
        oracle_out["s"] = time.time() - t

    th = threading.Thread(target=run_oracle)
    th.start()
    Yd = torch.as_tensor(Y, device="cuda")
    # FP32 storage: single-frame MLEM per frame over A and its transposed copy
    # (no LDS copies, so that both fit next to each other)
    os.environ["EPET_LDS"] = "0"
    try:
        ctx32 = Context(0)
    finally:
        del os.environ["EPET_LDS"]
    A.attach_transpose()
    X32 = np.zeros((V, F), np.float32)
    torch.cuda.synchronize()
    t0 = time.time()
    for f in range(F):
        x = torch.full((V,), 1.0 / V, device="cuda")
        ctx32.mlem(A, Yd[:, f].contiguous(), s, x, niter, use_graph=True)
        X32[:, f] = x.cpu().numpy()
    torch.cuda.synchronize()
    fp32_ms = 1e3 * (time.time() - t0) / (F * niter)
    say(f"  FP32 storage, single-frame: {fp32_ms:.1f} ms per frame-iteration")
    A.close()
    del A
    torch.cuda.empty_cache()
    out = {"series_104": True, "workload": f"{N} channels, {grid.nx}^3 voxels of {grid.dx * 1e3:g} mm, "
                                             f"{nrays} rays per channel",
           "channels": N, "voxels": V, "measured_channels": m, "nnz": nnz, "nnz_all_channels": nnz_all,
           "frames": F, "iterations": niter, "frame_counts": [float(t) for t in totals],
           "generation_all_channels_s": gen_all_s, "generation_measured_s": gen_s,
           "fp32_single_frame_ms_per_frame_iteration": fp32_ms}
    results = {}
    for key, vt in (("fp16_row_scale", VAL_F16), ("bf16_row_scale", VAL_BF16)):
        # rebuilt (deterministic generation) rather than uploaded: a generated
        # matrix carries its channel ids and grid, which order the blocked copies
        A32 = ctx.sm_build(rows)
        M = A32.convert(vt)
        A32.close()
        del A32
        M.attach_transpose()
        X = torch.full((V, F), 1.0 / V, device="cuda")
        t0 = time.time()
        ctx.mlem(M, Yd, s, X, 1, use_graph=True)          # blocked copies and tuning
        torch.cuda.synchronize()
        prep_s = time.time() - t0
        X.fill_(1.0 / V)
        torch.cuda.synchronize()
        t0 = time.time()
        ctx.mlem(M, Yd, s, X, niter, use_graph=True)
        torch.cuda.synchronize()
        t_b = time.time() - t0
        Xh = X.cpu().numpy()
        e32 = [float(np.linalg.norm(Xh[:, f] - X32[:, f]) / np.linalg.norm(X32[:, f])) for f in range(F)]
        results[key] = {"ms_per_iteration": 1e3 * t_b / niter, "frame_iter_per_s": niter * F / t_b,
                        "first_call_s": prep_s, "kernels": ctx.batched_config(M, F),
                        "frame_rel_l2_vs_fp32_storage": e32, "X": Xh}
        say(f"  {key}: {1e3 * t_b / niter:.1f} ms per {F}-frame iteration "
            f"(first call {prep_s:.1f} s); rel-L2 vs FP32 storage " + " ".join(f"{e:.2e}" for e in e32))
        M.close()
        del M, X
        torch.cuda.empty_cache()
    th.join()
    X64 = oracle_out["X"]
    say(f"  FP64 oracle ({host['threads']} threads, concurrent): {oracle_out['s']:.0f} s")

    def errs(Xa):
        return [float(np.linalg.norm(Xa[:, f].astype(np.float64) - X64[:, f]) / np.linalg.norm(X64[:, f]))
                for f in range(F)]

    out["fp32_storage"] = {"frame_rel_l2_vs_fp64": errs(X32)}
    say("  FP32 storage vs FP64: " + " ".join(f"{e:.2e}" for e in out["fp32_storage"]["frame_rel_l2_vs_fp64"]))
    for key, r in results.items():
        Xh = r.pop("X")
        r["frame_rel_l2_vs_fp64"] = errs(Xh)
        r["meets_1e-3_vs_fp64"] = bool(max(r["frame_rel_l2_vs_fp64"]) <= 1e-3)
        say(f"  {key} vs FP64: " + " ".join(f"{e:.2e}" for e in r["frame_rel_l2_vs_fp64"]))
        out[key] = r
    out["oracle"] = {"seconds": oracle_out["s"], "threads": host["threads"], "host": host,
                     "reference": "FP64 oracle MLEM on the FP32 matrix, all frames"}
    print(json.dumps(out))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--frames", type=int, default=32)
    ap.add_argument("--batched-steps", type=int, default=5)
    ap.add_argument("--workload", default="full")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--accuracy-iters", type=int, default=20,
                    help="batched MLEM iterations of the accuracy check against the FP64 oracle "
                         "(0: skip)")
    ap.add_argument("--cpu-generation", type=int, default=1,
                    help="time the oracle's generation on a channel sample (0: skip)")
    ap.add_argument("--profile-batched", default="",
                    help="after building the problem, convert to this storage (bf16, f16, f32), "
                         "tune the batched products for --frames frames and run 3 batched MLEM "
                         "iterations (for kernel traces and counter runs)")
    ap.add_argument("--series-104", action="store_true",
                    help="the 104^3 dynamic series (all channels, 32 rays, 8 frames, 30 iterations, "
                         "FP16 and BF16 storage against FP32 storage and the FP64 oracle); one JSON line")
    ap.add_argument("--profile-kernels", action="store_true",
                    help="after building the problem, launch only the forward SpMV, the back "
                         "projection and a 1 GiB device copy (3 times each) for counter runs")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU: every box has one MI355X; "
                                   "the row-sharded multi-GPU design is described, not run",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    if args.series_104:
        series_104()
        return

    import numpy as np
    import torch
    from epetreco_amd import geometry as G
    from epetreco_amd import phantom as PH
    from epetreco_amd.api import Context, OP_N, OP_T, VAL_BF16, VAL_F16

    setup, grid, nrays = G.workload(args.workload)
    ctx = Context(0)
    ctx.set_geometry(setup, grid, nrays=nrays, seed=0)
    V = grid.nvoxels

    # ---- system matrix (one-off, outside the timed region)
    say(f"generating system matrix: {setup.nchannels} channels x {V} voxels, {nrays} rays")
    torch.cuda.synchronize()
    t0 = time.time()
    A_full = ctx.sm_build(None, m=setup.nchannels)
    torch.cuda.synchronize()
    gen_s = time.time() - t0
    say(f"  {A_full.nnz} nonzeros in {gen_s:.2f} s")
    prob = PH.make_problem(ctx, A_full, args.workload, seed=0)
    full_nnz = A_full.nnz
    A_full.close()
    del A_full
    gc.collect()
    A = prob.A
    m = A.m
    say(f"measured channels {m}, nonzeros {A.nnz}")

    # ---- transposed copy (one-off): the back projection streams its rows
    torch.cuda.synchronize()
    t0 = time.time()
    A.attach_transpose()
    torch.cuda.synchronize()
    tr_s = time.time() - t0
    say(f"transposed copy built in {tr_s:.2f} s")

    if args.profile_batched:
        F = args.frames
        vt = {"bf16": VAL_BF16, "f16": VAL_F16, "f32": None}[args.profile_batched]
        M = A if vt is None else A.convert(vt)
        M.attach_transpose()
        Yp = torch.poisson(torch.full((m, F), 3.0, device="cuda")).contiguous()
        Xp = torch.full((V, F), 1.0 / V, device="cuda")
        t0 = time.perf_counter()
        cfg, tms = ctx.mlem_tune(M, F)
        torch.cuda.synchronize()
        say(f"tuned in {time.perf_counter() - t0:.2f} s (incl. blocked copies): {cfg} {tms}")
        ctx.mlem(M, Yp, prob.s, Xp, 3, use_graph=False)
        torch.cuda.synchronize()
        src = torch.ones(1 << 28, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            torch.mul(src, 1.0, out=dst)       # elementwise kernel: 1 GiB read, 1 GiB written
        torch.cuda.synchronize()
        print(json.dumps({"profile_batched": args.profile_batched, "frames": F,
                          "config": ctx.batched_config(M, F), "iterations": 3,
                          "copy_bytes_each_way": src.numel() * 4}))
        return

    if args.profile_kernels:
        xv = torch.rand(V, device="cuda")
        ev = torch.rand(m, device="cuda")
        yv = torch.zeros(m, device="cuda")
        cv = torch.zeros(V, device="cuda")
        for _ in range(3):
            A.spmv(OP_N, xv, yv)
        for _ in range(3):
            A.spmv(OP_T, ev, cv, beta=1.0)
        src = torch.ones(1 << 28, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            torch.mul(src, 1.0, out=dst)       # elementwise kernel: 1 GiB read, 1 GiB written
        torch.cuda.synchronize()
        print(json.dumps({"profile_kernels": True, "nnz": A.nnz, "rows": m, "voxels": V,
                          "algorithmic_bytes_fwd": spmv_bytes(A.nnz, m, V),
                          "copy_bytes_each_way": src.numel() * 4}))
        return

    # ---- plain FP32 MLEM: the timed steps
    x0 = torch.full((V,), 1.0 / V, device="cuda")
    x = x0.clone()
    ctx.mlem(A, prob.y, prob.s, x, args.warmup, use_graph=True)
    x = x0.clone()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ctx.mlem(A, prob.y, prob.s, x, args.steps, use_graph=True)
    torch.cuda.synchronize()
    t_plain = time.perf_counter() - t0
    ms_step = 1e3 * t_plain / args.steps
    plain_fis = args.steps / t_plain
    x_plain = x.clone()
    say(f"plain FP32 MLEM: {ms_step:.3f} ms/iteration, {plain_fis:.2f} frame-iter/s")

    # ---- per-kernel roofline: forward SpMV and back projection, HIP events on
    # the context's stream (torch's current stream)
    xv = torch.rand(V, device="cuda")
    ev = torch.rand(m, device="cuda")
    yv = torch.zeros(m, device="cuda")
    cv = torch.zeros(V, device="cuda")
    A.spmv(OP_N, xv, yv)
    fwd_ms = event_time_ms(lambda: A.spmv(OP_N, xv, yv), 10)
    bwd_ms = event_time_ms(lambda: A.spmv(OP_T, ev, cv, beta=1.0), 5)
    B_fwd = spmv_bytes(A.nnz, m, V)
    B_bwd = spmv_bytes(A.nnz, m, V)
    fwd_gbps = B_fwd / fwd_ms / 1e6
    bwd_gbps = B_bwd / bwd_ms / 1e6
    say(f"forward SpMV {fwd_ms:.3f} ms = {fwd_gbps:.0f} GB/s; back projection {bwd_ms:.3f} ms "
        f"= {bwd_gbps:.0f} GB/s")

    # ---- batched multi-frame MLEM, 32 frames: BF16 storage (headline) and
    # FP32 storage, through the transposed copy
    batched = None
    host = host_info()
    if not args.no_batched:
        F = args.frames
        imgs, totals = PH.dynamic_series(grid, args.workload, F, seed=0)
        X = torch.as_tensor(imgs.T.astype(np.float32), device="cuda").contiguous()
        Yb = A.spmm(OP_N, X)
        Yb = Yb * torch.as_tensor((totals / Yb.double().sum(0).cpu().numpy()).astype(np.float32),
                                  device="cuda")
        gen = torch.Generator(device="cuda").manual_seed(1)
        Yc = torch.poisson(Yb, generator=gen).contiguous()
        del Yb, X
        # accuracy frames: the fewest and the most counts, the first and the last
        pick = sorted({int(np.argmin(totals)), int(np.argmax(totals)), 0, F - 1})
        acc_iters = args.accuracy_iters

        def run_batched(M, label, key, sval, scaled):
            t0 = time.perf_counter()
            M.attach_transpose()
            torch.cuda.synchronize()
            t_tr = time.perf_counter() - t0
            Xb = torch.full((V, F), 1.0 / V, device="cuda")
            ctx.mlem(M, Yc, prob.s, Xb, 1, use_graph=True)       # one-off copies, warm-up
            Xb.fill_(1.0 / V)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            ctx.mlem(M, Yc, prob.s, Xb, args.batched_steps, use_graph=True)
            torch.cuda.synchronize()
            t_b = time.perf_counter() - t0
            fis = args.batched_steps * F / t_b
            say(f"  {label} x{F}: {1e3 * t_b / args.batched_steps:.2f} ms/iteration, "
                f"{fis:.1f} frame-iter/s ({fis / plain_fis:.2f}x plain FP32); transpose {t_tr:.2f} s")
            ms_it = 1e3 * t_b / args.batched_steps
            B_it = batched_bytes(M.nnz, m, V, F, sval, scaled)
            tr = batched_traffic(args.workload, key)
            res = {"frame_iter_per_s": fis, "ms_per_iteration": ms_it,
                   "speedup_vs_plain_fp32": fis / plain_fis, "iterations_timed": args.batched_steps,
                   "kernels": ctx.batched_config(M, F),
                   "roofline": {"bound": "hbm", "achieved": B_it / ms_it / 1e6,
                                "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                                "frac": B_it / ms_it / 1e6 / HBM_PEAK_GBPS,
                                "algorithmic_bytes": B_it, "traffic": tr,
                                "scope": "one F-frame MLEM iteration (graph replay, all kernels)"}}
            say(f"    {B_it / 1e9:.1f} GB algorithmic per iteration: {B_it / ms_it / 1e6:.0f} GB/s, "
                f"{B_it / ms_it / 1e6 / HBM_PEAK_GBPS:.1%} of 8 TB/s"
                + (f"; counters {tr / 1e9:.1f} GB" if tr else ""))
            Xa = None
            if acc_iters > 0:
                Xa = torch.full((V, len(pick)), 1.0 / V, device="cuda")
                ctx.mlem(M, Yc[:, pick].contiguous(), prob.s, Xa, acc_iters, use_graph=True)
                Xa = Xa.cpu().numpy()
            return Xb, res, Xa

        say(f"batched MLEM, {F} frames")
        runs = {}
        for key, vt, label in (("bf16_row_scale", VAL_BF16, "BF16 storage"),
                               ("fp16_row_scale", VAL_F16, "FP16 storage")):
            M = A.convert(vt)
            runs[key] = run_batched(M, label, key, 2, True)
            M.close()
            del M
            gc.collect()
            torch.cuda.empty_cache()
        runs["fp32"] = run_batched(A, "FP32 storage", "fp32", 4, False)
        Xf = runs["fp32"][0]
        batched = {"frames": F}
        for key, (Xk, rk, _) in runs.items():
            batched[key] = rk
            if key != "fp32":
                err = ((Xk - Xf).norm(dim=0) / Xf.norm(dim=0)).cpu().numpy()
                rk["max_frame_rel_l2_vs_fp32_storage"] = float(err.max())
                say(f"  max frame rel-L2, {key} vs FP32 storage after {args.batched_steps} "
                    f"iterations: {err.max():.2e}")
        rb, rh, rf = runs["bf16_row_scale"][1], runs["fp16_row_scale"][1], runs["fp32"][1]
        Ab_acc, Ah_acc, Af_acc = (runs[k][2] for k in ("bf16_row_scale", "fp16_row_scale", "fp32"))
        del runs, Xf
        if acc_iters > 0:
            # FP64 oracle MLEM on the FP32 matrix for the picked frames
            from oracle import oracle as O
            say(f"  FP64 oracle: frames {pick}, {acc_iters} iterations, {host['threads']} threads")
            t0 = time.perf_counter()
            rp, col, val, _ = A.to_torch()
            rp, col, val = rp.cpu().numpy(), col.cpu().numpy(), val.cpu().numpy()
            torch.cuda.empty_cache()
            X64 = O.mlem_f64_frames_par(A.m, V, val, rp, col, Yc[:, pick].cpu().numpy(),
                                        prob.s.cpu().numpy(), np.full((V, len(pick)), 1.0 / V),
                                        acc_iters)
            del rp, col, val
            t_o = time.perf_counter() - t0

            def errs(Xg):
                return [float(np.linalg.norm(Xg[:, k].astype(np.float64) - X64[:, k]) /
                              np.linalg.norm(X64[:, k])) for k in range(len(pick))]

            acc = {"frames": pick, "frame_counts": [float(totals[k]) for k in pick],
                   "iterations": acc_iters, "bar": 1e-3,
                   "reference": f"FP64 oracle MLEM on the FP32 matrix (OpenMP, {host['threads']} threads)"}
            for name, rk, Xa in (("BF16", rb, Ab_acc), ("FP16", rh, Ah_acc), ("FP32", rf, Af_acc)):
                ek = errs(Xa)
                rk["max_frame_rel_l2_vs_fp64"] = max(ek)
                rk["frame_rel_l2_vs_fp64"] = ek
                rk["meets_1e-3_vs_fp64"] = bool(max(ek) <= 1e-3)
                say(f"  rel-L2 vs FP64 after {acc_iters} iterations, frames {pick}, {name}: "
                    + " ".join(f"{e:.2e}" for e in ek))
            say(f"  (oracle {t_o:.1f} s)")
            batched["accuracy"] = acc

    cpu = None
    gen_cpu = None
    if not args.no_cpu:
        say("CPU baseline (oracle) on a row sample ...")
        cpu = cpu_baseline(A, prob.y, prob.s, host)
        say(f"  {cpu['value']:.3f} frame-iter/s on {cpu['cores']} threads ({host['cpu_model']})")
        if args.cpu_generation:
            say("CPU generation (oracle) on a channel sample ...")
            gen_cpu = cpu_generation(setup, grid, nrays, host)
            say(f"  {gen_cpu['sample_channels']} channels in {gen_cpu['sample_seconds']:.1f} s; "
                f"all channels ~{gen_cpu['seconds_all_channels']:.0f} s")
            cpu["generation"] = gen_cpu

    algo_frac = fwd_gbps / HBM_PEAK_GBPS
    traffic = measured_traffic(args.workload, "k_rows_lds<ValF32,4,1,0>")
    out = {
        "metric": "MLEM frame-iter/s + SpMV HBM GB/s at 52^3/5.1M ch; batched BF16 vs plain FP32",
        "value": plain_fis,
        "unit": "frame-iter/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded ellipsoid phantom, Poisson counts)",
        "config": {"workload": f"plain MLEM, {setup.nchannels} channels, {grid.nx}^3 voxels of "
                               f"{grid.dx * 1e3:g} mm, {nrays} rays/channel, FP32 ECSR",
                   "channels": setup.nchannels, "measured_channels": m, "voxels": V,
                   "rays_per_channel": nrays, "nnz": A.nnz, "nnz_all_channels": full_nnz},
        "spmv_fwd_gbps": fwd_gbps,
        "spmv_bwd_gbps": bwd_gbps,
        "generation_s": gen_s,
        "transpose_s": tr_s,
        "batched": batched,
        "roofline": {"bound": "hbm", "achieved": fwd_gbps, "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                     "frac": algo_frac, "traffic": traffic,
                     "frac_physical": (traffic / fwd_ms / 1e6 / HBM_PEAK_GBPS) if traffic else None,
                     "kernel": "forward SpMV (k_rows_lds<ValF32,4,1,0> + k_lds_reduce, timed together)",
                     "algorithmic_bytes": B_fwd, "avg_ms": fwd_ms},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
