"""Shared helpers of the GPU parity tests."""
import numpy as np


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def decoded(M):
    """A library matrix's stored operator as host FP32 CSR (rowptr, col, val):
    BF16/FP16 values decoded and multiplied by their row scale."""
    from epetreco_amd.api import VAL_BF16, VAL_F32
    rp, col, val, sc = M.to_torch()
    rp_h = rp.cpu().numpy()
    if M.value_type == VAL_F32:
        return rp_h, col.cpu().numpy(), val.cpu().numpy()
    raw = val.cpu().numpy().view(np.uint16)
    if M.value_type == VAL_BF16:
        v = (raw.astype(np.uint32) << 16).view(np.float32)
    else:
        v = raw.view(np.float16).astype(np.float32)
    rows = np.repeat(np.arange(M.m), np.diff(rp_h))
    return rp_h, col.cpu().numpy(), (v * sc.cpu().numpy()[rows]).astype(np.float32)


def env_ctx(monkeypatch, **env):
    """A fresh Context created with the given EPET_* variables set: a context
    reads its path choices once, when it is created."""
    from epetreco_amd.api import Context
    for k, v in env.items():
        monkeypatch.setenv(k, v)
    ctx = Context(0)
    for k in env:
        monkeypatch.delenv(k)
    return ctx
