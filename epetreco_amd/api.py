"""Python host surface over the C-ABI, mirroring the reference's host calls.

Device buffers are torch tensors on the context's device; every call enqueues
on the context's stream (torch's current stream at creation unless given).
Names follow the reference's functions: getWorkqueue (getWorkqueue.hpp:129-223),
getSystemMatrixFromWorkqueue (getSystemMatrixFromWorkqueue.cu:97-102),
getSystemMatrix (getSystemMatrixDeviceOnly.cu:324-331),
systemMatrixCalculation (wrappers.hpp:172-247), cooSort (cooSort.hpp:53-71),
convertCoo2Csr (cusparseWrapper.hpp:38-45), convertCsr2Ecsr
(convertCsr2Ecsr.hpp:34-47), csrmv (csrmv.hpp:27-35) and the
mlemOperations (mlemOperations.hpp:22-139).  Errors raise EpetError carrying
the library's file(line) message.
"""
from __future__ import annotations

import ctypes
from typing import Optional

import numpy as np
import torch

from . import _lib
from ._lib import OP_N, OP_T, VAL_BF16, VAL_F16, VAL_F32, EpetError, check, lib

EPET_ERR_ARG = 1
from .geometry import MeasurementSetup, VoxelGrid

__all__ = ["Context", "Matrix", "EpetError", "OP_N", "OP_T", "VAL_F32", "VAL_BF16", "VAL_F16"]


def _p(t: Optional[torch.Tensor]):
    """Device pointer of a tensor (None -> NULL)."""
    if t is None:
        return None
    if not t.is_cuda:
        raise ValueError("expected a device tensor")
    if not t.is_contiguous():
        raise ValueError("expected a contiguous tensor")
    return ctypes.c_void_p(t.data_ptr())


def _check_dtype(t: torch.Tensor, dtype, name: str):
    if t.dtype != dtype:
        raise TypeError(f"{name} must be {dtype}, got {t.dtype}")


def _check_vec(t, n: int, name: str, device) -> None:
    """A float32 device tensor of exactly n elements on `device` (checked on the
    host before any device call: a wrong size would make a kernel read or write
    past the allocation)."""
    if not isinstance(t, torch.Tensor):
        raise EpetError(EPET_ERR_ARG, f"{name}: expected a torch tensor")
    if t.dtype != torch.float32:
        raise EpetError(EPET_ERR_ARG, f"{name} must be float32, got {t.dtype}")
    if t.device != device:
        raise EpetError(EPET_ERR_ARG, f"{name} is on {t.device}, the context is on {device}")
    if not t.is_contiguous():
        raise EpetError(EPET_ERR_ARG, f"{name} must be contiguous")
    if t.numel() != n:
        raise EpetError(EPET_ERR_ARG, f"{name} has {t.numel()} elements, expected {n}")


class Context:
    """epet_ctx: device, stream, geometry, rays per channel, Philox seed."""

    def __init__(self, device: int = 0, stream: Optional[torch.cuda.Stream] = None):
        self.device = torch.device("cuda", device)
        self._stream = stream if stream is not None else torch.cuda.current_stream(self.device)
        h = ctypes.c_void_p()
        check(lib().epet_ctx_create(ctypes.byref(h), device,
                                    ctypes.c_void_p(self._stream.cuda_stream)))
        self._h = h
        self.setup: Optional[MeasurementSetup] = None
        self.grid: Optional[VoxelGrid] = None
        self.nrays = 8
        self.seed = 0

    @property
    def handle(self):
        return self._h

    def close(self):
        if self._h:
            lib().epet_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def synchronize(self):
        check(lib().epet_ctx_synchronize(self._h))

    # -- geometry (device_constant_memory.hpp:14-26) ------------------------
    def set_geometry(self, setup: MeasurementSetup, grid: VoxelGrid, nrays: int = None,
                     seed: int = None):
        cs, cg = setup.c(), grid.c()
        check(lib().epet_set_geometry(self._h, ctypes.byref(cs), ctypes.byref(cg)))
        self.setup, self.grid = setup, grid
        if nrays is not None:
            self.set_nrays(nrays)
        if seed is not None:
            self.set_seed(seed)

    def set_nrays(self, nrays: int):
        check(lib().epet_set_nrays(self._h, int(nrays)))
        self.nrays = int(nrays)

    def set_seed(self, seed: int):
        check(lib().epet_set_seed(self._h, int(seed)))
        self.seed = int(seed)

    def _i32(self, a) -> torch.Tensor:
        if isinstance(a, torch.Tensor):
            t = a.to(device=self.device, dtype=torch.int32)
        else:
            t = torch.as_tensor(np.asarray(a, dtype=np.int32), device=self.device)
        return t.contiguous()

    # -- rays -------------------------------------------------------------
    def rays(self, cnl) -> torch.Tensor:
        c = self._i32(cnl)
        out = torch.empty((c.numel(), self.nrays, 6), dtype=torch.float32, device=self.device)
        check(lib().epet_rays(self._h, _p(c), c.numel(), _p(out)))
        return out

    # -- workqueue ----------------------------------------------------------
    def getWorkqueue(self, list_ids, listId: int = 0, vxlId: int = 0, cap: int = -1,
                     capacity: Optional[int] = None):
        """Returns (cnl, vxl, listId, vxlId) like getWorkqueue's vectors and cursor."""
        lst = self._i32(list_ids)
        if capacity is None:
            capacity = cap if cap >= 0 else lst.numel() * self.grid.nvoxels
        capacity = max(int(capacity), 1)
        oc = torch.empty(capacity, dtype=torch.int32, device=self.device)
        ov = torch.empty(capacity, dtype=torch.int32, device=self.device)
        li, vi, n = ctypes.c_int(listId), ctypes.c_int(vxlId), ctypes.c_longlong(0)
        check(lib().epet_getWorkqueue(self._h, _p(lst), lst.numel(), ctypes.byref(li),
                                      ctypes.byref(vi), cap, _p(oc), _p(ov), capacity,
                                      ctypes.byref(n)))
        return oc[:n.value], ov[:n.value], li.value, vi.value

    def getSystemMatrixFromWorkqueue(self, cnl: torch.Tensor, vxl: torch.Tensor,
                                     nrays: Optional[int] = None) -> torch.Tensor:
        c, v = self._i32(cnl), self._i32(vxl)
        val = torch.empty(c.numel(), dtype=torch.float32, device=self.device)
        check(lib().epet_getSystemMatrixFromWorkqueue(self._h, _p(c), _p(v), _p(val), c.numel(),
                                                      int(nrays or self.nrays)))
        return val

    def getSystemMatrix(self, ml, capacity: int, keep_zeros: bool = True):
        """COO (sme, vxlId, cnlId) of the listed channels, row-major order."""
        m = self._i32(ml)
        sme = torch.empty(capacity, dtype=torch.float32, device=self.device)
        vxl = torch.empty(capacity, dtype=torch.int32, device=self.device)
        cnl = torch.empty(capacity, dtype=torch.int32, device=self.device)
        dest = torch.zeros(1, dtype=torch.int64, device=self.device)
        check(lib().epet_getSystemMatrix(self._h, _p(sme), _p(vxl), _p(cnl), _p(m), m.numel(),
                                         _p(dest), capacity, int(keep_zeros)))
        n = int(dest.item())
        return sme[:n], vxl[:n], cnl[:n], n

    def systemMatrixCalculation(self, yRowId, nchannels: int, capacity: int,
                                keep_zeros: bool = True):
        """ECSR (ecsrPtr int32[nY+1], vxlId, val) of the listed rows, plus nnz."""
        y = self._i32(yRowId)
        nY = y.numel()
        ecsr = torch.empty(nY + 1, dtype=torch.int32, device=self.device)
        vxl = torch.empty(capacity, dtype=torch.int32, device=self.device)
        val = torch.empty(capacity, dtype=torch.float32, device=self.device)
        cnl = torch.empty(capacity, dtype=torch.int32, device=self.device)
        csr = torch.empty(nchannels + 1, dtype=torch.int32, device=self.device)
        nnz = torch.zeros(1, dtype=torch.int64, device=self.device)
        check(lib().epet_systemMatrixCalculation(self._h, _p(ecsr), _p(vxl), _p(val), _p(nnz),
                                                 _p(cnl), _p(csr), _p(y), nY, nchannels,
                                                 capacity, int(keep_zeros)))
        n = int(nnz.item())
        return ecsr, vxl[:n], val[:n], n

    def sm_build(self, list_ids=None, m: Optional[int] = None) -> "Matrix":
        """Fused nonzero system matrix of the listed channels (library-owned ECSR)."""
        if list_ids is None:
            if m is None:
                m = self.setup.nchannels
            lst = None
        else:
            lst = self._i32(list_ids)
            m = lst.numel()
        h = ctypes.c_void_p()
        check(lib().epet_sm_build(self._h, _p(lst), int(m), ctypes.byref(h)))
        return Matrix(self, h)

    # -- sparse formats -------------------------------------------------------
    def cooSort(self, val: torch.Tensor, row: torch.Tensor, col: torch.Tensor):
        _check_dtype(val, torch.float32, "val")
        _check_dtype(row, torch.int32, "row")
        _check_dtype(col, torch.int32, "col")
        check(lib().epet_cooSort(self._h, _p(val), _p(row), _p(col), val.numel()))

    def convertCoo2Csr(self, cooRowId: torch.Tensor, m: int) -> torch.Tensor:
        _check_dtype(cooRowId, torch.int32, "cooRowId")
        out = torch.empty(m + 1, dtype=torch.int32, device=self.device)
        check(lib().epet_convertCoo2Csr(self._h, _p(out), _p(cooRowId), cooRowId.numel(), m))
        return out

    def convertCsr2Ecsr(self, yRowId: torch.Tensor, csr: torch.Tensor) -> torch.Tensor:
        _check_dtype(yRowId, torch.int32, "yRowId")
        _check_dtype(csr, torch.int32, "csr")
        out = torch.empty(yRowId.numel() + 1, dtype=torch.int32, device=self.device)
        check(lib().epet_convertCsr2Ecsr(self._h, _p(out), _p(yRowId), yRowId.numel(), _p(csr),
                                         csr.numel() - 1))
        return out

    # -- algebra ------------------------------------------------------------------
    def csrmv(self, op: int, m: int, n: int, alpha: float, val, rowptr, col, x, beta: float,
              y: torch.Tensor) -> torch.Tensor:
        a, b = ctypes.c_float(alpha), ctypes.c_float(beta)
        check(lib().epet_csrmv(self._h, op, m, n, val.numel(), ctypes.byref(a), _p(val),
                               _p(rowptr), _p(col), _p(x), ctypes.byref(b), _p(y)))
        return y

    def divides(self, res, a, b):
        check(lib().epet_divides(self._h, _p(res), _p(a), _p(b), res.numel()))
        return res

    def dividesMultiplies(self, res, a, b, c):
        check(lib().epet_dividesMultiplies(self._h, _p(res), _p(a), _p(b), _p(c), res.numel()))
        return res

    def scales(self, x, alpha: float):
        check(lib().epet_scales(self._h, _p(x), float(alpha), x.numel()))
        return x

    def sum(self, x) -> float:
        out = ctypes.c_double(0.0)
        check(lib().epet_sum(self._h, _p(x), x.numel(), ctypes.byref(out)))
        return out.value

    def getChords(self, rays: torch.Tensor, nChords: int):
        """Siddon chords of rays (n x 6) on the context's grid: (chords (n,
        nChords), voxelIds (n, nChords, 3), counts (n), -1 = capacity short)."""
        rays = rays.contiguous()
        n = rays.numel() // 6
        ch = torch.zeros(n, max(nChords, 1), dtype=torch.float32, device=self.device)
        vid = torch.zeros(n, max(nChords, 1), 3, dtype=torch.int32, device=self.device)
        cnt = torch.zeros(n, dtype=torch.int32, device=self.device)
        check(lib().epet_getChords(self._h, _p(rays), n, int(nChords), _p(ch), _p(vid), _p(cnt)))
        return ch, vid, cnt

    def mlem_tune(self, A: "Matrix", frames: int):
        """Autotune the multi-frame products for `frames`; returns
        ((rb_fwd, P_fwd, rb_bwd, P_bwd), (ms_fwd, ms_bwd)) for the blocked
        products, ((R_fwd, P_fwd, R_bwd, P_bwd), ...) for the gather kernels."""
        cfg = np.zeros(4, dtype=np.int32)
        ms = np.zeros(2, dtype=np.float32)
        check(lib().epet_mlem_tune(self._h, A.handle, int(frames),
                                   cfg.ctypes.data_as(_lib.c_int_p),
                                   ms.ctypes.data_as(_lib.c_float_p)))
        return tuple(int(v) for v in cfg), tuple(float(v) for v in ms)

    def batched_config(self, A: "Matrix", frames: int) -> dict:
        """The multi-frame path and launch shape epet_mlem runs for `frames`
        frames on A (as tuned by its first multi-frame call)."""
        info = np.zeros(10, dtype=np.int32)
        ms = np.zeros(2, dtype=np.float32)
        check(lib().epet_mlem_kernels(self._h, A.handle, int(frames),
                                      info.ctypes.data_as(_lib.c_int_p),
                                      ms.ctypes.data_as(_lib.c_float_p)))
        path = {2: "blocked", 1: "gather", 0: "untuned"}[int(info[0])]
        if path == "blocked":
            def shape(o, t):
                return {"frames_per_lane": int(info[o]), "lanes_per_row_slot": int(info[o + 1]),
                        "rows_per_block": int(info[o + 2]), "splits": int(info[o + 3]),
                        "tuned_ms": float(ms[t])}
            return {"path": "blocked (LDS-staged column tiles, entry-balanced slots, k_blk)",
                    "forward": shape(1, 0), "back_projection": shape(5, 1)}
        if path == "gather":
            return {"path": "gather kernels (k_mf_fwd / k_mf_bwd)",
                    "forward": {"frame_tile": int(info[1]), "rows_per_wavefront": int(info[2]),
                                "tuned_ms": float(ms[0])},
                    "back_projection": {"frame_tile": int(info[5]),
                                        "rows_per_wavefront": int(info[6]),
                                        "tuned_ms": float(ms[1])}}
        return {"path": path}

    def mlem(self, A: "Matrix", y: torch.Tensor, s: torch.Tensor, x: torch.Tensor,
             niter: int, use_graph: bool = False, tol: float = 0.0):
        """In-place MLEM on x (n x F, frame-minor); returns (loglik[niter, F], iters[F])."""
        F = 1 if x.dim() == 1 else x.shape[1]
        if not 1 <= F <= 64:
            raise EpetError(EPET_ERR_ARG, f"frames must be in [1, 64], got {F}")
        if y.dim() != x.dim() or (x.dim() == 2 and y.shape[1] != F):
            raise EpetError(EPET_ERR_ARG, f"y {tuple(y.shape)} and x {tuple(x.shape)} disagree on the frames")
        _check_vec(y, A.m * F, "y", self.device)
        _check_vec(s, A.n, "s", self.device)
        _check_vec(x, A.n * F, "x", self.device)
        ll = np.zeros((max(niter, 1), F), dtype=np.float64)
        iters = np.zeros(F, dtype=np.int32)
        check(lib().epet_mlem(self._h, A.handle, F, _p(y), _p(s), _p(x), niter, int(use_graph),
                              float(tol), ll.ctypes.data_as(_lib.c_double_p),
                              iters.ctypes.data_as(_lib.c_int_p)))
        return ll[:niter], iters

    def mlem_noalloc(self, A: "Matrix", y: torch.Tensor, s: torch.Tensor, x: torch.Tensor,
                     niter: int):
        """In-place single-frame MLEM on x that allocates no device memory
        (epet_mlem_noalloc); returns loglik[niter]."""
        _check_vec(y, A.m, "y", self.device)
        _check_vec(s, A.n, "s", self.device)
        _check_vec(x, A.n, "x", self.device)
        ll = np.zeros(max(niter, 1), dtype=np.float64)
        check(lib().epet_mlem_noalloc(self._h, A.handle, _p(y), _p(s), _p(x), niter,
                                      ll.ctypes.data_as(_lib.c_double_p)))
        return ll[:niter]


class Matrix:
    """Library-owned ECSR matrix (epet_matrix)."""

    def __init__(self, ctx: Context, handle):
        self.ctx = ctx
        self._h = handle
        m, n, vt = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        nnz = ctypes.c_longlong()
        check(lib().epet_matrix_info(handle, ctypes.byref(m), ctypes.byref(n), ctypes.byref(nnz),
                                     ctypes.byref(vt)))
        self.m, self.n, self.nnz, self.value_type = m.value, n.value, nnz.value, vt.value

    @property
    def handle(self):
        return self._h

    def close(self):
        if self._h:
            lib().epet_matrix_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @classmethod
    def from_csr(cls, ctx: Context, rowptr: torch.Tensor, col: torch.Tensor, val: torch.Tensor,
                 n: int) -> "Matrix":
        rp = rowptr.to(device=ctx.device, dtype=torch.int64).contiguous()
        c = col.to(device=ctx.device, dtype=torch.int32).contiguous()
        v = val.to(device=ctx.device, dtype=torch.float32).contiguous()
        h = ctypes.c_void_p()
        check(lib().epet_matrix_create(ctx.handle, rp.numel() - 1, n, c.numel(), _p(rp), _p(c),
                                       _p(v), ctypes.byref(h)))
        return cls(ctx, h)

    def arrays(self):
        """(rowptr, col, val, rowscale) device pointers as ints."""
        rp, c, v, s = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
        check(lib().epet_matrix_arrays(self._h, ctypes.byref(rp), ctypes.byref(c),
                                       ctypes.byref(v), ctypes.byref(s)))
        return rp.value, c.value, v.value, s.value

    def to_torch(self):
        """Copies (rowptr int64, col int32, val float32|int16 bits, rowscale|None)."""
        rp, c, v, s = self.arrays()
        dev = self.ctx.device
        trp = torch.empty(self.m + 1, dtype=torch.int64, device=dev)
        tc = torch.empty(max(self.nnz, 1), dtype=torch.int32, device=dev)
        vdt = torch.float32 if self.value_type == VAL_F32 else torch.int16
        tv = torch.empty(max(self.nnz, 1), dtype=vdt, device=dev)
        L, h = lib(), self.ctx.handle
        check(L.epet_memcpy_d2d(h, _p(trp), ctypes.c_void_p(rp), 8 * (self.m + 1)))
        if self.nnz:
            check(L.epet_memcpy_d2d(h, _p(tc), ctypes.c_void_p(c), 4 * self.nnz))
            check(L.epet_memcpy_d2d(h, _p(tv), ctypes.c_void_p(v), tv.element_size() * self.nnz))
        ts = None
        if s:
            ts = torch.empty(max(self.m, 1), dtype=torch.float32, device=dev)
            check(L.epet_memcpy_d2d(h, _p(ts), ctypes.c_void_p(s), 4 * self.m))
        self.ctx.synchronize()
        return trp, tc[:self.nnz], tv[:self.nnz], ts

    def spmv(self, op: int, x: torch.Tensor, y: Optional[torch.Tensor] = None,
             alpha: float = 1.0, beta: float = 0.0, ctx: Optional[Context] = None) -> torch.Tensor:
        """y = alpha op(A) x + beta y (on `ctx`, the matrix's own context by default)."""
        ctx = ctx or self.ctx
        if op not in (OP_N, OP_T):
            raise EpetError(EPET_ERR_ARG, f"unknown operation {op}")
        nin, nout = (self.n, self.m) if op == OP_N else (self.m, self.n)
        if y is None:
            y = torch.zeros(nout, dtype=torch.float32, device=self.ctx.device)
        _check_vec(x, nin, "x", self.ctx.device)
        _check_vec(y, nout, "y", self.ctx.device)
        check(lib().epet_matrix_spmv(ctx.handle, self._h, op, alpha, _p(x), beta, _p(y)))
        return y

    def spmm(self, op: int, X: torch.Tensor, Y: Optional[torch.Tensor] = None,
             alpha: float = 1.0, beta: float = 0.0) -> torch.Tensor:
        if op not in (OP_N, OP_T):
            raise EpetError(EPET_ERR_ARG, f"unknown operation {op}")
        if X.dim() != 2:
            raise EpetError(EPET_ERR_ARG, "X must be 2-D (rows x frames, frame-minor)")
        F = X.shape[1]
        nin, nout = (self.n, self.m) if op == OP_N else (self.m, self.n)
        if Y is None:
            Y = torch.zeros((nout, F), dtype=torch.float32, device=self.ctx.device)
        if Y.dim() != 2 or Y.shape[1] != F:
            raise EpetError(EPET_ERR_ARG, f"Y {tuple(Y.shape)} must have {F} frames")
        _check_vec(X, nin * F, "X", self.ctx.device)
        _check_vec(Y, nout * F, "Y", self.ctx.device)
        check(lib().epet_matrix_spmm(self.ctx.handle, self._h, op, F, alpha, _p(X), beta, _p(Y)))
        return Y

    def select_rows(self, rows) -> "Matrix":
        r = self.ctx._i32(rows)
        h = ctypes.c_void_p()
        check(lib().epet_matrix_select_rows(self.ctx.handle, self._h, _p(r), r.numel(),
                                            ctypes.byref(h)))
        return Matrix(self.ctx, h)

    def convert(self, value_type: int) -> "Matrix":
        h = ctypes.c_void_p()
        check(lib().epet_matrix_convert(self.ctx.handle, self._h, value_type, ctypes.byref(h)))
        return Matrix(self.ctx, h)

    def attach_transpose(self) -> "Matrix":
        """Build the transposed copy used by the multi-frame back projection."""
        check(lib().epet_matrix_transpose(self.ctx.handle, self._h))
        return self

    def colsum_f64(self) -> torch.Tensor:
        s = torch.empty(self.n, dtype=torch.float64, device=self.ctx.device)
        check(lib().epet_matrix_colsum_f64(self.ctx.handle, self._h, _p(s)))
        return s
