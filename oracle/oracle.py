"""ctypes wrapper of the CPU oracle (liboracle.so) and of the host-compiled
reference checker (_ref/libepetref.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke check and by the
CPU-baseline leg of bench.py, never by the product package.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
_REF = None

c_int_p = ctypes.POINTER(ctypes.c_int)
c_float_p = ctypes.POINTER(ctypes.c_float)
c_double_p = ctypes.POINTER(ctypes.c_double)
c_i64_p = ctypes.POINTER(ctypes.c_int64)
c_u32_p = ctypes.POINTER(ctypes.c_uint32)


def build(force: bool = False) -> None:
    """Compile liboracle.so (and _ref when the reference sources exist)."""
    if force or not os.path.exists(os.path.join(_HERE, "liboracle.so")):
        subprocess.check_call(["make", "-s", "-C", _HERE, "liboracle.so"])
    if os.path.isdir("/root/reference/src"):
        subprocess.check_call(["make", "-s", "-C", _HERE, "ref"])


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        P = ctypes.c_void_p
        sig = {
            "eo_ms_linid": (ctypes.c_int, [P] + [ctypes.c_int] * 5),
            "eo_ms_ida": (ctypes.c_int, [P, ctypes.c_int]),
            "eo_ms_id0z": (ctypes.c_int, [P, ctypes.c_int]),
            "eo_ms_id0y": (ctypes.c_int, [P, ctypes.c_int]),
            "eo_ms_id1z": (ctypes.c_int, [P, ctypes.c_int]),
            "eo_ms_id1y": (ctypes.c_int, [P, ctypes.c_int]),
            "eo_vg_linid": (ctypes.c_int, [P] + [ctypes.c_int] * 3),
            "eo_vg_idx": (ctypes.c_int, [P, ctypes.c_int]),
            "eo_vg_idy": (ctypes.c_int, [P, ctypes.c_int]),
            "eo_vg_idz": (ctypes.c_int, [P, ctypes.c_int]),
            "eo_angle_sincos": (None, [P, ctypes.c_int, c_float_p, c_float_p]),
            "eo_trafo": (None, [P, ctypes.c_int, c_float_p, ctypes.c_int,
                                ctypes.c_int, ctypes.c_int, c_float_p]),
            "eo_distance": (ctypes.c_float, [c_float_p, c_float_p, c_float_p]),
            "eo_is_candidate": (ctypes.c_int, [P, P, ctypes.c_int, ctypes.c_int]),
            "eo_workqueue": (ctypes.c_longlong, [P, P, c_int_p, ctypes.c_int, c_int_p,
                                                 c_int_p, ctypes.c_longlong, c_int_p,
                                                 c_int_p]),
            "eo_philox4x32_10": (None, [c_u32_p, c_u32_p, c_u32_p]),
            "eo_u01": (ctypes.c_float, [ctypes.c_uint32]),
            "eo_ray": (None, [P, ctypes.c_int, ctypes.c_int, ctypes.c_uint64, c_float_p]),
            "eo_ray_center": (None, [P, ctypes.c_int, c_float_p]),
            "eo_intersection_length": (ctypes.c_float, [c_float_p, c_float_p]),
            "eo_voxel_box": (None, [P, ctypes.c_int, c_float_p]),
            "eo_calc_sme": (ctypes.c_float, [P, P, ctypes.c_int, ctypes.c_int,
                                             ctypes.c_int, ctypes.c_uint64]),
            "eo_calc_sme_rays": (ctypes.c_float, [P, ctypes.c_int, c_float_p, ctypes.c_int]),
            "eo_sm_from_workqueue": (None, [P, P, c_int_p, c_int_p, c_float_p,
                                            ctypes.c_longlong, ctypes.c_int,
                                            ctypes.c_uint64]),
            "eo_sm_rows": (ctypes.c_longlong, [P, P, c_int_p, ctypes.c_int, ctypes.c_int,
                                               ctypes.c_uint64, ctypes.c_int, c_i64_p,
                                               c_int_p, c_float_p, ctypes.c_longlong]),
            "eo_get_chords": (ctypes.c_int, [c_float_p, c_int_p, ctypes.c_int, c_float_p,
                                             c_float_p, c_float_p, c_int_p]),
            "eo_coo_sort": (None, [c_float_p, c_int_p, c_int_p, ctypes.c_longlong]),
            "eo_coo2csr": (None, [c_int_p, c_int_p, ctypes.c_longlong, ctypes.c_int]),
            "eo_csr2ecsr": (None, [c_int_p, c_int_p, ctypes.c_int, c_int_p, ctypes.c_int]),
            "eo_csrmv": (None, [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_float,
                                c_float_p, c_i64_p, c_int_p, c_float_p, ctypes.c_float,
                                c_float_p]),
            "eo_divides": (None, [c_float_p, c_float_p, c_float_p, ctypes.c_longlong]),
            "eo_divides_multiplies": (None, [c_float_p, c_float_p, c_float_p, c_float_p,
                                             ctypes.c_longlong]),
            "eo_scales": (None, [c_float_p, ctypes.c_float, ctypes.c_longlong]),
            "eo_sum": (ctypes.c_float, [c_float_p, ctypes.c_longlong]),
            "eo_sum_f64": (ctypes.c_double, [c_float_p, ctypes.c_longlong]),
            "eo_mlem_f32": (None, [ctypes.c_int, ctypes.c_int, c_float_p, c_i64_p, c_int_p,
                                   c_float_p, c_float_p, c_float_p, ctypes.c_int]),
            "eo_mlem_f64": (None, [ctypes.c_int, ctypes.c_int, c_float_p, c_i64_p, c_int_p,
                                   c_float_p, c_float_p, c_double_p, ctypes.c_int]),
            "eo_mlem_f64_frames": (None, [ctypes.c_int, ctypes.c_int, ctypes.c_int, c_float_p,
                                          c_i64_p, c_int_p, c_float_p, c_float_p, c_double_p,
                                          c_int_p, ctypes.c_int]),
            "eo_mlem_f64_frames_par": (None, [ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                              c_float_p, c_i64_p, c_int_p, c_float_p,
                                              c_float_p, c_double_p, c_int_p, ctypes.c_int]),
            "eo_colsum_f64": (None, [ctypes.c_int, ctypes.c_int, c_float_p, c_i64_p, c_int_p,
                                     c_double_p, ctypes.c_int]),
            "eo_mlem_iter_omp": (None, [ctypes.c_int, ctypes.c_int, c_float_p, c_i64_p,
                                        c_int_p, c_float_p, c_float_p, c_float_p,
                                        c_float_p, c_float_p, ctypes.c_int]),
            "eo_loglik": (ctypes.c_double, [ctypes.c_int, c_float_p, c_i64_p, c_int_p,
                                            c_float_p, c_double_p]),
            "eo_nInChunk": (ctypes.c_uint32, [ctypes.c_uint32] * 3),
            "eo_chunkPtr": (ctypes.c_uint32, [ctypes.c_uint32] * 2),
            "eo_nChunks": (ctypes.c_uint32, [ctypes.c_uint64] * 2),
            "eo_max_threads": (ctypes.c_int, []),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _LIB = L
    return _LIB


def ref():
    """The host-compiled reference checker, or None where it is not built."""
    global _REF
    if _REF is None:
        path = os.path.join(_HERE, "_ref", "libepetref.so")
        if not os.path.exists(path):
            return None
        L = ctypes.CDLL(path)
        P = ctypes.c_void_p
        L.ref_workqueue.restype = ctypes.c_longlong
        L.ref_workqueue.argtypes = [P, P, c_int_p, ctypes.c_int, c_int_p, c_int_p,
                                    ctypes.c_int, c_int_p, c_int_p, ctypes.c_longlong]
        L.ref_trafo.restype = None
        L.ref_trafo.argtypes = [P, ctypes.c_int, ctypes.c_int, c_float_p, ctypes.c_int,
                                ctypes.c_int, ctypes.c_int, c_float_p]
        L.ref_distance.restype = ctypes.c_float
        L.ref_distance.argtypes = [c_float_p, c_float_p, c_float_p]
        L.ref_ms_linid.restype = ctypes.c_int
        L.ref_ms_linid.argtypes = [P] + [ctypes.c_int] * 5
        L.ref_ms_decompose.restype = None
        L.ref_ms_decompose.argtypes = [P, ctypes.c_int, c_int_p]
        L.ref_vg_linid.restype = ctypes.c_int
        L.ref_vg_linid.argtypes = [P] + [ctypes.c_int] * 3
        L.ref_vg_decompose.restype = None
        L.ref_vg_decompose.argtypes = [P, ctypes.c_int, c_int_p]
        _REF = L
    return _REF


# ---------------------------------------------------------------- helpers
def ptr(a: np.ndarray, ctype):
    return a.ctypes.data_as(ctypes.POINTER(ctype))


def _sg(setup, grid):
    cs, cg = setup.c(), grid.c()
    return cs, cg, ctypes.byref(cs), ctypes.byref(cg)


def philox(ctr, key) -> np.ndarray:
    c = np.asarray(ctr, dtype=np.uint32)
    k = np.asarray(key, dtype=np.uint32)
    out = np.zeros(4, dtype=np.uint32)
    lib().eo_philox4x32_10(ptr(c, ctypes.c_uint32), ptr(k, ctypes.c_uint32),
                           ptr(out, ctypes.c_uint32))
    return out


def rays(setup, cnl: int, nrays: int, seed: int) -> np.ndarray:
    cs = setup.c()
    out = np.zeros((nrays, 6), dtype=np.float32)
    for r in range(nrays):
        lib().eo_ray(ctypes.byref(cs), int(cnl), r, seed, ptr(out[r], ctypes.c_float))
    return out


def trafo(setup, det, box, idz, idy, ida) -> np.ndarray:
    cs = setup.c()
    b = np.ascontiguousarray(box, dtype=np.float32)
    out = np.zeros(3, dtype=np.float32)
    lib().eo_trafo(ctypes.byref(cs), det, ptr(b, ctypes.c_float), idz, idy, ida,
                   ptr(out, ctypes.c_float))
    return out


def intersection_length(vxl, ray) -> float:
    v = np.ascontiguousarray(vxl, dtype=np.float32)
    r = np.ascontiguousarray(ray, dtype=np.float32)
    return lib().eo_intersection_length(ptr(v, ctypes.c_float), ptr(r, ctypes.c_float))


def workqueue(setup, grid, lst, listId=0, vxlId=0, cap=-1, want=True):
    """Returns (cnl, vxl, listId, vxlId) of the oracle's getWorkqueue."""
    cs, cg, ps, pg = _sg(setup, grid)
    lst = np.ascontiguousarray(lst, dtype=np.int32)
    li, vi = ctypes.c_int(listId), ctypes.c_int(vxlId)
    n = lib().eo_workqueue(ps, pg, ptr(lst, ctypes.c_int), len(lst), ctypes.byref(li),
                           ctypes.byref(vi), cap, None, None)
    if not want:
        return n
    cnl = np.zeros(max(n, 1), dtype=np.int32)
    vxl = np.zeros(max(n, 1), dtype=np.int32)
    li, vi = ctypes.c_int(listId), ctypes.c_int(vxlId)
    lib().eo_workqueue(ps, pg, ptr(lst, ctypes.c_int), len(lst), ctypes.byref(li),
                       ctypes.byref(vi), cap, ptr(cnl, ctypes.c_int), ptr(vxl, ctypes.c_int))
    return cnl[:n], vxl[:n], li.value, vi.value


def ref_workqueue(setup, grid, lst, listId=0, vxlId=0, cap=-1, out_cap=None):
    R = ref()
    if R is None:
        return None
    cs, cg, ps, pg = _sg(setup, grid)
    lst = np.ascontiguousarray(lst, dtype=np.int32)
    if out_cap is None:
        out_cap = len(lst) * grid.nvoxels if cap < 0 else cap
    cnl = np.zeros(max(out_cap, 1), dtype=np.int32)
    vxl = np.zeros(max(out_cap, 1), dtype=np.int32)
    li, vi = ctypes.c_int(listId), ctypes.c_int(vxlId)
    n = R.ref_workqueue(ps, pg, ptr(lst, ctypes.c_int), len(lst), ctypes.byref(li),
                        ctypes.byref(vi), cap, ptr(cnl, ctypes.c_int),
                        ptr(vxl, ctypes.c_int), out_cap)
    k = min(n, out_cap)
    return cnl[:k], vxl[:k], li.value, vi.value


def sm_rows(setup, grid, lst, nrays, seed, keep_zeros=False):
    """Oracle system matrix of the listed channels: (rowptr int64, col, val)."""
    cs, cg, ps, pg = _sg(setup, grid)
    lst = np.ascontiguousarray(lst, dtype=np.int32)
    m = len(lst)
    rowptr = np.zeros(m + 1, dtype=np.int64)
    nnz = lib().eo_sm_rows(ps, pg, ptr(lst, ctypes.c_int), m, nrays, seed,
                           int(keep_zeros), ptr(rowptr, ctypes.c_int64), None, None, -1)
    col = np.zeros(max(nnz, 1), dtype=np.int32)
    val = np.zeros(max(nnz, 1), dtype=np.float32)
    n2 = lib().eo_sm_rows(ps, pg, ptr(lst, ctypes.c_int), m, nrays, seed, int(keep_zeros),
                          ptr(rowptr, ctypes.c_int64), ptr(col, ctypes.c_int),
                          ptr(val, ctypes.c_float), nnz)
    assert n2 == nnz
    return rowptr, col[:nnz], val[:nnz]


def sm_from_workqueue(setup, grid, cnl, vxl, nrays, seed):
    cs, cg, ps, pg = _sg(setup, grid)
    cnl = np.ascontiguousarray(cnl, dtype=np.int32)
    vxl = np.ascontiguousarray(vxl, dtype=np.int32)
    val = np.zeros(len(cnl), dtype=np.float32)
    lib().eo_sm_from_workqueue(ps, pg, ptr(cnl, ctypes.c_int), ptr(vxl, ctypes.c_int),
                               ptr(val, ctypes.c_float), len(cnl), nrays, seed)
    return val


def get_chords(ray, grid, cap=1024):
    r = np.ascontiguousarray(ray, dtype=np.float32)
    O = np.array([grid.ox, grid.oy, grid.oz], dtype=np.float32)
    D = np.array([grid.dx, grid.dy, grid.dz], dtype=np.float32)
    N = np.array([grid.nx, grid.ny, grid.nz], dtype=np.int32)
    ch = np.zeros(cap, dtype=np.float32)
    ids = np.zeros(3 * cap, dtype=np.int32)
    k = lib().eo_get_chords(ptr(ch, ctypes.c_float), ptr(ids, ctypes.c_int), cap,
                            ptr(r, ctypes.c_float), ptr(O, ctypes.c_float),
                            ptr(D, ctypes.c_float), ptr(N, ctypes.c_int))
    if k < 0:
        raise ValueError("chord capacity exceeded")
    return ch[:k], ids[:3 * k].reshape(k, 3)


def csrmv(trans, m, n, alpha, val, rowptr, col, x, beta, y):
    y = np.ascontiguousarray(y, dtype=np.float32).copy()
    val = np.ascontiguousarray(val, dtype=np.float32)
    rowptr = np.ascontiguousarray(rowptr, dtype=np.int64)
    col = np.ascontiguousarray(col, dtype=np.int32)
    x = np.ascontiguousarray(x, dtype=np.float32)
    lib().eo_csrmv(int(trans), m, n, alpha, ptr(val, ctypes.c_float),
                   ptr(rowptr, ctypes.c_int64), ptr(col, ctypes.c_int),
                   ptr(x, ctypes.c_float), beta, ptr(y, ctypes.c_float))
    return y


def mlem_f32(m, n, val, rowptr, col, y, s, x0, niter):
    x = np.ascontiguousarray(x0, dtype=np.float32).copy()
    lib().eo_mlem_f32(m, n, ptr(np.ascontiguousarray(val, np.float32), ctypes.c_float),
                      ptr(np.ascontiguousarray(rowptr, np.int64), ctypes.c_int64),
                      ptr(np.ascontiguousarray(col, np.int32), ctypes.c_int),
                      ptr(np.ascontiguousarray(y, np.float32), ctypes.c_float),
                      ptr(np.ascontiguousarray(s, np.float32), ctypes.c_float),
                      ptr(x, ctypes.c_float), niter)
    return x


def mlem_f64(m, n, val, rowptr, col, y, s, x0, niter):
    x = np.ascontiguousarray(x0, dtype=np.float64).copy()
    lib().eo_mlem_f64(m, n, ptr(np.ascontiguousarray(val, np.float32), ctypes.c_float),
                      ptr(np.ascontiguousarray(rowptr, np.int64), ctypes.c_int64),
                      ptr(np.ascontiguousarray(col, np.int32), ctypes.c_int),
                      ptr(np.ascontiguousarray(y, np.float32), ctypes.c_float),
                      ptr(np.ascontiguousarray(s, np.float32), ctypes.c_float),
                      ptr(x, ctypes.c_double), niter)
    return x


def mlem_f64_frames(m, n, val, rowptr, col, Y, s, X0, niter, nthreads=0):
    """F-frame FP64 MLEM (frame-minor Y (m, F), X0 (n, F)); niter is an int or
    one count per frame."""
    X = np.ascontiguousarray(X0, dtype=np.float64).copy()
    F = X.shape[1]
    its = np.broadcast_to(np.asarray(niter, np.int32), (F,)).copy()
    lib().eo_mlem_f64_frames(m, n, F, ptr(np.ascontiguousarray(val, np.float32), ctypes.c_float),
                             ptr(np.ascontiguousarray(rowptr, np.int64), ctypes.c_int64),
                             ptr(np.ascontiguousarray(col, np.int32), ctypes.c_int),
                             ptr(np.ascontiguousarray(Y, np.float32), ctypes.c_float),
                             ptr(np.ascontiguousarray(s, np.float32), ctypes.c_float),
                             ptr(X, ctypes.c_double), ptr(its, ctypes.c_int), nthreads)
    return X


def mlem_f64_frames_par(m, n, val, rowptr, col, Y, s, X0, niter, nthreads=0):
    """mlem_f64_frames for full-size matrices (per-thread scatter back
    projection, no transpose); same arguments."""
    X = np.ascontiguousarray(X0, dtype=np.float64).copy()
    if X.ndim == 1:
        X = X[:, None].copy()
    F = X.shape[1]
    Y = np.ascontiguousarray(Y, dtype=np.float32)
    if Y.ndim == 1:
        Y = Y[:, None]
    its = np.broadcast_to(np.asarray(niter, np.int32), (F,)).copy()
    lib().eo_mlem_f64_frames_par(m, n, F,
                                 ptr(np.ascontiguousarray(val, np.float32), ctypes.c_float),
                                 ptr(np.ascontiguousarray(rowptr, np.int64), ctypes.c_int64),
                                 ptr(np.ascontiguousarray(col, np.int32), ctypes.c_int),
                                 ptr(np.ascontiguousarray(Y), ctypes.c_float),
                                 ptr(np.ascontiguousarray(s, np.float32), ctypes.c_float),
                                 ptr(X, ctypes.c_double), ptr(its, ctypes.c_int), nthreads)
    return X


def colsum_f64(n, val, rowptr, col, nthreads=0):
    """Sensitivity image sum_i a_ij in FP64 over the rows of a CSR matrix."""
    rowptr = np.ascontiguousarray(rowptr, np.int64)
    out = np.zeros(n, dtype=np.float64)
    lib().eo_colsum_f64(len(rowptr) - 1, n, ptr(np.ascontiguousarray(val, np.float32), ctypes.c_float),
                        ptr(rowptr, ctypes.c_int64),
                        ptr(np.ascontiguousarray(col, np.int32), ctypes.c_int),
                        ptr(out, ctypes.c_double), nthreads)
    return out


def loglik(val, rowptr, col, y, x):
    x = np.ascontiguousarray(x, dtype=np.float64)
    m = len(rowptr) - 1
    return lib().eo_loglik(m, ptr(np.ascontiguousarray(val, np.float32), ctypes.c_float),
                           ptr(np.ascontiguousarray(rowptr, np.int64), ctypes.c_int64),
                           ptr(np.ascontiguousarray(col, np.int32), ctypes.c_int),
                           ptr(np.ascontiguousarray(y, np.float32), ctypes.c_float),
                           ptr(x, ctypes.c_double))
