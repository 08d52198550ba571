"""ctypes binding of libepet.so (the C-ABI declared in include/epet.h).

torch is imported before the library is loaded: torch ships its own HIP
runtime with the same SONAME (libamdhip64.so.7), so loading torch first makes
libepet.so bind to that one runtime instead of a second copy.  There is no CPU
fallback: a missing library raises at import of the symbol table.
"""
from __future__ import annotations

import ctypes
import os

import torch  # noqa: F401  (must precede the CDLL load, see above)

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libepet.so")

c_int_p = ctypes.POINTER(ctypes.c_int)
c_ll_p = ctypes.POINTER(ctypes.c_longlong)
c_double_p = ctypes.POINTER(ctypes.c_double)
c_float_p = ctypes.POINTER(ctypes.c_float)
P = ctypes.c_void_p

EPET_OK = 0
ERRORS = {1: "EPET_ERR_ARG", 2: "EPET_ERR_HIP", 3: "EPET_ERR_CAPACITY", 4: "EPET_ERR_STATE",
          5: "EPET_ERR_ALLOC", 6: "EPET_ERR_UNSUPPORTED"}
VAL_F32, VAL_BF16, VAL_F16 = 0, 1, 2
OP_N, OP_T = 0, 1

# name -> (restype, argtypes); the declarations of include/epet.h
SIGNATURES = {
    "epet_last_error": (ctypes.c_char_p, []),
    "epet_version": (ctypes.c_int, []),
    "epet_ctx_create": (ctypes.c_int, [ctypes.POINTER(P), ctypes.c_int, P]),
    "epet_ctx_destroy": (ctypes.c_int, [P]),
    "epet_ctx_set_stream": (ctypes.c_int, [P, P]),
    "epet_ctx_synchronize": (ctypes.c_int, [P]),
    "epet_set_geometry": (ctypes.c_int, [P, P, P]),
    "epet_set_nrays": (ctypes.c_int, [P, ctypes.c_int]),
    "epet_set_seed": (ctypes.c_int, [P, ctypes.c_uint64]),
    "epet_malloc": (ctypes.c_int, [P, ctypes.POINTER(P), ctypes.c_size_t]),
    "epet_free": (ctypes.c_int, [P, P]),
    "epet_memcpy_h2d": (ctypes.c_int, [P, P, P, ctypes.c_size_t]),
    "epet_memcpy_d2h": (ctypes.c_int, [P, P, P, ctypes.c_size_t]),
    "epet_memcpy_d2d": (ctypes.c_int, [P, P, P, ctypes.c_size_t]),
    "epet_rays": (ctypes.c_int, [P, P, ctypes.c_int, P]),
    "epet_getWorkqueue": (ctypes.c_int, [P, P, ctypes.c_int, c_int_p, c_int_p, ctypes.c_longlong,
                                         P, P, ctypes.c_longlong, c_ll_p]),
    "epet_getSystemMatrixFromWorkqueue": (ctypes.c_int, [P, P, P, P, ctypes.c_longlong,
                                                         ctypes.c_int]),
    "epet_getSystemMatrix": (ctypes.c_int, [P, P, P, P, P, ctypes.c_uint32, P, ctypes.c_longlong,
                                            ctypes.c_int]),
    "epet_systemMatrixCalculation": (ctypes.c_int, [P, P, P, P, P, P, P, P, ctypes.c_uint32,
                                                    ctypes.c_int, ctypes.c_longlong,
                                                    ctypes.c_int]),
    "epet_sm_build": (ctypes.c_int, [P, P, ctypes.c_int, ctypes.POINTER(P)]),
    "epet_matrix_info": (ctypes.c_int, [P, c_int_p, c_int_p, c_ll_p, c_int_p]),
    "epet_matrix_arrays": (ctypes.c_int, [P, ctypes.POINTER(P), ctypes.POINTER(P),
                                          ctypes.POINTER(P), ctypes.POINTER(P)]),
    "epet_matrix_destroy": (ctypes.c_int, [P]),
    "epet_matrix_create": (ctypes.c_int, [P, ctypes.c_int, ctypes.c_int, ctypes.c_longlong, P, P,
                                          P, ctypes.POINTER(P)]),
    "epet_matrix_select_rows": (ctypes.c_int, [P, P, P, ctypes.c_int, ctypes.POINTER(P)]),
    "epet_matrix_convert": (ctypes.c_int, [P, P, ctypes.c_int, ctypes.POINTER(P)]),
    "epet_matrix_colsum_f64": (ctypes.c_int, [P, P, P]),
    "epet_matrix_transpose": (ctypes.c_int, [P, P]),
    "epet_cooSort": (ctypes.c_int, [P, P, P, P, ctypes.c_longlong]),
    "epet_convertCoo2Csr": (ctypes.c_int, [P, P, P, ctypes.c_longlong, ctypes.c_int]),
    "epet_convertCsr2Ecsr": (ctypes.c_int, [P, P, P, ctypes.c_int, P, ctypes.c_int]),
    "epet_csrmv": (ctypes.c_int, [P, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_longlong,
                                  c_float_p, P, P, P, P, c_float_p, P]),
    "epet_matrix_spmv": (ctypes.c_int, [P, P, ctypes.c_int, ctypes.c_float, P, ctypes.c_float, P]),
    "epet_matrix_spmm": (ctypes.c_int, [P, P, ctypes.c_int, ctypes.c_int, ctypes.c_float, P,
                                        ctypes.c_float, P]),
    "epet_divides": (ctypes.c_int, [P, P, P, P, ctypes.c_longlong]),
    "epet_dividesMultiplies": (ctypes.c_int, [P, P, P, P, P, ctypes.c_longlong]),
    "epet_scales": (ctypes.c_int, [P, P, ctypes.c_float, ctypes.c_longlong]),
    "epet_sum": (ctypes.c_int, [P, P, ctypes.c_longlong, c_double_p]),
    "epet_mlem": (ctypes.c_int, [P, P, ctypes.c_int, P, P, P, ctypes.c_int, ctypes.c_int,
                                 ctypes.c_double, c_double_p, c_int_p]),
    "epet_getChords": (ctypes.c_int, [P, P, ctypes.c_int, ctypes.c_int, P, P, P]),
    "epet_mlem_noalloc": (ctypes.c_int, [P, P, P, P, P, ctypes.c_int, c_double_p]),
    "epet_mlem_tune": (ctypes.c_int, [P, P, ctypes.c_int, c_int_p, c_float_p]),
    "epet_mlem_kernels": (ctypes.c_int, [P, P, ctypes.c_int, c_int_p, c_float_p]),
    "epet_nInChunk": (ctypes.c_uint32, [ctypes.c_uint32] * 3),
    "epet_chunkPtr": (ctypes.c_uint32, [ctypes.c_uint32] * 2),
    "epet_nChunks": (ctypes.c_uint32, [ctypes.c_ulonglong] * 2),
}

_LIB = None


class EpetError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


def lib() -> ctypes.CDLL:
    """The loaded library; raises if it has not been built."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} not built: run `python -m epetreco_amd.build`")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _LIB = L
    return _LIB


def check(rc: int) -> None:
    if rc != EPET_OK:
        raise EpetError(rc, lib().epet_last_error().decode())
