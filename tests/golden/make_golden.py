"""Write the golden fixtures of tests/golden/ (run in a container that has the
reference sources, after `make -C oracle ref`).

Two kinds of fixture:
  * known answers held inline by the reference's own test programs
    (test_cooSort.cu:17-19, test_cusparseWrapper.cu:26-28,
    test_convertCsr2Ecsr.cu:7-8, test_mlemOperations.cu:39-44), as data;
  * input/output vectors of the reference's host-side geometry and workqueue
    templates, computed by oracle/_ref/libepetref.so (the reference's own
    headers compiled from its sources), for the small (11,250 channels, 16^3)
    and full (5,140,980 channels, 52^3) workloads.
Only data is written; no reference source text.
"""
from __future__ import annotations

import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402
from epetreco_amd import geometry as G  # noqa: E402


def pairs_sha256(cnl: np.ndarray, vxl: np.ndarray) -> str:
    a = np.stack([cnl.astype(np.int32), vxl.astype(np.int32)], axis=1)
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def kats():
    n = 10
    A = [(i + 1) ** 2 for i in range(n)]
    B = [10] * n
    C = [i + 1 for i in range(n)]
    return {
        "cooSort": {"row": [1, 3, 2, 2, 1, 3], "col": [2, 2, 1, 2, 1, 1],
                    "val": [2, 6, 3, 4, 1, 5],
                    "expect_row": [1, 1, 2, 2, 3, 3], "expect_col": [1, 2, 1, 2, 1, 2],
                    "expect_val": [1, 2, 3, 4, 5, 6]},
        "coo2csr": {"row": [0, 1, 1, 1, 2, 2, 3, 3, 3], "m": 4, "expect": [0, 1, 4, 6, 9]},
        "csr2ecsr": {"yRowId": [2, 4, 6], "csr": [0, 0, 0, 3, 3, 3, 3, 5], "m": 7,
                     "expect": [0, 3, 3, 5]},
        "mlemOperations": {"A": A, "B": B, "C": C,
                           "expect_divides": [a / c for a, c in zip(A, C)],
                           "expect_dividesMultiplies": [a * b / c for a, b, c in zip(A, B, C)],
                           "expect_sum": 550.0,
                           "expect_scaled": [a * b / c / 550.0 for a, b, c in zip(A, B, C)]},
    }


def trafo_points(R, setup, rng, nsample=None):
    cs = setup.c()
    boxes = [(0.5, 0.5, 0.5)] + [(a, b, c) for a in (0, 1) for b in (0, 1) for c in (0, 1)]
    rows = []
    if nsample is None:
        for det in (0, 1):
            nz = setup.n1z if det else setup.n0z
            ny = setup.n1y if det else setup.n0y
            for ida in range(setup.na):
                for iz in range(nz):
                    for iy in range(ny):
                        for b in boxes:
                            rows.append((det, iz, iy, ida) + b)
    else:
        for _ in range(nsample):
            det = int(rng.integers(2))
            rows.append((det, int(rng.integers(setup.n0z)), int(rng.integers(setup.n0y)),
                         int(rng.integers(setup.na)), *rng.random(3).astype(np.float32)))
    ids = np.array([r[:4] for r in rows], dtype=np.int32)
    box = np.array([r[4:] for r in rows], dtype=np.float32)
    out = np.zeros((len(rows), 3), dtype=np.float32)
    out_inp = np.zeros((len(rows), 3), dtype=np.float32)
    import ctypes
    for k in range(len(rows)):
        b = np.ascontiguousarray(box[k])
        R.ref_trafo(ctypes.byref(cs), int(ids[k, 0]), 0, O.ptr(b, ctypes.c_float),
                    int(ids[k, 1]), int(ids[k, 2]), int(ids[k, 3]), O.ptr(out[k], ctypes.c_float))
        R.ref_trafo(ctypes.byref(cs), int(ids[k, 0]), 1, O.ptr(b, ctypes.c_float),
                    int(ids[k, 1]), int(ids[k, 2]), int(ids[k, 3]),
                    O.ptr(out_inp[k], ctypes.c_float))
    return ids, box, out, out_inp


def main():
    R = O.ref()
    if R is None:
        raise SystemExit("oracle/_ref not built: make -C oracle ref")
    rng = np.random.default_rng(20141009)
    with open(os.path.join(HERE, "ref_kats.json"), "w") as f:
        json.dump(kats(), f, indent=1)

    # ---- small workload: full workqueue hash, sampled lists, capped cursors
    s, g, _ = G.workload("small")
    lst = np.arange(s.nchannels, dtype=np.int32)
    cnl, vxl, li, vi = O.ref_workqueue(s, g, lst)
    sample = np.sort(rng.choice(s.nchannels, 50, replace=False)).astype(np.int32)
    per = {int(c): vxl[cnl == c] for c in sample}
    caps = []
    cur = (0, 0)
    for cap in (1000, 1, 4095, 12345, 1):
        c2, v2, l2, w2 = O.ref_workqueue(s, g, lst, cur[0], cur[1], cap)
        caps.append((cur[0], cur[1], cap, len(c2), l2, w2, pairs_sha256(c2, v2)))
        cur = (l2, w2)
    # a cursor in the middle of a row and one at the last voxel of a row
    for start in ((7, 2000), (11, g.nvoxels - 1)):
        c2, v2, l2, w2 = O.ref_workqueue(s, g, lst, start[0], start[1], 500)
        caps.append((start[0], start[1], 500, len(c2), l2, w2, pairs_sha256(c2, v2)))
    ids, box, out, out_inp = trafo_points(R, s, rng)
    np.savez_compressed(
        os.path.join(HERE, "small_geometry.npz"),
        wq_count=np.int64(len(cnl)), wq_sha256=np.array(pairs_sha256(cnl, vxl)),
        wq_end_cursor=np.array([li, vi], dtype=np.int32),
        sample_cnl=sample,
        sample_ptr=np.cumsum([0] + [len(per[int(c)]) for c in sample]).astype(np.int64),
        sample_vxl=np.concatenate([per[int(c)] for c in sample]).astype(np.int32),
        caps=np.array([c[:6] for c in caps], dtype=np.int64),
        caps_sha256=np.array([c[6] for c in caps]),
        trafo_ids=ids, trafo_box=box, trafo_out=out, trafo_out_inplace=out_inp)

    # ---- full workload: sampled channels' candidate lists, sampled geometry
    s2, g2, _ = G.workload("full")
    sample2 = np.sort(rng.choice(s2.nchannels, 8, replace=False)).astype(np.int32)
    lists = []
    for c in sample2:
        c3, v3, _, _ = O.ref_workqueue(s2, g2, np.array([c], dtype=np.int32))
        lists.append(v3)
    ids2, box2, out2, out2_inp = trafo_points(R, s2, rng, nsample=4000)
    # distances of random points from random lines
    a = rng.uniform(-0.25, 0.25, (2000, 3)).astype(np.float32)
    b = rng.uniform(-0.25, 0.25, (2000, 3)).astype(np.float32)
    p = rng.uniform(-0.03, 0.03, (2000, 3)).astype(np.float32)
    import ctypes
    d = np.array([R.ref_distance(O.ptr(np.ascontiguousarray(a[k]), ctypes.c_float),
                                 O.ptr(np.ascontiguousarray(b[k]), ctypes.c_float),
                                 O.ptr(np.ascontiguousarray(p[k]), ctypes.c_float))
                  for k in range(len(a))], dtype=np.float32)
    # channel / voxel index functors
    lin = rng.integers(0, s2.nchannels, 5000).astype(np.int32)
    cs2 = s2.c()
    dec = np.zeros((len(lin), 5), dtype=np.int32)
    for k, v in enumerate(lin):
        R.ref_ms_decompose(ctypes.byref(cs2), int(v), O.ptr(dec[k], ctypes.c_int))
    np.savez_compressed(
        os.path.join(HERE, "full_geometry.npz"),
        sample_cnl=sample2,
        sample_ptr=np.cumsum([0] + [len(v) for v in lists]).astype(np.int64),
        sample_vxl=np.concatenate(lists).astype(np.int32),
        trafo_ids=ids2, trafo_box=box2, trafo_out=out2, trafo_out_inplace=out2_inp,
        dist_a=a, dist_b=b, dist_p=p, dist_out=d,
        ms_lin=lin, ms_dec=dec)
    print("golden fixtures written:", len(cnl), "small-workload candidates;",
          [len(v) for v in lists], "full-workload candidates per sampled channel")


if __name__ == "__main__":
    main()
