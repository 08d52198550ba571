"""CPU checks that pin the oracle (oracle/epet_oracle.c) to the reference.

Each oracle function is checked against something other than itself: the
reference's own host-compiled templates (golden fixtures written by
tests/golden/make_golden.py from oracle/_ref), the reference tests' known
answers, published known-answer vectors, an independent implementation, or a
closed form / brute force chosen so that a dropped term, swapped operand or
wrong index fails.
"""
import json
import os
import shutil
import subprocess
import tempfile

import numpy as np
import pytest

from conftest import GOLDEN, ROOT
from epetreco_amd import geometry as G

SMALL = G.workload("small")
FULL = G.workload("full")
KATS = json.load(open(os.path.join(GOLDEN, "ref_kats.json")))


# ------------------------------------------------------------------ Philox
# Random123 known-answer vectors for philox4x32_10 (kat_vectors file).
PHILOX_KAT = [
    ((0, 0, 0, 0), (0, 0), (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)),
    ((0xffffffff,) * 4, (0xffffffff,) * 2, (0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)),
    ((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344), (0xa4093822, 0x299f31d0),
     (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1)),
]


@pytest.mark.parametrize("ctr,key,expect", PHILOX_KAT)
def test_philox_known_answers(oracle, ctr, key, expect):
    assert tuple(int(v) for v in oracle.philox(ctr, key)) == expect


def test_philox_matches_rocrand_header(oracle):
    hipcc = "/opt/rocm/bin/hipcc"
    if not os.path.exists(hipcc) or not os.path.exists("/opt/rocm/include/rocrand"):
        pytest.skip("hipcc / rocRAND headers not present")
    with tempfile.TemporaryDirectory() as d:
        exe = os.path.join(d, "philox_check")
        r = subprocess.run([hipcc, "-x", "hip", "--offload-host-only", "-O1", "-o", exe,
                            os.path.join(ROOT, "tests", "philox_rocrand_check.cpp")],
                           capture_output=True, text=True)
        if r.returncode != 0:
            pytest.skip("could not build the rocRAND checker: " + r.stderr[-300:])
        rng = np.random.default_rng(5)
        cases = rng.integers(0, 2 ** 32, (500, 6), dtype=np.uint64)
        inp = "\n".join(" ".join(f"{int(v):x}" for v in c) for c in cases) + "\n"
        out = subprocess.run([exe], input=inp, capture_output=True, text=True, check=True)
        lines = out.stdout.strip().split("\n")
        assert len(lines) == len(cases)
        for c, line in zip(cases, lines):
            got = oracle.philox(c[:4], c[4:])
            assert [int(t, 16) for t in line.split()] == [int(v) for v in got]


def test_u01_is_open_interval_of_odd_multiples(oracle):
    L = oracle.lib()
    for x in (0, 1, 255, 256, 2 ** 31, 2 ** 32 - 1, 0x12345678):
        u = L.eo_u01(x)
        assert 0.0 < u < 1.0
        assert ((x >> 8) | 1) == int(u * 2 ** 24)


def test_rays_endpoints_inside_segments(oracle):
    """Ray endpoints lie inside their segment boxes: invert the closed-form trafo."""
    s, g, nr = FULL
    rng = np.random.default_rng(2)
    for cnl in rng.integers(0, s.nchannels, 30):
        ida, i0z, i0y, i1z, i1y = (int(v) for v in s.decompose(cnl))
        th = np.pi * ida * s.da / 180.0
        rays = oracle.rays(s, int(cnl), nr, 99).astype(np.float64)
        for det, p, iz, iy, ny, nz in ((0, rays[:, :3], i0z, i0y, s.n0y, s.n0z),
                                        (1, rays[:, 3:], i1z, i1y, s.n1y, s.n1z)):
            pos = np.array([s.pos1x if det else s.pos0x, (iy - 0.5 * (ny - 1)) * s.segy,
                            (iz - 0.5 * (nz - 1)) * s.segz])
            # undo R_y(theta): world = R (pos + (b - 1/2) e)
            c, sn = np.cos(th), np.sin(th)
            local = np.stack([c * p[:, 0] - sn * p[:, 2], p[:, 1], sn * p[:, 0] + c * p[:, 2]], 1)
            b = (local - pos) / np.array([s.segx, s.segy, s.segz]) + 0.5
            assert np.all(b > -1e-5) and np.all(b < 1 + 1e-5)
        assert len(np.unique(rays, axis=0)) == nr   # distinct Philox streams per ray


# ------------------------------------------------------------------ index maps
def test_channel_index_roundtrip(oracle):
    """Every channel id round-trips (test_MeasurementSetupLinIndex.cu:37-46)."""
    s = FULL[0]
    lin = np.arange(s.nchannels)
    ida, a, b, c, d = s.decompose(lin)
    assert np.array_equal(s.lin_id(a, b, c, d, ida), lin)
    z = np.load(os.path.join(GOLDEN, "full_geometry.npz"))
    cs = s.c()
    import ctypes
    L = oracle.lib()
    for v, dec in zip(z["ms_lin"][:500], z["ms_dec"][:500]):
        mine = [L.eo_ms_ida(ctypes.byref(cs), int(v)), L.eo_ms_id0z(ctypes.byref(cs), int(v)),
                L.eo_ms_id0y(ctypes.byref(cs), int(v)), L.eo_ms_id1z(ctypes.byref(cs), int(v)),
                L.eo_ms_id1y(ctypes.byref(cs), int(v))]
        assert mine == list(dec)
        assert list(dec) == [int(t) for t in s.decompose(int(v))]
        assert L.eo_ms_linid(ctypes.byref(cs), *[int(t) for t in dec[1:]], int(dec[0])) == v


def test_voxel_index_roundtrip_z_fastest(oracle):
    """test_VoxelGridLinIndex.cu:37-44; voxel id = (idx*ny + idy)*nz + idz."""
    g = FULL[1]
    lin = np.arange(g.nvoxels)
    ix, iy, iz = g.decompose(lin)
    assert np.array_equal(g.lin_id(ix, iy, iz), lin)
    assert g.lin_id(0, 0, 1) == 1 and g.lin_id(0, 1, 0) == g.nz and g.lin_id(1, 0, 0) == g.ny * g.nz
    import ctypes
    cg = g.c()
    L = oracle.lib()
    for v in (0, 1, 51, 52, 2703, 2704, 140607, 77777):
        assert (L.eo_vg_idx(ctypes.byref(cg), v), L.eo_vg_idy(ctypes.byref(cg), v),
                L.eo_vg_idz(ctypes.byref(cg), v)) == tuple(int(t) for t in g.decompose(v))


# ------------------------------------------------------------------ trafo / distance
@pytest.mark.parametrize("which", ["small", "full"])
def test_trafo_bit_exact_vs_reference(oracle, which):
    z = np.load(os.path.join(GOLDEN, f"{which}_geometry.npz"))
    s = (SMALL if which == "small" else FULL)[0]
    ids, box, ref, ref_inp = z["trafo_ids"], z["trafo_box"], z["trafo_out"], z["trafo_out_inplace"]
    got = np.array([oracle.trafo(s, int(i[0]), b, int(i[1]), int(i[2]), int(i[3]))
                    for i, b in zip(ids, box)])
    assert np.array_equal(got, ref)
    assert np.array_equal(ref_inp, ref)   # inplace and non-inplace functors agree
    # closed form R_y(theta) (pos + (b - 1/2) e) in double
    th = np.pi * ids[:, 3] * s.da / 180.0
    det = ids[:, 0]
    ny = np.where(det == 1, s.n1y, s.n0y)
    nz = np.where(det == 1, s.n1z, s.n0z)
    loc = np.stack([np.where(det == 1, s.pos1x, s.pos0x) + (box[:, 0] - 0.5) * s.segx,
                    (ids[:, 2] - 0.5 * (ny - 1)) * s.segy + (box[:, 1] - 0.5) * s.segy,
                    (ids[:, 1] - 0.5 * (nz - 1)) * s.segz + (box[:, 2] - 0.5) * s.segz], 1)
    world = np.stack([np.cos(th) * loc[:, 0] + np.sin(th) * loc[:, 2], loc[:, 1],
                      -np.sin(th) * loc[:, 0] + np.cos(th) * loc[:, 2]], 1)
    np.testing.assert_allclose(got, world, atol=2e-7)


def test_distance_bit_exact_and_closed_form(oracle):
    z = np.load(os.path.join(GOLDEN, "full_geometry.npz"))
    import ctypes
    L = oracle.lib()
    a, b, p, ref = z["dist_a"], z["dist_b"], z["dist_p"], z["dist_out"]
    got = np.array([L.eo_distance(oracle.ptr(np.ascontiguousarray(a[k]), ctypes.c_float),
                                  oracle.ptr(np.ascontiguousarray(b[k]), ctypes.c_float),
                                  oracle.ptr(np.ascontiguousarray(p[k]), ctypes.c_float))
                    for k in range(len(a))], dtype=np.float32)
    assert np.array_equal(got, ref)
    A, B, P = (t.astype(np.float64) for t in (a, b, p))
    d = np.linalg.norm(np.cross(B - A, P - A), axis=1) / np.linalg.norm(B - A, axis=1)
    np.testing.assert_allclose(got, d, rtol=2e-5, atol=1e-8)


# ------------------------------------------------------------------ workqueue
def _sha(c, v):
    import hashlib
    a = np.stack([np.asarray(c, np.int32), np.asarray(v, np.int32)], axis=1)
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def test_workqueue_small_vs_reference(oracle):
    s, g, _ = SMALL
    z = np.load(os.path.join(GOLDEN, "small_geometry.npz"))
    lst = np.arange(s.nchannels, dtype=np.int32)
    cnl, vxl, li, vi = oracle.workqueue(s, g, lst)
    assert len(cnl) == int(z["wq_count"]) and _sha(cnl, vxl) == str(z["wq_sha256"])
    assert [li, vi] == list(z["wq_end_cursor"])
    for k, c in enumerate(z["sample_cnl"]):
        np.testing.assert_array_equal(vxl[cnl == c],
                                      z["sample_vxl"][z["sample_ptr"][k]:z["sample_ptr"][k + 1]])
    for row, sha in zip(z["caps"], z["caps_sha256"]):
        l0, v0, cap, n, l1, v1 = (int(t) for t in row)
        c2, w2, l2, u2 = oracle.workqueue(s, g, lst, l0, v0, cap)
        assert (len(c2), l2, u2) == (n, l1, v1) and _sha(c2, w2) == str(sha)


def test_workqueue_full_sampled_vs_reference(oracle):
    s, g, _ = FULL
    z = np.load(os.path.join(GOLDEN, "full_geometry.npz"))
    for k, c in enumerate(z["sample_cnl"]):
        _, vxl, li, vi = oracle.workqueue(s, g, np.array([c], np.int32))
        np.testing.assert_array_equal(vxl, z["sample_vxl"][z["sample_ptr"][k]:z["sample_ptr"][k + 1]])


def test_workqueue_capped_is_prefix(oracle):
    """test_getWorkqueue.cu:186-198: a capped run is a prefix of the full run,
    and resuming from the cursor continues it."""
    s, g, _ = SMALL
    lst = np.arange(40, 90, dtype=np.int32)
    cnl, vxl, _, _ = oracle.workqueue(s, g, lst)
    got_c, got_v, cur = [], [], (0, 0)
    while True:
        c2, v2, l2, w2 = oracle.workqueue(s, g, lst, cur[0], cur[1], 777)
        got_c.append(c2); got_v.append(v2)
        cur = (l2, w2)
        if len(c2) < 777 or l2 >= len(lst):
            break
    np.testing.assert_array_equal(np.concatenate(got_c), cnl)
    np.testing.assert_array_equal(np.concatenate(got_v), vxl)


def test_candidate_brute_force_double(oracle):
    """Candidates = voxel centres within sumRadii of the pixel-centre line, in double,
    away from ties."""
    s, g, _ = FULL
    rng = np.random.default_rng(4)
    cen = g.centres()
    sr = 0.5 * (np.linalg.norm([g.dx, g.dy, g.dz]) + np.linalg.norm([s.segx, s.segy, s.segz]))
    for cnl in rng.integers(0, s.nchannels, 4):
        ida, i0z, i0y, i1z, i1y = (int(v) for v in s.decompose(cnl))
        p0 = oracle.trafo(s, 0, [0.5, 0.5, 0.5], i0z, i0y, ida).astype(np.float64)
        p1 = oracle.trafo(s, 1, [0.5, 0.5, 0.5], i1z, i1y, ida).astype(np.float64)
        d = np.linalg.norm(np.cross(p1 - p0, cen - p0), axis=1) / np.linalg.norm(p1 - p0)
        _, vxl, _, _ = oracle.workqueue(s, g, np.array([cnl], np.int32))
        inside = np.zeros(g.nvoxels, bool)
        inside[vxl] = True
        clear = np.abs(d - sr) > 1e-6
        assert np.array_equal(inside[clear], (d < sr)[clear])


# ------------------------------------------------------------------ intersection length
def _seg_box_length(box, ray):
    """Exact (double) length of the infinite line through ray inside box."""
    p = ray[:3].astype(np.float64)
    dvec = (ray[3:] - ray[:3]).astype(np.float64)
    lo, hi = -np.inf, np.inf
    for k in range(3):
        if dvec[k] == 0:
            continue
        a, b = (box[k] - p[k]) / dvec[k], (box[3 + k] - p[k]) / dvec[k]
        lo, hi = max(lo, min(a, b)), min(hi, max(a, b))
    return max(0.0, hi - lo) * np.linalg.norm(dvec)


def test_intersection_closed_forms(oracle):
    d = 0.001
    box = np.array([0, 0, 0, d, d, d], np.float32)
    # through the centre parallel to x (y, z not crossed); short rays keep the
    # reference's alpha cancellation (|err| ~ ulp(0.5) * ray length) small
    assert oracle.intersection_length(box, [-d, d / 2, d / 2, 2 * d, d / 2, d / 2]) == pytest.approx(d, rel=1e-6)
    # a 2 m ray: error bounded by a few ulp(0.5) times its length
    assert abs(oracle.intersection_length(box, [-1, d / 2, d / 2, 1, d / 2, d / 2]) - d) < 3e-7
    # main diagonal: d*sqrt(3)
    assert oracle.intersection_length(box, [-d, -d, -d, 2 * d, 2 * d, 2 * d]) == pytest.approx(d * 3 ** 0.5, rel=1e-5)
    # a miss
    assert oracle.intersection_length(box, [-1, 2 * d, 2 * d, 1, 2 * d, 3 * d]) == 0.0
    # degenerate ray (start == end)
    assert oracle.intersection_length(box, [0.5 * d] * 6) == 0.0
    # reference quirk: axis-parallel lines outside the slab are not rejected
    # (getSystemMatrixDeviceOnly.cu:96-118) -- y is ignored here
    assert oracle.intersection_length(box, [-d, 5 * d, d / 2, 2 * d, 5 * d, d / 2]) == pytest.approx(d, rel=1e-6)


def test_intersection_brute_force(oracle):
    rng = np.random.default_rng(8)
    g = FULL[1]
    for _ in range(3000):
        v = int(rng.integers(g.nvoxels))
        ix, iy, iz = (int(t) for t in g.decompose(v))
        box = np.array([g.ox + ix * g.dx, g.oy + iy * g.dy, g.oz + iz * g.dz,
                        g.ox + (ix + 1) * g.dx, g.oy + (iy + 1) * g.dy, g.oz + (iz + 1) * g.dz],
                       np.float32)
        c = (box[:3] + box[3:]) / 2
        # long rays (0.46 m) passing within ~1.5 mm of the voxel centre
        u = rng.normal(size=3); u /= np.linalg.norm(u)
        q = c + rng.uniform(-1.5e-3, 1.5e-3, 3)
        ray = np.concatenate([q - 0.23 * u, q + 0.23 * u]).astype(np.float32)
        got = oracle.intersection_length(box, ray)
        ref = _seg_box_length(box.astype(np.float64), ray)
        assert abs(got - ref) < 1e-7, (got, ref)


# ------------------------------------------------------------------ system matrix rows
def test_sm_rows_structures(oracle):
    """keep-zeros rows = workqueue candidates; nonzero rows = candidates with value != 0;
    values = mean over rays of exact intersection lengths."""
    s, g, nr = SMALL
    lst = np.array([3, 500, 4000, 9999], np.int32)
    rp0, c0, v0 = oracle.sm_rows(s, g, lst, nr, 7, keep_zeros=True)
    cnl, vxl, _, _ = oracle.workqueue(s, g, lst)
    np.testing.assert_array_equal(c0, vxl)
    rp1, c1, v1 = oracle.sm_rows(s, g, lst, nr, 7, keep_zeros=False)
    keep = v0 != 0
    np.testing.assert_array_equal(c1, c0[keep])
    np.testing.assert_array_equal(v1, v0[keep])
    # values against the double-precision mean of exact lengths
    for i, c in enumerate(lst):
        rays = oracle.rays(s, int(c), nr, 7)
        for k in range(rp1[i], rp1[i + 1], 7):
            ix, iy, iz = (int(t) for t in g.decompose(int(c1[k])))
            box = np.array([g.ox + ix * g.dx, g.oy + iy * g.dy, g.oz + iz * g.dz,
                            g.ox + (ix + 1) * g.dx, g.oy + (iy + 1) * g.dy,
                            g.oz + (iz + 1) * g.dz], np.float32).astype(np.float64)
            ref = np.mean([_seg_box_length(box, r) for r in rays])
            assert abs(v1[k] - ref) < 2e-7
    # FromWorkqueue values of the same pairs are the same numbers
    vw = oracle.sm_from_workqueue(s, g, cnl, vxl, nr, 7)
    np.testing.assert_array_equal(vw, v0)


# ------------------------------------------------------------------ chords
def _siddon_double(ray, g):
    """Independent double-precision Siddon: chord lengths and voxel ids of the
    segment of the ray inside the grid, in ray order."""
    p = ray[:3].astype(np.float64)
    d = (ray[3:] - ray[:3]).astype(np.float64)
    O = np.array([g.ox, g.oy, g.oz]); D = np.array([g.dx, g.dy, g.dz])
    N = np.array([g.nx, g.ny, g.nz])
    al = [0.0, 1.0]
    lo, hi = 0.0, 1.0
    for k in range(3):
        if d[k] == 0:
            continue
        a = (O[k] + np.arange(N[k] + 1) * D[k] - p[k]) / d[k]
        lo, hi = max(lo, a.min()), min(hi, a.max())
        al.extend(a)
    al = np.unique(np.clip(al, lo, hi))
    L = np.linalg.norm(d)
    mids = 0.5 * (al[1:] + al[:-1])
    ids = np.floor((p + mids[:, None] * d - O) / D).astype(int)
    return np.diff(al) * L, ids


def test_get_chords_corner_to_corner(oracle):
    """Rays entering within the first voxel and leaving within the last voxel of
    every axis (the rays getChords is written for): chords and voxel ids equal an
    independent double Siddon."""
    g = G.VoxelGrid(0.0, 0.0, 0.0, 1.0, 1.0, 1.0, 6, 5, 4)
    rng = np.random.default_rng(3)
    for _ in range(100):
        a = rng.uniform(0.1, 0.9, 3)
        b = np.array([g.nx, g.ny, g.nz]) - rng.uniform(0.1, 0.9, 3)
        u = b - a
        ray = np.concatenate([a - 0.3 * u, b + 0.3 * u]).astype(np.float32)
        ch, ids = oracle.get_chords(ray, g, cap=256)
        ref_ch, ref_ids = _siddon_double(ray, g)
        keep = ref_ch > 1e-4
        ref_ch, ref_ids = ref_ch[keep], ref_ids[keep]
        kk = ch > 1e-4
        ch, ids = ch[kk], ids[kk]
        n = len(ref_ch)
        np.testing.assert_allclose(ch[:n], ref_ch, atol=2e-5)
        np.testing.assert_array_equal(ids[:n], ref_ids)
        # quirk (3): the loop may write one chord past the exit plane; its voxel
        # id lies outside the grid
        assert len(ch) <= n + 1
        if len(ch) == n + 1:
            assert not np.all((ids[n] >= 0) & (ids[n] < [g.nx, g.ny, g.nz]))


def test_get_chords_reference_quirks(oracle):
    """Behaviour kept from ChordsCalc_lowlevel.hpp: (1) an axis that is not
    crossed still takes part in the crossing comparison (:312), so a ray
    parallel to y and z through a 2x2x2 grid yields two extra zero chords;
    (2) aDimnext starts at aDimmin + aDimup (:242), so a ray entering through a
    face away from the first plane of another axis starts with a negative chord;
    (3) the incremental parameter update can leave aCurr just below aMax, and the
    loop (:292) then writes one chord beyond the exit plane."""
    g = G.VoxelGrid(0, 0, 0, 1, 1, 1, 2, 2, 2)
    ch, ids = oracle.get_chords(np.array([-1, 0.5, 0.5, 3, 0.5, 0.5], np.float32), g)
    np.testing.assert_allclose(ch, [1.0, 0.0, 0.0, 1.0])
    assert ch.sum() == pytest.approx(2.0)
    g = G.VoxelGrid(-0.026, -0.026, -0.026, 0.001, 0.001, 0.001, 52, 52, 52)
    ray = np.array([-0.2, 0.01, 0.003, 0.2, -0.004, 0.008], np.float32)
    ch, _ = oracle.get_chords(ray, g, cap=4096)
    assert ch[0] < 0


# ------------------------------------------------------------------ sparse formats
def test_coo_sort_kat(oracle):
    k = KATS["cooSort"]
    r, c, v = (np.array(k[n], t) for n, t in (("row", np.int32), ("col", np.int32),
                                              ("val", np.float32)))
    import ctypes
    oracle.lib().eo_coo_sort(oracle.ptr(v, ctypes.c_float), oracle.ptr(r, ctypes.c_int),
                             oracle.ptr(c, ctypes.c_int), len(v))
    assert r.tolist() == k["expect_row"] and c.tolist() == k["expect_col"]
    assert v.tolist() == k["expect_val"]


def test_coo2csr_and_csr2ecsr_kats(oracle):
    import ctypes
    k = KATS["coo2csr"]
    rows = np.array(k["row"], np.int32)
    out = np.zeros(k["m"] + 1, np.int32)
    oracle.lib().eo_coo2csr(oracle.ptr(out, ctypes.c_int), oracle.ptr(rows, ctypes.c_int),
                            len(rows), k["m"])
    assert out.tolist() == k["expect"]
    k = KATS["csr2ecsr"]
    y, csr = np.array(k["yRowId"], np.int32), np.array(k["csr"], np.int32)
    out = np.zeros(len(y) + 1, np.int32)
    oracle.lib().eo_csr2ecsr(oracle.ptr(out, ctypes.c_int), oracle.ptr(y, ctypes.c_int), len(y),
                             oracle.ptr(csr, ctypes.c_int), k["m"])
    assert out.tolist() == k["expect"]


def test_mlem_operations_kats(oracle):
    import ctypes
    k = KATS["mlemOperations"]
    A, B, C = (np.array(k[n], np.float32) for n in "ABC")
    D = np.zeros_like(A)
    L = oracle.lib()
    f = lambda a: oracle.ptr(a, ctypes.c_float)
    L.eo_divides(f(D), f(A), f(C), len(A))
    assert D.tolist() == k["expect_divides"]
    L.eo_divides_multiplies(f(D), f(A), f(B), f(C), len(A))
    assert D.tolist() == k["expect_dividesMultiplies"]
    assert L.eo_sum(f(D), len(D)) == k["expect_sum"]
    L.eo_scales(f(D), np.float32(1 / 550.0), len(D))
    np.testing.assert_allclose(D, k["expect_scaled"], rtol=1e-6)


def _rand_csr(rng, m, n, p):
    M = (rng.random((m, n)) < p) * rng.random((m, n))
    rp = np.concatenate([[0], np.cumsum((M != 0).sum(1))]).astype(np.int64)
    col = np.nonzero(M)[1].astype(np.int32)
    val = M[M != 0].astype(np.float32)
    return M.astype(np.float32), rp, col, val


def test_csrmv_dense_and_adjoint(oracle):
    rng = np.random.default_rng(0)
    M, rp, col, val = _rand_csr(rng, 60, 80, 0.2)
    x, e = rng.random(80).astype(np.float32), rng.random(60).astype(np.float32)
    y = oracle.csrmv(0, 60, 80, 1.5, val, rp, col, x, 0.0, np.zeros(60))
    np.testing.assert_allclose(y, 1.5 * M.astype(np.float64) @ x, rtol=1e-6)
    c = oracle.csrmv(1, 60, 80, 1.0, val, rp, col, e, 0.0, np.zeros(80))
    np.testing.assert_allclose(c, M.astype(np.float64).T @ e, rtol=1e-6)
    assert abs(np.dot(y / 1.5, e) - np.dot(x, c)) < 1e-4 * abs(np.dot(x, c))


# ------------------------------------------------------------------ MLEM
def test_mlem_identity_gives_normalised_data(oracle):
    n = 50
    rp, col, val = np.arange(n + 1), np.arange(n, dtype=np.int32), np.ones(n, np.float32)
    y = np.random.default_rng(1).random(n).astype(np.float32) + 0.1
    x0 = np.full(n, 1.0 / n)
    x = oracle.mlem_f64(n, n, val, rp, col, y, np.ones(n, np.float32), x0, 1)
    y64 = y.astype(np.float64)
    np.testing.assert_allclose(x, y64 / y64.sum(), rtol=1e-12)
    x32 = oracle.mlem_f32(n, n, val, rp, col, y, np.ones(n, np.float32), x0, 1)
    np.testing.assert_allclose(x32, y / y.sum(), rtol=1e-6)


def test_mlem_invariants_and_convergence(oracle):
    rng = np.random.default_rng(2)
    m, n = 400, 60
    M, rp, col, val = _rand_csr(rng, m, n, 0.3)
    xt = rng.random(n) + 0.2
    y = (M.astype(np.float64) @ xt).astype(np.float32)   # consistent, noise-free
    s = M.sum(0).astype(np.float32)
    x0 = np.full(n, 1.0 / n)
    lls, x = [], x0
    for it in range(8):
        lls.append(oracle.loglik(val, rp, col, y, x * y.sum()))
        x = oracle.mlem_f64(m, n, val, rp, col, y, s, x, 1)
        assert abs(x.sum() - 1) < 1e-12 and (x >= 0).all()
    assert np.all(np.diff(lls) >= -1e-9 * np.abs(lls[:-1]))
    x = oracle.mlem_f64(m, n, val, rp, col, y, s, x0, 3000)
    np.testing.assert_allclose(x, xt / xt.sum(), rtol=2e-3)
    # the FP32 reference loop stays within 1e-4 of FP64
    x32 = oracle.mlem_f32(m, n, val, rp, col, y, s, x0.astype(np.float32), 20)
    x64 = oracle.mlem_f64(m, n, val, rp, col, y, s, x0, 20)
    assert np.linalg.norm(x32 - x64) / np.linalg.norm(x64) < 1e-4


def test_mlem_omp_iteration_matches_serial(oracle):
    import ctypes
    rng = np.random.default_rng(6)
    m, n = 3000, 500
    M, rp, col, val = _rand_csr(rng, m, n, 0.05)
    y = rng.poisson(50, m).astype(np.float32)
    s = M.sum(0).astype(np.float32)
    s = (s / s.sum()).astype(np.float32)
    x = np.full(n, 1.0 / n, np.float32)
    yh, c = np.zeros(m, np.float32), np.zeros(n, np.float32)
    f = lambda a, t: oracle.ptr(a, t)
    oracle.lib().eo_mlem_iter_omp(m, n, f(val, ctypes.c_float), f(rp, ctypes.c_int64),
                                  f(col, ctypes.c_int), f(y, ctypes.c_float), f(s, ctypes.c_float),
                                  f(x, ctypes.c_float), f(yh, ctypes.c_float),
                                  f(c, ctypes.c_float), 4)
    ref = oracle.mlem_f64(m, n, val, rp, col, y, s, np.full(n, 1.0 / n), 1)
    assert np.linalg.norm(x - ref) / np.linalg.norm(ref) < 1e-5


def test_mlem_frames_matches_single_frame_bitwise(oracle):
    """Each frame of the multi-frame FP64 oracle equals eo_mlem_f64 on that
    frame alone, bit for bit, including frames frozen at their own counts and
    an empty column (zero sensitivity)."""
    rng = np.random.default_rng(7)
    m, n, F = 700, 90, 5
    M, rp, col, val = _rand_csr(rng, m, n, 0.1)
    keep = col != 17                       # column 17 stays empty
    M[:, 17] = 0
    rp = np.concatenate([[0], np.cumsum((M != 0).sum(1))]).astype(np.int64)
    col, val = col[keep], val[keep]
    Y = rng.poisson(20, (m, F)).astype(np.float32)
    s = M.sum(0).astype(np.float32)
    X0 = rng.random((n, F)) + 0.1
    its = np.array([0, 1, 4, 9, 6], np.int32)
    X = oracle.mlem_f64_frames(m, n, val, rp, col, Y, s, X0, its, nthreads=3)
    for f in range(F):
        xf = oracle.mlem_f64(m, n, val, rp, col, Y[:, f], s, X0[:, f], int(its[f]))
        assert np.array_equal(X[:, f], xf), f


def test_mlem_frames_par_matches_transpose_path(oracle):
    """The full-size FP64 oracle (per-thread scatter back projection) equals
    the bitwise-pinned transpose form up to FP64 summation order, for several
    thread counts, frames frozen at their own counts and an empty column."""
    rng = np.random.default_rng(8)
    m, n, F = 1500, 120, 4
    M, rp, col, val = _rand_csr(rng, m, n, 0.1)
    keep = col != 33
    M[:, 33] = 0
    rp = np.concatenate([[0], np.cumsum((M != 0).sum(1))]).astype(np.int64)
    col, val = col[keep], val[keep]
    Y = rng.poisson(20, (m, F)).astype(np.float32)
    s = M.sum(0).astype(np.float32)
    X0 = rng.random((n, F)) + 0.1
    its = np.array([0, 3, 7, 5], np.int32)
    Xa = oracle.mlem_f64_frames(m, n, val, rp, col, Y, s, X0, its, nthreads=2)
    for nt in (1, 3, 5):
        Xb = oracle.mlem_f64_frames_par(m, n, val, rp, col, Y, s, X0, its, nthreads=nt)
        np.testing.assert_allclose(Xb, Xa, rtol=1e-12, atol=0)
    np.testing.assert_array_equal(Xa[:, 0], X0[:, 0])   # zero iterations: unchanged


def test_colsum_f64_closed_form(oracle):
    """FP64 sensitivity image = dense column sums (all-ones back projection);
    an empty column gives exactly 0; the thread count changes nothing beyond
    FP64 summation order."""
    rng = np.random.default_rng(9)
    m, n = 900, 75
    M, rp, col, val = _rand_csr(rng, m, n, 0.2)
    keep = col != 5
    M[:, 5] = 0
    rp = np.concatenate([[0], np.cumsum((M != 0).sum(1))]).astype(np.int64)
    col, val = col[keep], val[keep]
    ref = M.astype(np.float64).sum(0)
    for nt in (1, 4):
        cs = oracle.colsum_f64(n, val, rp, col, nthreads=nt)
        np.testing.assert_allclose(cs, ref, rtol=1e-13)
        assert cs[5] == 0.0
    ones = np.ones(m, np.float32)
    # the same image as the oracle's transposed product with a ones vector
    np.testing.assert_allclose(cs, oracle.csrmv(1, m, n, 1.0, val, rp, col, ones, 0.0,
                                                np.zeros(n, np.float32)), rtol=2e-6)


# ------------------------------------------------------------------ chunk helpers
def test_chunk_helpers(oracle):
    """wrappers.hpp:405-445 semantics: full chunks, remainder, zero past the end."""
    L = oracle.lib()
    nAll, full = 1000, 300
    assert [L.eo_nInChunk(i, nAll, full) for i in range(5)] == [300, 300, 300, 100, 0]
    assert [L.eo_nInChunk(i, 900, 300) for i in range(4)] == [300, 300, 300, 0]
    assert L.eo_chunkPtr(3, full) == 900
    assert L.eo_nChunks(10 ** 12 + 1, 10 ** 6) == 10 ** 6 + 1
    assert sum(L.eo_nInChunk(i, nAll, full) for i in range(L.eo_nChunks(nAll, full))) == nAll
