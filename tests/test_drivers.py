"""Command-line drivers with the reference's HDF5 formats (tests/reco.cu:44-60,
tests/backprojection.cu:39-50; H5Reader.hpp, H5DensityWriter.hpp).

CPU: the HDF5 round trip of a measurement ("messung") and of a density
("density" with xmin..zmax / xnbin..znbin attributes) through the epet_h5 tool.
GPU: `reco` on a small-workload measurement file matches the FP64 oracle's
MLEM on the oracle's own matrix of the same channels, iteration by iteration
(rel-L2 <= 1e-4), with the in-tree sensitivity and with a sensitivity file;
`backprojection` matches the oracle matrix's FP64 A^T y / sum (<= 1e-5).
"""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "epetreco_amd", "bin")


def _tool(name):
    exe = os.path.join(BIN, name)
    if not os.path.exists(exe):
        from epetreco_amd import build as B
        if not B.build_drivers():
            pytest.skip("HDF5 headers not found: drivers not built")
    return exe


def _run(args, **kw):
    r = subprocess.run(args, capture_output=True, text=True, **kw)
    assert r.returncode == 0, r.stdout + r.stderr
    return r.stdout


def test_h5_roundtrip(tmp_path):
    h5 = _tool("epet_h5")
    rng = np.random.default_rng(0)
    meas = rng.poisson(2.0, 11250).astype(np.float32)
    meas.tofile(tmp_path / "m.f32")
    _run([h5, "mkmeas", str(tmp_path / "m.h5"), str(tmp_path / "m.f32")])
    _run([h5, "dump", str(tmp_path / "m.h5"), "messung", str(tmp_path / "m2.f32")])
    assert np.array_equal(np.fromfile(tmp_path / "m2.f32", np.float32), meas)
    dens = rng.random(16 ** 3).astype(np.float32)
    dens.tofile(tmp_path / "d.f32")
    _run([h5, "mkdens", str(tmp_path / "d.h5"), str(tmp_path / "d.f32"), "small"])
    _run([h5, "dump", str(tmp_path / "d.h5"), "density", str(tmp_path / "d2.f32")])
    assert np.array_equal(np.fromfile(tmp_path / "d2.f32", np.float32), dens)
    attrs = dict(l.split() for l in _run([h5, "attrs", str(tmp_path / "d.h5")]).split("\n") if l)
    assert int(attrs["xnbin"]) == 16 and int(attrs["znbin"]) == 16
    np.testing.assert_allclose(float(attrs["xmin"]), -0.024, rtol=1e-6)
    np.testing.assert_allclose(float(attrs["zmax"]), -0.024 + 16 * 0.003, rtol=1e-6)


@pytest.mark.gpu
def test_reco_and_backprojection_drivers(gpu_ctx_factory, oracle, tmp_path):
    """The drivers against the oracle, not against the library's own calls:
    the oracle builds the listed channels' matrix and the all-channel
    sensitivity itself (full voxel scan, same Philox rays), runs FP64 MLEM
    from the homogeneous start, and computes A^T y in FP64.  Every
    `<it>_<out>` image (tests/reco.cu:283-290) is within 1e-4 relative L2 of
    the oracle's image after it + 1 iterations, with the in-tree sensitivity
    and with a sensitivity file; the backprojection output (normalised to sum 1,
    tests/backprojection.cu:115-161) within 1e-5."""
    import torch
    from epetreco_amd import phantom
    from epetreco_amd.api import OP_N
    h5, reco, bp = _tool("epet_h5"), _tool("reco"), _tool("backprojection")
    ctx = gpu_ctx_factory("small", seed=0)
    s, g = ctx.setup, ctx.grid
    A_full = ctx.sm_build(None, m=s.nchannels)
    xt = torch.as_tensor(phantom.phantom(g, "small", 0).astype(np.float32), device="cuda")
    ybar = A_full.spmv(OP_N, xt).double()
    ybar = (ybar * (1e5 / ybar.sum())).cpu().numpy()
    A_full.close()
    meas = phantom.poisson_counts(ybar, 3)
    meas.tofile(tmp_path / "m.f32")
    _run([h5, "mkmeas", str(tmp_path / "m.h5"), str(tmp_path / "m.f32")])
    # the oracle's problem: all-channel sensitivity and the measured rows
    arp, acol, aval = oracle.sm_rows(s, g, np.arange(s.nchannels, dtype=np.int32), 8, 0)
    sens = oracle.colsum_f64(g.nvoxels, aval, arp, acol).astype(np.float32)
    sens.tofile(tmp_path / "s.f32")
    _run([h5, "mkdens", str(tmp_path / "s.h5"), str(tmp_path / "s.f32"), "small"])
    rows = np.nonzero(meas)[0].astype(np.int32)
    orp, ocol, oval = oracle.sm_rows(s, g, rows, 8, 0)
    y = meas[rows]
    V = g.nvoxels
    ref = [oracle.mlem_f64(len(rows), V, oval, orp, ocol, y, sens, np.full(V, 1.0 / V), it + 1)
           for it in range(3)]
    for sfile in ("in-tree", str(tmp_path / "s.h5")):
        print("reco with sensitivity", sfile, flush=True)
        _run([reco, str(tmp_path / "m.h5"), "out.h5", "8", sfile, "3", "no-guess.h5",
              "--workload=small"], cwd=tmp_path)
        for it in range(3):
            _run([h5, "dump", str(tmp_path / f"{it}_out.h5"), "density", str(tmp_path / "o.f32")])
            got = np.fromfile(tmp_path / "o.f32", np.float32).astype(np.float64)
            err = np.linalg.norm(got - ref[it]) / np.linalg.norm(ref[it])
            assert err < 1e-4, (sfile, it, err)
    _run([bp, str(tmp_path / "m.h5"), str(tmp_path / "bp.h5"), "8", "--workload=small"])
    _run([h5, "dump", str(tmp_path / "bp.h5"), "density", str(tmp_path / "bp.f32")])
    got = np.fromfile(tmp_path / "bp.f32", np.float32).astype(np.float64)
    c = np.bincount(ocol, weights=oval.astype(np.float64) * np.repeat(y, np.diff(orp)),
                    minlength=V)
    c /= c.sum()
    assert np.linalg.norm(got - c) / np.linalg.norm(c) < 1e-5
