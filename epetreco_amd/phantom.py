"""Seeded synthetic phantoms and measurements for the benchmark workloads.

A measurement is generated the way the reference's inputs are used
(tests/reco.cu:76-183): Poisson counts of the forward projection of a phantom
over every channel, the list of nonzero channels (readHDF5_MeasVct drops zero
channels, wrappers.hpp:335-349), and a sensitivity image s = A^T 1 over all
channels (what tests/backprojection.cu computes from an all-ones measurement;
here the column sums are accumulated in FP64).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import torch

from .geometry import VoxelGrid


@dataclass
class Ellipsoid:
    centre: tuple   # metres
    semi: tuple     # metres (axis-aligned)
    activity: float


def ellipsoids(name: str, seed: int):
    """The ellipsoid set of a workload (semi-axes and activities as specified)."""
    rng = np.random.default_rng(seed)
    if name in ("small", "projector"):
        n, smin, smax, amin, amax, cmax = 3, 0.005, 0.020, 0.2, 1.0, 0.010
    else:
        n, smin, smax, amin, amax, cmax = 8, 0.002, 0.015, 0.1, 1.0, 0.012
    out = []
    for _ in range(n):
        semi = tuple(rng.uniform(smin, smax, 3))
        centre = tuple(rng.uniform(-cmax, cmax, 3))
        out.append(Ellipsoid(centre, semi, float(rng.uniform(amin, amax))))
    return out


def voxelise(grid: VoxelGrid, ells, activities=None, background=0.0, cyl_radius=None):
    """Activity at voxel centres (voxel-id order, z fastest), float64."""
    c = grid.centres()
    x = np.zeros(grid.nvoxels)
    if background and cyl_radius is not None:
        # cylinder about the rotation axis (y)
        x[(c[:, 0] ** 2 + c[:, 2] ** 2) <= cyl_radius ** 2] = background
    for k, e in enumerate(ells):
        a = e.activity if activities is None else activities[k]
        inside = np.sum(((c - np.asarray(e.centre)) / np.asarray(e.semi)) ** 2, axis=1) <= 1.0
        x[inside] += a
    return x


def phantom(grid: VoxelGrid, name: str, seed: int = 0) -> np.ndarray:
    ells = ellipsoids(name, seed)
    if name in ("small", "projector"):
        return voxelise(grid, ells)
    return voxelise(grid, ells, background=0.05, cyl_radius=0.024)


def dynamic_series(grid: VoxelGrid, name: str, frames: int, seed: int = 0,
                   counts: tuple = (1e5, 1e7)):
    """Frames t = 0..F-1 with per-ellipsoid activity a * exp(-t / tau), tau ~ U(5, 50) s,
    and total counts log-uniform in `counts`.  Returns (images (F, V), totals (F,))."""
    ells = ellipsoids(name, seed)
    rng = np.random.default_rng(seed + 1)
    tau = rng.uniform(5.0, 50.0, len(ells))
    imgs = []
    for t in range(frames):
        acts = [e.activity * np.exp(-t / tau[k]) for k, e in enumerate(ells)]
        imgs.append(voxelise(grid, ells, activities=acts))
    totals = np.exp(rng.uniform(np.log(counts[0]), np.log(counts[1]), frames))
    return np.stack(imgs), totals


TOTAL_COUNTS = {"small": 1e5, "projector": 1e5, "full": 1e7, "large": 1e7}


@dataclass
class Problem:
    A: object            # Matrix restricted to the measured channels (ECSR)
    y: torch.Tensor      # counts of the measured channels (float32, device)
    s: torch.Tensor      # sensitivity over all channels (float32, device)
    rows: np.ndarray     # measured channel ids
    x_true: np.ndarray   # phantom (float64)


def poisson_counts(ybar: np.ndarray, seed: int) -> np.ndarray:
    return np.random.default_rng(seed).poisson(ybar).astype(np.float32)


def make_problem(ctx, A_full, name: str, seed: int = 0, total: float = None) -> Problem:
    """Synthetic measurement of a workload's phantom through A_full (all channels)."""
    from .api import OP_N
    grid = ctx.grid
    x_true = phantom(grid, name, seed)
    xt = torch.as_tensor(x_true.astype(np.float32), device=ctx.device)
    ybar = A_full.spmv(OP_N, xt).double()
    total = TOTAL_COUNTS.get(name, 1e5) if total is None else total
    ybar = (ybar * (total / ybar.sum())).cpu().numpy()
    counts = poisson_counts(ybar, seed + 17)
    rows = np.nonzero(counts)[0].astype(np.int32)
    A = A_full.select_rows(rows)
    y = torch.as_tensor(counts[rows], device=ctx.device)
    s = A_full.colsum_f64().float()
    return Problem(A, y, s, rows, x_true)
