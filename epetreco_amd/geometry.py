"""Scanner and voxel-grid descriptions, mirroring the reference's value types.

``MeasurementSetup`` and ``VoxelGrid`` take their constructor arguments in the
order of ``DefaultMeasurementSetup<T>(pos0x, pos1x, na, n0z, n0y, n1z, n1y, da,
segx, segy, segz)`` (MeasurementSetup.hpp:157-163) and ``DefaultVoxelGrid<T>(ox,
oy, oz, dx, dy, dz, nx, ny, nz)`` (VoxelGrid.hpp:103-110).  Units are metres.
Their ``ctypes`` images have the reference's member layout
(MeasurementSetup.hpp:129-139: pos0x, pos1x, na, da, n0z, n0y, n1z, n1y, segx,
segy, segz; VoxelGrid.hpp:88-91: O[3], D[3], N[3]), which is also the layout
``include/epet.h`` declares.

The index maps follow MeasurementSetupLinIndex.hpp:32-165 (channel id =
(((ida*n0z + id0z)*n0y + id0y)*n1z + id1z)*n1y + id1y) and
VoxelGridLinIndex.hpp:193-282 (voxel id = (idx*ny + idy)*nz + idz, z fastest).
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np


class CSetup(ctypes.Structure):
    _fields_ = [("pos0x", ctypes.c_float), ("pos1x", ctypes.c_float),
                ("na", ctypes.c_int), ("da", ctypes.c_float),
                ("n0z", ctypes.c_int), ("n0y", ctypes.c_int),
                ("n1z", ctypes.c_int), ("n1y", ctypes.c_int),
                ("segx", ctypes.c_float), ("segy", ctypes.c_float),
                ("segz", ctypes.c_float)]


class CGrid(ctypes.Structure):
    _fields_ = [("o", ctypes.c_float * 3), ("d", ctypes.c_float * 3),
                ("n", ctypes.c_int * 3)]


@dataclass(frozen=True)
class MeasurementSetup:
    pos0x: float
    pos1x: float
    na: int
    n0z: int
    n0y: int
    n1z: int
    n1y: int
    da: float
    segx: float
    segy: float
    segz: float

    @property
    def nchannels(self) -> int:
        return self.na * self.n0z * self.n0y * self.n1z * self.n1y

    def c(self) -> CSetup:
        return CSetup(self.pos0x, self.pos1x, self.na, self.da, self.n0z,
                      self.n0y, self.n1z, self.n1y, self.segx, self.segy,
                      self.segz)

    # MeasurementSetupLinIndex.hpp:40-44
    def lin_id(self, id0z, id0y, id1z, id1y, ida):
        return ((((np.asarray(ida) * self.n0z + id0z) * self.n0y + id0y)
                 * self.n1z + id1z) * self.n1y + id1y)

    # MeasurementSetupLinIndex.hpp:65-163
    def decompose(self, lin):
        lin = np.asarray(lin)
        p1 = self.n1y
        p2 = p1 * self.n1z
        p3 = p2 * self.n0y
        p4 = p3 * self.n0z
        t = lin % p4
        return (lin // p4, t // p3, (t % p3) // p2, (t % p2) // p1, t % p1)


@dataclass(frozen=True)
class VoxelGrid:
    ox: float
    oy: float
    oz: float
    dx: float
    dy: float
    dz: float
    nx: int
    ny: int
    nz: int

    @property
    def nvoxels(self) -> int:
        return self.nx * self.ny * self.nz

    def c(self) -> CGrid:
        g = CGrid()
        g.o[:] = (self.ox, self.oy, self.oz)
        g.d[:] = (self.dx, self.dy, self.dz)
        g.n[:] = (self.nx, self.ny, self.nz)
        return g

    # VoxelGridLinIndex.hpp:201-203
    def lin_id(self, idx, idy, idz):
        return (np.asarray(idx) * self.ny + idy) * self.nz + idz

    # VoxelGridLinIndex.hpp:227-280
    def decompose(self, lin):
        lin = np.asarray(lin)
        return (lin // (self.ny * self.nz), (lin % (self.ny * self.nz)) // self.nz,
                lin % self.nz)

    def centres(self):
        """Voxel centres (V, 3) in voxel-id order, float64."""
        ix, iy, iz = self.decompose(np.arange(self.nvoxels))
        return np.stack([self.ox + (ix + 0.5) * self.dx,
                         self.oy + (iy + 0.5) * self.dy,
                         self.oz + (iz + 0.5) * self.dz], axis=1)


# Two heads 460 mm apart (segment centres at x = -/+0.230 m), segments of
# 20 x 4 x 4 mm (x depth, y, z), rotating about y.
def setup_small() -> MeasurementSetup:
    """5x5 segments per head, 18 angles of 10 deg: 11,250 channels."""
    return MeasurementSetup(-0.230, 0.230, 18, 5, 5, 5, 5, 10.0, 0.020, 0.004, 0.004)


def setup_full() -> MeasurementSetup:
    """13x13 segments per head, 180 angles of 2 deg: 5,140,980 channels."""
    return MeasurementSetup(-0.230, 0.230, 180, 13, 13, 13, 13, 2.0, 0.020, 0.004, 0.004)


def grid_16_3mm() -> VoxelGrid:
    return VoxelGrid(-0.024, -0.024, -0.024, 0.003, 0.003, 0.003, 16, 16, 16)


def grid_52_1mm() -> VoxelGrid:
    """GRID52 of the reference (voxelgrid_defines.h:28-39)."""
    return VoxelGrid(-0.026, -0.026, -0.026, 0.001, 0.001, 0.001, 52, 52, 52)


def grid_104_half_mm() -> VoxelGrid:
    return VoxelGrid(-0.026, -0.026, -0.026, 0.0005, 0.0005, 0.0005, 104, 104, 104)


# name -> (setup, grid, rays per channel)
WORKLOADS = {
    "small": (setup_small, grid_16_3mm, 8),          # plain MLEM, small
    "full": (setup_full, grid_52_1mm, 64),           # plain MLEM, 52^3 / 5.14M ch
    "projector": (setup_small, grid_52_1mm, 8),      # differentiable projector
    "large": (setup_full, grid_104_half_mm, 32),     # reduced-precision, 104^3
}


def workload(name: str):
    s, g, nrays = WORKLOADS[name]
    return s(), g(), nrays
