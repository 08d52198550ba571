"""MI355X-native MLEM reconstruction hot path of epetreco.

Host surface over libepet.so (hand-written HIP for gfx950, C-ABI in
include/epet.h).  Importing the package does not touch the GPU.
"""
from .geometry import MeasurementSetup, VoxelGrid, workload  # noqa: F401


def __getattr__(name):
    # the device API loads torch and libepet.so on first use
    if name in ("Context", "Matrix", "EpetError", "OP_N", "OP_T", "VAL_F32", "VAL_BF16",
                "VAL_F16"):
        from . import api
        return getattr(api, name)
    raise AttributeError(name)
