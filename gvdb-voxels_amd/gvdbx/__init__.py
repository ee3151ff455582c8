"""gvdbx -- MI355X (gfx950) point-to-sparse-grid path of GVDB Voxels.

Python mirror of the reference VolumeGVDB point API over the C ABI in include/gvdbx.h
(libgvdbx.so, hand-written HIP).  The product path is the HIP library: there is no CPU
fallback, and loading fails loudly when libgvdbx.so is missing.
"""
from .lib import GvdbxError, load_library, LIB_PATH, NODE_DTYPE, DeviceArray  # noqa: F401
from .volume import VolumeGVDB, T_FLOAT, T_UCHAR4  # noqa: F401
