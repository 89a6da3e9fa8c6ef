"""gpvmc_amd: MI355X-native Metropolis VMC hot path of EPFL-LQM/gpvmc.

The compute path is libgpvmc_hip.so (hand-written gfx950 kernels behind the C ABI in
include/gpvmc.h).  This package only binds that library; there is no CPU fallback:
importing the binding without the built library raises.
"""
from .binding import (  # noqa: F401
    ACC_STRIDE, FP64, MIXED, MEAS_OVERLAP, MEAS_PROJHEIS, MEAS_STAGMAGNZ, MEAS_STATSPIN,
    Batch, GpvmcError, library_path, load_library,
)
from .model import (  # noqa: F401
    edges, ground_occ, neighbours, orbital_table, philox_uniforms, random_spins,
)
