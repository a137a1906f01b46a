"""MI355X-native SU(2) DMRG inner loop (Hubbard / t-J with flux).

The compute lives in libhhdmrg.so (HIP kernels for gfx950 behind the C ABI of
include/hhdmrg.h); this package is its host-side mirror of the reference
subroutine interface.
"""
from ._lib import LIB_PATH, load  # noqa: F401
from .dmrg import (Batch, Block, Context, Driver, HHError, StepRecord, Superblock, Trunc,  # noqa: F401
                   structure_factor_cut)
from .model import WORKLOADS, Model, Workload  # noqa: F401
