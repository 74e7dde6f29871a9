"""parafem_amd: ParaFEM's EBE-PCG solve and element assembly on MI355X.

Modules:
  api       -- ctypes binding of libparafem_amd.so (the C ABI in include/)
  mesh      -- structured meshes, boundary conditions, loads, deck I/O
  problems  -- workload builders and the p121/p123 flows on a device context
  build     -- compiles csrc/*.hip for gfx950 into libparafem_amd.so
"""
from . import mesh  # noqa: F401
from .api import (Batch, Context, DeviceArray, LibraryNotBuilt, PFError, lib,  # noqa: F401
                  reduce_grid)

__version__ = "0.1.0"
