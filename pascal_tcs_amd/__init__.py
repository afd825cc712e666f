"""MI355X-native PaScaL_TCS MPM-STD time step (HIP/gfx950) behind a C-ABI.

Modules:
  _lib    ctypes binding of libpascal_tcs.so (include/pascal_tcs.h)
  config  mirror of the reference's module global (namelist + constants)
  ic      seeded initial states and the restart-file format
  solver  TCSSolver: the reference's module routines on one GPU
  ptdma   the PaScaL_TDMA many-system API on one GPU
  build   in-tree hipcc build of the library
"""
from .config import Properties, TCSConfig  # noqa: F401

__all__ = ["Properties", "TCSConfig"]
