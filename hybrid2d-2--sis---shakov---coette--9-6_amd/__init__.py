"""MI355X-native GSIS kinetic loop (discrete-velocity Shakhov solve + synthetic
NS correction) behind the C-ABI of include/gsis_capi.h.

Submodules:
  capi    -- ctypes mirror of the C-ABI and the libgsis.so loader
  cases   -- host-side producers: mesh geometry, velocity grids, constants,
             initial state, readers of the reference's input files
  solver  -- GsisSolver: the driver loop over the C-ABI
  multiblock -- meshes of several blocks, their links and a driver with one
             handle per block
  restart -- restart files in the reference's layout
  output  -- the reference's Tecplot files from downloaded fields
  fields  -- macroscopic fields and their relative L2 differences
  build   -- hipcc build of libgsis.so for gfx950 (and the flang host side)
"""
from . import capi, cases, fields, multiblock, output, restart  # noqa: F401

__all__ = ["capi", "cases", "fields"]
