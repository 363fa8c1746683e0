"""vir_amd: MI355X-native (gfx950) projector and SIRT hot path of vir.

Drop-in modules mirroring the reference's call surface:
  vir_amd.sys_mat.siddon   -- siddons(), SiddonProjector
  vir_amd.projection.proj  -- sd_f_proj, sd_b_proj (stored-matrix operators)
  vir_amd.proj_geom        -- geom_circular, geom_square
  vir_amd.core             -- Grid3d, censpace, boundspace, np_array
Compute runs in libvir_amd.so (hand-written HIP for gfx950) through ctypes.
"""
from .core import Grid3d, boundspace, censpace, np_array  # noqa: F401

__version__ = "0.1.0"
