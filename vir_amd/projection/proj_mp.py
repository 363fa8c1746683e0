"""The multiprocessing projection calls (projection/proj_mp.py:43-96) on the
GPU: the work the reference spreads over a pool of CPU processes sharing
the volume runs as one call of the device projectors instead.

  mp_sd_f_proj(phantom, sdlist, cpus=...)         proj_mp.py:43-62
  mp_sd_b_proj(sino, sdlist, nPixels, cpus=...)   proj_mp.py:78-96

sdlist is a siddons() object array ([x, y, z, len] per ray, None for misses,
as the workers index phantom[(ray[0], ray[1], ray[2])]; [ravel_idx, len]
rays are accepted too).  `cpus` is accepted for the call surface and not
used.

Return values follow the reference:
  * mp_sd_f_proj: the per-ray sums, shaped like sdlist, float32 -- or
    float64 when some ray is None, because the worker returns the Python
    float 0.0 for it and np.array() of the mixed list widens
    (proj_mp.py:32-40, 61);
  * mp_sd_b_proj: a float32 volume of shape nPixels; each worker applies
    phantom[(x, y, z)] += len * value under the lock (proj_mp.py:65-75),
    numpy's fancy-index `+=`, so a voxel repeated within a ray receives its
    last entry only -- sd_b_proj(compat=True) -- with the rays added in
    order in float32 (the reference's order depends on the pool's
    scheduling; its float32 rounding differs in the last bits).
The forward sums are accumulated in float64 where the workers use numpy's
float32 np.sum (results agree to float32 rounding).
"""
import multiprocessing as mp

import numpy as np

from .proj import sd_b_proj, sd_f_proj


def _rays(sdlist):
    return np.asarray(sdlist, dtype=object)


def _is_ravel(sd):
    first = next((r for r in sd.reshape(-1) if r is not None), None)
    return first is not None and len(first) == 2


def mp_sd_f_proj(phantom, sdlist, cpus=mp.cpu_count() - 1):
    sd = _rays(sdlist)
    sino = sd_f_proj(np.asarray(phantom), sd, ravel=_is_ravel(sd), flat=False, sino_shape=sd.shape)
    if any(r is None for r in sd.reshape(-1)):
        return np.asarray(sino, np.float64)
    return np.asarray(sino, np.float32)


def mp_sd_b_proj(sino, sdlist, nPixels, cpus=mp.cpu_count() - 1):
    sd = _rays(sdlist)
    out = sd_b_proj(np.asarray(sino), sd, nPixels, ravel=_is_ravel(sd), sequential=True, compat=True)
    return np.asarray(out, np.float32).reshape(tuple(int(v) for v in np.atleast_1d(nPixels)))
