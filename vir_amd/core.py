"""Grid helpers mirroring the reference's core/core_funcs.py.

Only what the projector path needs: np_array (core_funcs.py:10-34),
censpace / boundspace (:37-81) and Grid3d (:212-265), with identical dtype
behaviour so plane coordinates match the reference bit for bit.
"""
import numpy as np


def np_array(v, d, dtype=float):
    """Scalar or array_like -> numpy array of d entries (core_funcs.py:10-34)."""
    v = np.array(v, dtype=dtype)
    if v.size == 1:
        v = np.repeat(v, int(d))
    return v


def censpace(n, d=1.0, c=0.0, dtype=float):
    """Sample centres of n cells of width d centred at c (core_funcs.py:37-59)."""
    return np.linspace(-n + 1, n - 1, n, dtype=dtype) * d / 2.0 + c


def boundspace(n, d=1.0, c=0.0, dtype=float):
    """Cell boundaries of n cells of width d centred at c (core_funcs.py:62-81)."""
    return np.linspace(-n, n, n + 1, dtype=dtype) * d / 2.0 + c


class Grid3d:
    """Voxel grid: counts, spacings, origin, centres and plane coordinates
    (core_funcs.py:212-265)."""

    def __init__(self, nPixels=512, dPixels=1.0, origin=0.0, dtype=float):
        self.nPixels = np_array(nPixels, 3, dtype=int)
        self.nX, self.nY, self.nZ = self.nPixels
        self.dPixels = np_array(dPixels, 3, dtype=dtype)
        self.dX, self.dY, self.dZ = self.dPixels
        self.origin = np_array(origin, 3, dtype=dtype)
        self.oX, self.oY, self.oZ = self.origin
        self.X = censpace(self.nX, d=self.dX, c=self.oX, dtype=dtype)
        self.Y = censpace(self.nY, d=self.dY, c=self.oY, dtype=dtype)
        self.Z = censpace(self.nZ, d=self.dZ, c=self.oZ, dtype=dtype)
        self.Xb = boundspace(self.nX, d=self.dX, c=self.oX, dtype=dtype)
        self.Yb = boundspace(self.nY, d=self.dY, c=self.oY, dtype=dtype)
        self.Zb = boundspace(self.nZ, d=self.dZ, c=self.oZ, dtype=dtype)
