"""ORACLE: stored-matrix projections -- test infrastructure only.

* fp64 / bp64: textbook sparse products in float64 over a CSR system matrix
  (the mathematics of projection/proj.py:56-150 and projection_c.c:101-219).
* RefProjection: ctypes binding of the reference's own projection_c.c,
  compiled unmodified from /root/reference by oracle/Makefile into
  oracle/_ref/libprojection_ref.so (forward_proj_ravel_rays_struct,
  back_proj_ravel_rays_struct; projection_c.c:61-65,150-157,200-208).
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
REF_SO = os.path.join(_HERE, "_ref", "libprojection_ref.so")


def fp64(vol, ind, ln, off):
    """sino[r] = sum_e vol[ind[e]] * len[e] in float64."""
    v = np.asarray(vol, np.float64).reshape(-1)
    prod = v[np.asarray(ind, np.int64)] * np.asarray(ln, np.float64)
    out = np.zeros(len(off) - 1, np.float64)
    starts = np.asarray(off[:-1], np.int64)
    ends = np.asarray(off[1:], np.int64)
    nz = ends > starts
    if nz.any():
        # rows are contiguous; with the empty rows removed each reduceat
        # segment [start_i, start_{i+1}) is exactly one row
        out[nz] = np.add.reduceat(prod, starts[nz])
    return out


def bp64(sino, ind, ln, off, nvox):
    """vol[v] = sum_{e: ind[e]=v} len[e] * sino[row(e)] in float64."""
    off = np.asarray(off, np.int64)
    rows = np.repeat(np.arange(len(off) - 1), np.diff(off))
    w = np.asarray(ln, np.float64) * np.asarray(sino, np.float64).reshape(-1)[rows]
    return np.bincount(np.asarray(ind, np.int64), weights=w, minlength=nvox)


class CsrOperator:
    """A stored CSR matrix (int32 index, float32 length, int64 offsets) as a
    float64 operator with `A @ x`, `A.T @ y` and `A.shape`, the products
    computed by the C oracle on `nthreads` OpenMP threads
    (oracle_csr_fp64 / oracle_csr_bp64).  Lets oracle/sirt_ref.sirt run at
    workload size.  Pinned against scipy.sparse in tests/test_oracle_siddon.py."""

    def __init__(self, ind, ln, off, ncols, nthreads=None, _t=False):
        from . import siddon as osd
        self.ind = np.ascontiguousarray(ind, np.int32)
        self.ln = np.ascontiguousarray(ln, np.float32)
        self.off = np.ascontiguousarray(off, np.int64)
        self.nrows = self.off.size - 1
        self.ncols = int(ncols)
        self.nthreads = int(nthreads or max(1, min(16, len(os.sched_getaffinity(0)))))
        self._t = _t
        self._lib = osd._load()
        P = ctypes.c_void_p
        self._lib.oracle_csr_fp64.restype = None
        self._lib.oracle_csr_fp64.argtypes = [ctypes.c_int64, P, P, P, P, P, ctypes.c_int]
        self._lib.oracle_csr_bp64.restype = ctypes.c_int
        self._lib.oracle_csr_bp64.argtypes = [ctypes.c_int64, P, P, P, ctypes.c_int64, P, P, ctypes.c_int]

    @property
    def shape(self):
        return (self.ncols, self.nrows) if self._t else (self.nrows, self.ncols)

    @property
    def T(self):
        t = CsrOperator.__new__(CsrOperator)
        t.__dict__.update(self.__dict__)
        t._t = not self._t
        return t

    def __matmul__(self, v):
        v = np.ascontiguousarray(np.asarray(v, np.float64).reshape(-1))
        p = lambda a: a.ctypes.data_as(ctypes.c_void_p)  # noqa: E731
        if not self._t:
            out = np.empty(self.nrows)
            self._lib.oracle_csr_fp64(self.nrows, p(self.ind), p(self.ln), p(self.off), p(v), p(out),
                                      self.nthreads)
        else:
            out = np.empty(self.ncols)
            if self._lib.oracle_csr_bp64(self.nrows, p(self.ind), p(self.ln), p(self.off), self.ncols,
                                         p(v), p(out), self.nthreads):
                raise MemoryError("oracle_csr_bp64")
        return out


def sd_b_proj_numpy(sino, ind, ln, off, nPixels):
    """The reference's numpy backprojector (projection/proj.py:119-150,
    ravel=True branch :133-136) over a CSR matrix: float32 volume, rays in
    order, `phantom_flat[idx] += len * sino[r]` with numpy's buffered
    fancy-index `+=`, so a voxel repeated within a ray keeps one update."""
    phantom = np.zeros(tuple(nPixels), dtype=np.float32)
    flat = phantom.ravel()
    s = np.asarray(sino, np.float32).ravel()
    ind = np.asarray(ind)
    ln = np.asarray(ln, np.float32)
    for r in range(len(off) - 1):
        a, b = int(off[r]), int(off[r + 1])
        if b > a:
            flat[ind[a:b]] += ln[a:b] * s[r]
    return phantom


class _RaveledRay(ctypes.Structure):
    # projection_c.c:61-65
    _fields_ = [("n", ctypes.c_int), ("R", ctypes.POINTER(ctypes.c_int)),
                ("L", ctypes.POINTER(ctypes.c_float))]


class RefProjection:
    """The reference's compiled projection_c.c (built into oracle/_ref)."""

    def __init__(self):
        if not os.path.exists(REF_SO):
            raise FileNotFoundError(REF_SO)
        self.so = ctypes.CDLL(REF_SO)
        for name in ("forward_proj_ravel_rays_struct", "back_proj_ravel_rays_struct"):
            fn = getattr(self.so, name)
            fn.restype = None
            fn.argtypes = [ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float),
                           ctypes.POINTER(_RaveledRay), ctypes.c_int]

    @staticmethod
    def _rays(ind, ln, off):
        ind = np.ascontiguousarray(ind, np.int32)
        ln = np.ascontiguousarray(ln, np.float32)
        n = len(off) - 1
        rays = (_RaveledRay * n)()
        ip = ind.ctypes.data
        lp = ln.ctypes.data
        for r in range(n):
            s = int(off[r])
            rays[r].n = int(off[r + 1] - off[r])
            rays[r].R = ctypes.cast(ip + 4 * s, ctypes.POINTER(ctypes.c_int))
            rays[r].L = ctypes.cast(lp + 4 * s, ctypes.POINTER(ctypes.c_float))
        return rays, ind, ln

    def fp(self, vol, ind, ln, off):
        vol = np.ascontiguousarray(vol, np.float32).reshape(-1)
        sino = np.zeros(len(off) - 1, np.float32)
        rays, ind, ln = self._rays(ind, ln, off)
        self.so.forward_proj_ravel_rays_struct(
            vol.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
            sino.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), rays, len(off) - 1)
        return sino

    def bp(self, sino, ind, ln, off, nvox):
        vol = np.zeros(nvox, np.float32)
        sino = np.ascontiguousarray(sino, np.float32).reshape(-1)
        rays, ind, ln = self._rays(ind, ln, off)
        self.so.back_proj_ravel_rays_struct(
            vol.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
            sino.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), rays, len(off) - 1)
        return vol
