"""ORACLE: the reference's own sys_mat/siddons.c, compiled unmodified by
oracle/Makefile into oracle/_ref/libsiddons_ref.so -- test infrastructure
only.  ctypes calls of its calc_grid_lines (:137-148), siddons (:290-341),
siddons_proj (:345-396) and forward_project (:403-460), one ray per call as
the C API is.  Its debug printf output is sent to /dev/null.

The scratch arrays are sized with room to spare (the reference's own
callers size x_alpha at N+1 values although calc_alphas writes up to N+2),
so calls stay inside their buffers; callers must keep the rays where
siddons.c itself stays inside its plane arrays (rays whose plane index
range lies in [0, N] on every axis) and, for forward_project, nz >= nx/2 +
ny/2 + 1 (its merged list buffer holds 2 nz + 3 values).
"""
import contextlib
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
REF_SO = os.path.join(_HERE, "_ref", "libsiddons_ref.so")


class VectI(ctypes.Structure):
    _fields_ = [("x", ctypes.c_int), ("y", ctypes.c_int), ("z", ctypes.c_int)]


class VectF(ctypes.Structure):
    _fields_ = [("x", ctypes.c_float), ("y", ctypes.c_float), ("z", ctypes.c_float)]


class VoxelLength(ctypes.Structure):
    _fields_ = [("idx", ctypes.c_int), ("length", ctypes.c_float)]


FP = ctypes.POINTER(ctypes.c_float)


@contextlib.contextmanager
def _quiet():
    libc = ctypes.CDLL(None)
    libc.fflush(None)
    saved = os.dup(1)
    null = os.open(os.devnull, os.O_WRONLY)
    try:
        os.dup2(null, 1)
        yield
    finally:
        libc.fflush(None)
        os.dup2(saved, 1)
        os.close(null)
        os.close(saved)


class SiddonsC:
    def __init__(self, nPixels, dPixels):
        if not os.path.exists(REF_SO):
            raise FileNotFoundError(REF_SO)
        self.so = ctypes.CDLL(REF_SO)
        n = np.broadcast_to(np.asarray(nPixels, np.int64), (3,))
        d = np.broadcast_to(np.asarray(dPixels, np.float32), (3,))
        self.n = VectI(*(int(v) for v in n))
        self.d = VectF(*(float(v) for v in d))
        self.shape = tuple(int(v) for v in n)
        self.lines = [np.zeros(int(v) + 1, np.float32) for v in n]
        self.so.calc_grid_lines.restype = None
        self.so.calc_grid_lines(*(a.ctypes.data_as(FP) for a in self.lines), self.n, self.d)
        m = int(n.sum()) + 8
        self.alpha = [np.zeros(int(v) + 3, np.float32) for v in n]
        self.all_alpha = np.zeros(m, np.float32)
        self.rows = (VoxelLength * m)()
        self.so.siddons.restype = None
        self.so.siddons_proj.restype = None
        self.so.forward_project.restype = None

    def _bufs(self):
        return ([a.ctypes.data_as(FP) for a in self.lines] + [a.ctypes.data_as(FP) for a in self.alpha] +
                [self.all_alpha.ctypes.data_as(FP)])

    def ray(self, src, trg):
        """The rows of one ray: (idx int32, length float32); empty when
        siddons() leaves the row buffer untouched (a ray that misses)."""
        for k in range(len(self.rows)):
            self.rows[k].idx, self.rows[k].length = -7, float("nan")
        with _quiet():
            self.so.siddons(VectF(*map(float, src)), VectF(*map(float, trg)), self.n, self.d, *self._bufs(),
                            self.rows)
        if self.rows[0].idx == -7:
            return np.zeros(0, np.int32), np.zeros(0, np.float32)
        k = 0
        while self.rows[k].idx != -1:
            k += 1
        idx = np.array([self.rows[i].idx for i in range(k)], np.int32)
        ln = np.array([self.rows[i].length for i in range(k)], np.float32)
        return idx, ln

    def proj(self, iArray, src, trg):
        vol = np.ascontiguousarray(np.asarray(iArray, np.float32).reshape(-1))
        out = ctypes.c_float(0.0)
        with _quiet():
            self.so.siddons_proj(ctypes.byref(out), vol.ctypes.data_as(FP), VectF(*map(float, src)),
                                 VectF(*map(float, trg)), self.n, self.d, *self._bufs())
        return np.float32(out.value)

    def forward_project(self, iArray):
        nx, ny, nz = self.shape
        if 2 * nz + 3 < nx + ny + 5:
            raise ValueError("forward_project's alpha buffer (2 nz + 3) is too small for this grid")
        vol = np.ascontiguousarray(np.asarray(iArray, np.float32).reshape(-1))
        out = np.zeros(64, np.float32)
        with _quiet():
            self.so.forward_project(vol.ctypes.data_as(FP), out.ctypes.data_as(FP), self.n, self.d)
        return out
