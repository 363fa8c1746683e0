"""ORACLE: the reference's own sys_mat/siddons.c, compiled unmodified by
oracle/Makefile into oracle/_ref/libsiddons_ref.so -- test infrastructure
only.  ctypes calls of its calc_grid_lines (:137-148), siddons (:290-341),
siddons_proj (:345-396) and forward_project (:403-460), one ray per call as
the C API is.  Its debug printf output is sent to /dev/null.

siddons.c reads outside its grid-line arrays on some rays: its first plane
index floor(N - (plane[N] - entry)/d) is -1 when rounding puts a ray's entry
a hair before plane 0, and its range ignores the z slab.  The grid lines
are therefore held between guard values, every call runs twice with two
different guards, and a result that changes with the guard is reported as
not in bounds (undefined in siddons.c).  Scratch arrays are sized with room
to spare (the reference's own callers size x_alpha at N+1 values although
calc_alphas writes up to N+2).  forward_project allocates its own arrays:
callers use it only on rays the per-ray calls report in bounds, and with
nz >= nx/2 + ny/2 + 1 (its merged list buffer holds 2 nz + 3 values).
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
        # grid lines at [1, N+2) of buffers with one guard value on each side
        self._guarded = [np.zeros(int(v) + 3, np.float32) for v in n]
        self.lines = [b[1:-1] for b in self._guarded]
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

    def _guard(self, value):
        for b in self._guarded:
            b[0] = b[-1] = value

    def _twice(self, call):
        """call() with both guards; (result, in_bounds)."""
        self._guard(np.float32(-3.0e4))
        a = call()
        self._guard(np.float32(7.5e3))
        b = call()
        same = all(np.array_equal(np.asarray(x).view(np.int32), np.asarray(y).view(np.int32))
                   for x, y in zip(a if isinstance(a, tuple) else (a,), b if isinstance(b, tuple) else (b,)))
        return a, same

    def ray(self, src, trg):
        """((idx int32, length float32), in_bounds) of one ray; rows empty
        when siddons() leaves the row buffer untouched (a ray that misses)."""
        return self._twice(lambda: self._ray(src, trg))

    def proj(self, iArray, src, trg):
        """(ray sum float32, in_bounds) of one ray; siddons_proj is called
        only when the ray's rows are in bounds and every index lies inside
        the volume (its reads are then defined), else (None, False)."""
        (idx, _), ok = self.ray(src, trg)
        nvox = self.shape[0] * self.shape[1] * self.shape[2]
        if not ok or (idx.size and (idx.min() < 0 or idx.max() >= nvox)):
            return None, False
        return self._proj(iArray, src, trg), True

    def _ray(self, src, trg):
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

    def _proj(self, iArray, src, trg):
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
