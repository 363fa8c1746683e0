"""ORACLE: Siddon system matrix, CPU restatement -- test infrastructure only.

Python face of oracle/siddon_oracle.c, a plain-C restatement of the
reference's ``siddons()`` (sys_mat/siddon.py:636-833).  Output formats mirror
the reference exactly:

* ``flat=True``  -> ``(flat_ind int32[nnz], flat_len float32[nnz],
  flat_off int32[nRays+1])`` (siddon.py:698-704, 813-831)
* ``ravel=True`` -> object array of ``[ravel_idx int32, len float32]`` or
  ``None`` for rays that miss (siddon.py:818-822)
* otherwise      -> object array of ``[x int32, y int32, z int32, len f32]``
  (siddon.py:823-824)

Pins (tests/test_oracle_siddon.py): golden vectors captured from the
reference itself (tests/golden/make_siddon_golden.py), FP64 brute-force
Liang-Barsky voxel sets on small grids, and chord-length / count invariants.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_lib = None


def load_library(path):
    """Open an oracle build (the default one, or a planted-mistake variant)."""
    if not os.path.exists(path):
        raise RuntimeError(f"{path} missing: run `make -C oracle`")
    lib = ctypes.CDLL(path)
    P = ctypes.c_void_p
    if hasattr(lib, "oracle_fp_bp"):
        lib.oracle_fp_bp.restype = ctypes.c_int64
        lib.oracle_fp_bp.argtypes = [ctypes.c_int64, P, P, P, P, P, ctypes.c_double, P, P, P,
                                     ctypes.c_int]
    if True:
        lib.oracle_siddons.restype = ctypes.c_int64
        lib.oracle_siddons.argtypes = [ctypes.c_int64, P, P, P, P, P,
                                       ctypes.c_double, P, P, P, P, P, P,
                                       ctypes.c_int64]
        lib.oracle_planes.restype = None
        lib.oracle_planes.argtypes = [ctypes.c_int64, ctypes.c_float,
                                      ctypes.c_float, P]
    return lib


def _load():
    global _lib
    if _lib is None:
        _lib = load_library(os.path.join(_HERE, "liboracle.so"))
    return _lib


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def grid_params(nPixels, dPixels, origin):
    """Grid3d(nPixels, dPixels, origin, float32) parameters
    (core/core_funcs.py:235-265 via np_array core_funcs.py:10-34)."""
    n = np.array(nPixels, dtype=np.int64)
    if n.size == 1:
        n = np.repeat(n, 3)
    d = np.array(dPixels, dtype=np.float32)
    if d.size == 1:
        d = np.repeat(d, 3)
    o = np.array(origin, dtype=np.float32)
    if o.size == 1:
        o = np.repeat(o, 3)
    return n.reshape(3), d.reshape(3), o.reshape(3)


def planes(n, d, o):
    lib = _load()
    out = np.empty(int(n) + 1, dtype=np.float32)
    lib.oracle_planes(int(n), ctypes.c_float(d), ctypes.c_float(o), _ptr(out))
    return out


def trace(src, trg, nPixels=128, dPixels=1.0, origin=0.0, epsilon=1e-5, lib=None):
    """Raw trace: returns (shape, x, y, z, len, off int64, flags uint8, n)."""
    lib = lib or _load()
    n, d, o = grid_params(nPixels, dPixels, origin)
    trg = np.array(trg, dtype=np.float32)
    if trg.ndim == 1:
        trg = trg[np.newaxis, :]
    src = np.array(src, dtype=np.float32)
    if src.ndim == 1:
        src = src[np.newaxis, :]
    src = np.ascontiguousarray(np.broadcast_to(src, trg.shape), dtype=np.float32)
    trg = np.ascontiguousarray(trg)
    shape = trg.shape[:-1]
    nr = int(np.prod(shape))
    per = int(n.sum() + 4)
    cap = max(1, nr * per)
    x = np.empty(cap, np.int32)
    y = np.empty(cap, np.int32)
    z = np.empty(cap, np.int32)
    ln = np.empty(cap, np.float32)
    off = np.zeros(nr + 1, np.int64)
    bad = np.zeros(max(nr, 1), np.uint8)
    tot = lib.oracle_siddons(nr, _ptr(src), _ptr(trg), _ptr(n), _ptr(d), _ptr(o),
                             float(epsilon), _ptr(x), _ptr(y), _ptr(z), _ptr(ln),
                             _ptr(off), _ptr(bad), cap)
    if tot < 0:
        raise RuntimeError("oracle capacity exceeded")
    return shape, x[:tot], y[:tot], z[:tot], ln[:tot], off, bad[:nr], n


def fp_bp_threads(src, trg, vol, nPixels=128, dPixels=1.0, origin=0.0, epsilon=1e-5,
                  nthreads=1):
    """The reference's CPU path on `nthreads` OpenMP threads: trace, float64
    CSR forward projection, float32 backprojection of that sinogram.
    Returns (sino float32 (nrays,), back float32 nPixels, nnz)."""
    lib = _load()
    n, d, o = grid_params(nPixels, dPixels, origin)
    src = np.ascontiguousarray(np.asarray(src, np.float32).reshape(-1, 3))
    trg = np.ascontiguousarray(np.asarray(trg, np.float32).reshape(-1, 3))
    vol = np.ascontiguousarray(np.asarray(vol, np.float32).reshape(tuple(int(v) for v in n)))
    sino = np.zeros(src.shape[0], np.float32)
    back = np.zeros(vol.shape, np.float32)
    nnz = lib.oracle_fp_bp(src.shape[0], _ptr(src), _ptr(trg), _ptr(n), _ptr(d), _ptr(o),
                           float(epsilon), _ptr(vol), _ptr(sino), _ptr(back), int(nthreads))
    if nnz < 0:
        raise RuntimeError("oracle_fp_bp failed")
    return sino, back, int(nnz)


def siddons(src, trg, nPixels=128, dPixels=1.0, origin=0.0, ravel=False,
            flat=False, epsilon=1e-5):
    """Reference-format output (see module docstring)."""
    shape, x, y, z, ln, off, bad, n = trace(src, trg, nPixels, dPixels, origin,
                                            epsilon)
    if (flat or ravel) and (bad & 1).any():
        raise ValueError("invalid entry in coordinates array")
    if flat:
        ind = ((x.astype(np.int64) * n[1] + y) * n[2] + z).astype(np.int32)
        return ind, ln, off.astype(np.int32)
    rays = np.empty(shape, dtype=object)
    flat_rays = rays.reshape(-1)
    for r in range(flat_rays.size):
        s, e = off[r], off[r + 1]
        if not (bad[r] & 2):
            continue
        if ravel:
            ind = ((x[s:e].astype(np.int64) * n[1] + y[s:e]) * n[2] + z[s:e])
            flat_rays[r] = [ind.astype(np.int32), ln[s:e]]
        else:
            flat_rays[r] = [x[s:e], y[s:e], z[s:e], ln[s:e]]
    return rays

