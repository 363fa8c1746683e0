"""The ctypes projection calls (projection/proj_ctype.py:27-90) on the GPU.

  sd_f_proj_c(phantom, sino, sdlist_c, ravel=True, C=False, dims=None, nRays=None)
  sd_b_proj_c(phantom, sino, sdlist_c, ravel=True, C=False, dims=None, nRays=None)

sdlist_c is what the reference passes to projection.so: an array of C ray
records (mpct.list_ctypes_object(sdlist, ravel, flat=True), a byref of it,
or a pointer to its first record).  Like forward_proj_*_rays_struct and
back_proj_*_rays_struct (projection_c.c:150-219) the calls return None and
ADD into the caller's arrays: sino[r] += sum_i phantom[R_i] * L_i, then
phantom[R_i] += sino[r] * L_i over the rays in order, in float32 with the C
code's rounding sequence (the stored-matrix kernels' VIR_SEQUENTIAL mode;
bit-identical to the compiled reference, tested).

C=False: phantom and sino are numpy arrays passed through mpct.ctypes_vars,
as in the reference, so an array that is not float32 is converted to a
float32 copy and the result lands in that copy (it does not reach the
caller's array, proj_ctype.py:38-39); nRays = sino.size, dims =
phantom.shape.  C=True: phantom and sino are float pointers (ctypes) with
dims = (nX, nY, nZ) and nRays given; the calls read and write through them.

The ray records are read on the host into a stored matrix; the products run
in libvir_amd (vir_csr_fp, vir_csr_transpose + vir_csc_bp).
"""
import ctypes

import numpy as np
import torch

from .. import _lib
from .. import mpct
from ..device import require_gpu, stream
from .proj import check_csr


def _records(sdlist_c, ravel, nrays):
    """The first nrays records of sdlist_c as a ctypes array."""
    rec = mpct.RaveledRay if ravel else mpct.UnraveledRay
    obj = getattr(sdlist_c, "_obj", sdlist_c)        # byref(...) -> the object
    if isinstance(obj, ctypes.Array):
        if not issubclass(obj._type_, rec):
            raise TypeError(f"sdlist_c holds {obj._type_.__name__}, expected {rec.__name__}")
        if len(obj) < nrays:
            raise IndexError(f"sdlist_c has {len(obj)} rays, {nrays} requested")
        return obj
    if isinstance(obj, ctypes._Pointer):
        return ctypes.cast(obj, ctypes.POINTER(rec * nrays)).contents
    if isinstance(obj, rec):
        if nrays > 1:
            raise IndexError("a single ray record for several rays")
        return (rec * 1)(obj)
    raise TypeError("sdlist_c must be an array of ray records (mpct.list_ctypes_object)")


def _csr(recs, nrays, ravel, dims):
    """(ind int32, len float32, off int64) of the records, flat index
    nY*nZ*x + nZ*y + z (ravel_ray_struct, projection_c.c:95-98)."""
    nX, nY, nZ = dims
    counts = np.zeros(nrays, np.int64)
    inds, lens = [], []
    for r in range(nrays):
        n = int(recs[r].n)
        if n <= 0:
            continue
        counts[r] = n
        L = np.ctypeslib.as_array(recs[r].L, shape=(n,)).astype(np.float32)
        if ravel:
            idx = np.ctypeslib.as_array(recs[r].R, shape=(n,)).astype(np.int64)
        else:
            x = np.ctypeslib.as_array(recs[r].X, shape=(n,)).astype(np.int64)
            y = np.ctypeslib.as_array(recs[r].Y, shape=(n,)).astype(np.int64)
            z = np.ctypeslib.as_array(recs[r].Z, shape=(n,)).astype(np.int64)
            idx = nY * nZ * x + nZ * y + z
        inds.append(idx)
        lens.append(L)
    off = np.zeros(nrays + 1, np.int64)
    np.cumsum(counts, out=off[1:])
    ind = np.concatenate(inds).astype(np.int32) if inds else np.zeros(0, np.int32)
    ln = np.concatenate(lens) if lens else np.zeros(0, np.float32)
    return ind, ln, off


def _views(phantom, sino, C, dims, nRays):
    """float32 numpy views of the buffers the C call would write, dims,
    nRays."""
    if C:
        if dims is None or nRays is None:
            raise ValueError("C=True needs dims and nRays")
        nX, nY, nZ = (int(v) for v in dims)
        ph = np.ctypeslib.as_array(ctypes.cast(phantom, mpct.c_float_p), shape=(nX * nY * nZ,))
        sn = np.ctypeslib.as_array(ctypes.cast(sino, mpct.c_float_p), shape=(int(nRays),))
        return ph, sn, (nX, nY, nZ), int(nRays)
    sino32, _ = mpct.ctypes_vars(sino)
    phantom32, _ = mpct.ctypes_vars(phantom)
    dims = tuple(int(v) for v in phantom32.shape)
    if len(dims) != 3:
        raise ValueError("phantom must be 3-D (nX, nY, nZ)")
    if not (phantom32.flags.c_contiguous and sino32.flags.c_contiguous):
        raise ValueError("phantom and sino must be C-contiguous")
    return phantom32.reshape(-1), sino32.reshape(-1), dims, int(sino32.size)


def sd_f_proj_c(phantom, sino, sdlist_c, ravel=True, C=False, dims=None, nRays=None):
    """sino += A phantom (forward_proj_{ravel,unravel}_rays_struct)."""
    ph, sn, dims, n = _views(phantom, sino, C, dims, nRays)
    recs = _records(sdlist_c, ravel, n)
    ind, ln, off = _csr(recs, n, ravel, dims)
    check_csr(ind, ln, off, ph.size)
    require_gpu()
    d = [torch.from_numpy(a).cuda() for a in (ind, ln, off, ph, sn)]
    _lib.check(_lib.lib().vir_csr_fp(_lib.ptr(d[0]), _lib.ptr(d[1]), _lib.ptr(d[2]), n, _lib.ptr(d[3]),
                                     _lib.ptr(d[4]), _lib.VIR_SEQUENTIAL | _lib.VIR_ACCUMULATE, stream()),
               "vir_csr_fp")
    sn[:] = d[4].cpu().numpy()


def sd_b_proj_c(phantom, sino, sdlist_c, ravel=True, C=False, dims=None, nRays=None):
    """phantom += A^T sino (back_proj_{ravel,unravel}_rays_struct)."""
    from ..sys_mat.siddon import csr_transpose
    ph, sn, dims, n = _views(phantom, sino, C, dims, nRays)
    recs = _records(sdlist_c, ravel, n)
    ind, ln, off = _csr(recs, n, ravel, dims)
    check_csr(ind, ln, off, ph.size)
    require_gpu()
    ind_d, ln_d, off_d, ph_d, sn_d = (torch.from_numpy(a).cuda() for a in (ind, ln, off, ph, sn))
    colptr, row, clen = csr_transpose(ind_d, ln_d, off_d, ph.size)
    _lib.check(_lib.lib().vir_csc_bp(_lib.ptr(colptr), _lib.ptr(row), _lib.ptr(clen), ph.size, _lib.ptr(sn_d),
                                     _lib.ptr(ph_d), _lib.VIR_SEQUENTIAL | _lib.VIR_ACCUMULATE, stream()),
               "vir_csc_bp")
    ph[:] = ph_d.cpu().numpy()
