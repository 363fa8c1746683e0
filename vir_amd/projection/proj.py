"""Stored-system-matrix forward/back projection on the GPU.

Drop-ins for the reference's projection operators over a Siddon system matrix
(the output of siddons()):

  sd_f_proj(phantom, sdlist, ravel=True, flat=True, sino_shape=None)
        projection/proj.py:56-116 (numpy), proj_numba.py:35-56 (numba),
        proj_cuda.py:30-62 (numba.cuda), proj_ctype.py:27-55 (ctypes C)
  sd_b_proj(sino, sdlist, nPixels, ravel=True)
        projection/proj.py:119-150, proj_ctype.py:57-90

`sdlist` is any siddons() output: the CSR tuple (flat=True), an object array
of [ravel_idx, len] (ravel=True) or of [x, y, z, len].  The forward projection
is a CSR gather (vir_csr_fp); the backprojection is an atomic-free gather
over the transposed (voxel-major) matrix (vir_csr_transpose + vir_csc_bp),
which accumulates repeated voxels of one ray like projection_c.c:176-219.
numpy arrays in -> numpy arrays out; CUDA tensors in -> CUDA tensors out.
"""
import numpy as np
import torch

from .. import _lib
from ..device import as_device, as_device_f32, require_gpu, stream
from ..sys_mat.siddon import csr_transpose


def _is_flat(sdlist):
    return isinstance(sdlist, (tuple, list)) and len(sdlist) == 3 and \
        not isinstance(sdlist, np.ndarray)


def _to_csr(sdlist, nPixels, ravel):
    """CSR (ind int32, len f32, off int64) numpy arrays of an sdlist."""
    if _is_flat(sdlist):
        ind, ln, off = sdlist
        if isinstance(ind, torch.Tensor):
            return ind, ln, off
        return (np.asarray(ind, np.int32), np.asarray(ln, np.float32),
                np.asarray(off, np.int64))
    flat = np.asarray(sdlist, dtype=object).reshape(-1)
    counts = np.zeros(flat.size, np.int64)
    inds, lens = [], []
    for i, ray in enumerate(flat):
        if ray is None:
            continue
        if ravel:
            idx = np.asarray(ray[0], np.int64)
            ln = np.asarray(ray[1], np.float32)
        else:
            idx = np.ravel_multi_index((np.asarray(ray[0]), np.asarray(ray[1]),
                                        np.asarray(ray[2])), tuple(nPixels))
            ln = np.asarray(ray[3], np.float32)
        counts[i] = idx.size
        inds.append(idx.astype(np.int32))
        lens.append(ln)
    off = np.zeros(flat.size + 1, np.int64)
    np.cumsum(counts, out=off[1:])
    ind = np.concatenate(inds) if inds else np.zeros(0, np.int32)
    ln = np.concatenate(lens) if lens else np.zeros(0, np.float32)
    return ind, ln, off


def check_csr(ind, ln, off, nvox):
    """Reject a stored matrix whose entries the kernels would read out of
    bounds, before anything is launched: every index in [0, nvox) (the
    reference's fancy indexing raises IndexError, proj.py:92-93,127-128),
    off[0] == 0, off non-decreasing, off[-1] == len(ind) == len(ln)
    (ValueError).  numpy arrays are checked on the host, device tensors with
    torch reductions on the device."""
    if isinstance(off, torch.Tensor):
        nnz, nl, n_off = ind.numel(), ln.numel(), off.numel()
        first = int(off[0].item()) if n_off else 0
        last = int(off[-1].item()) if n_off else 0
        mono = n_off < 2 or bool((off[1:] >= off[:-1]).all().item())
        lo = int(ind.min().item()) if nnz else 0
        hi = int(ind.max().item()) if nnz else 0
    else:
        off = np.asarray(off)
        nnz, nl, n_off = np.size(ind), np.size(ln), off.size
        first = int(off[0]) if n_off else 0
        last = int(off[-1]) if n_off else 0
        mono = n_off < 2 or bool(np.all(off[1:] >= off[:-1]))
        lo = int(np.min(ind)) if nnz else 0
        hi = int(np.max(ind)) if nnz else 0
    if n_off < 1:
        raise ValueError("offsets must hold nrays + 1 >= 1 values")
    if first != 0 or not mono or last != nnz or nl != nnz:
        raise ValueError("offsets must start at 0, not decrease and end at the number of entries "
                         f"(off[0]={first}, off[-1]={last}, {nnz} indices, {nl} lengths)")
    if nnz and (lo < 0 or hi >= nvox):
        bad = lo if lo < 0 else hi
        raise IndexError(f"index {bad} is out of bounds for a volume of size {nvox}")


def _device_csr(csr, nvox):
    ind, ln, off = csr
    if isinstance(off, torch.Tensor):
        check_csr(ind, ln, off, nvox)
    return (as_device(ind, torch.int32), as_device_f32(ln), as_device(off, torch.int64))


def _ray_shape(sdlist, sino_shape):
    if sino_shape is not None:
        return tuple(np.atleast_1d(sino_shape).tolist()) if not isinstance(sino_shape, tuple) \
            else sino_shape
    if _is_flat(sdlist):
        return (len(sdlist[2]) - 1,)
    return np.asarray(sdlist, dtype=object).shape


def sd_f_proj(phantom, sdlist, ravel=True, flat=True, sino_shape=None, sequential=False):
    """Forward projection sino[r] = sum_j phantom[idx_j] * len_j.

    sequential=True sums each ray in entry order in plain float32, the
    rounding of projection_c.c:101-157; the default accumulates in float64 as
    proj_numba.py:17 / proj_cuda.py:22 do."""
    torch_in = isinstance(phantom, torch.Tensor)
    nPixels = tuple(phantom.shape)
    shape = _ray_shape(sdlist, sino_shape)
    nvox = int(np.prod(nPixels))
    csr = _to_csr(sdlist, nPixels, ravel or _is_flat(sdlist))
    check_csr(*csr, nvox)
    require_gpu()
    ind, ln, off = _device_csr(csr, nvox)
    nrays = off.numel() - 1
    vol = as_device_f32(phantom)
    sino = torch.zeros(int(np.prod(shape)), dtype=torch.float32, device=vol.device)
    if nrays > sino.numel():
        raise ValueError("sino_shape smaller than the number of rays")
    fl = _lib.VIR_SEQUENTIAL if sequential else 0
    _lib.check(_lib.lib().vir_csr_fp(_lib.ptr(ind), _lib.ptr(ln), _lib.ptr(off), nrays,
                                     _lib.ptr(vol), _lib.ptr(sino), fl, stream()), "vir_csr_fp")
    sino = sino.reshape(shape)
    return sino if torch_in else sino.cpu().numpy()


def last_of_repeats(ind, ln, off):
    """The stored matrix with every entry whose voxel appears again later in
    the same ray removed (host arrays).  numpy's buffered fancy-index `+=`
    (proj.py:134, 139-140) writes each repeated voxel once, from its last
    entry, so this is the matrix the reference's sd_b_proj applies."""
    ind = np.asarray(ind)
    off = np.asarray(off, np.int64)
    nnz = ind.size
    if nnz == 0:
        return ind, np.asarray(ln), off
    ray = np.repeat(np.arange(off.size - 1), np.diff(off))
    pos = np.arange(nnz)
    order = np.lexsort((pos, ind, ray))          # by ray, voxel, then position
    last = np.ones(nnz, bool)
    same = (ray[order][1:] == ray[order][:-1]) & (ind[order][1:] == ind[order][:-1])
    last[order[:-1][same]] = False               # a later entry of the same voxel follows
    counts = np.bincount(ray[last], minlength=off.size - 1)
    new_off = np.zeros_like(off)
    np.cumsum(counts, out=new_off[1:])
    return ind[last], np.asarray(ln)[last], new_off


def sd_b_proj(sino, sdlist, nPixels, ravel=True, sequential=False, compat=False):
    """Backprojection phantom[v] = sum over entries (r, v) of len * sino[r]
    (the exact transpose of sd_f_proj).  sequential=True reproduces the
    float32 rounding sequence of projection_c.c:176-208.

    compat=True reproduces the reference's numpy backprojector
    (proj.py:119-150): a voxel repeated within one ray receives only its last
    entry's value (fancy-index `+=`).  With sequential=True as well the
    result is the reference's bit for bit (rays added in order in float32,
    each product len * sino in float32)."""
    torch_in = isinstance(sino, torch.Tensor)
    nPixels = tuple(int(v) for v in np.atleast_1d(nPixels))
    nvox = int(np.prod(nPixels))
    csr = _to_csr(sdlist, nPixels, ravel or _is_flat(sdlist))
    if compat:
        csr = last_of_repeats(*(c.cpu().numpy() if isinstance(c, torch.Tensor) else c for c in csr))
    check_csr(*csr, nvox)
    if int(np.prod(np.shape(sino))) < len(csr[2]) - 1:
        # sino_flat[i] of proj.py:139-140 raises for a short sinogram
        raise IndexError(f"sinogram has {int(np.prod(np.shape(sino)))} values for {len(csr[2]) - 1} rays")
    require_gpu()
    ind, ln, off = _device_csr(csr, nvox)
    s = as_device_f32(sino).reshape(-1)
    colptr, row, clen = csr_transpose(ind, ln, off, nvox)
    vol = torch.empty(nvox, dtype=torch.float32, device=s.device)
    fl = _lib.VIR_SEQUENTIAL if sequential else 0
    _lib.check(_lib.lib().vir_csc_bp(_lib.ptr(colptr), _lib.ptr(row), _lib.ptr(clen), nvox,
                                     _lib.ptr(s), _lib.ptr(vol), fl, stream()), "vir_csc_bp")
    vol = vol.reshape(nPixels)
    return vol if torch_in else vol.cpu().numpy()


def sd_f_proj_numba(phantom, sdlist, sino_shape):
    """proj_numba.sd_f_proj_numba (proj_numba.py:35-44): CSR forward projection."""
    return sd_f_proj(phantom, sdlist, flat=True, sino_shape=sino_shape)


def sd_f_proj_numba_p(phantom, sdlist, sino_shape):
    """proj_numba.sd_f_proj_numba_p (proj_numba.py:47-56)."""
    return sd_f_proj(phantom, sdlist, flat=True, sino_shape=sino_shape)


def sd_f_proj_numba_g(phantom, sdlist, sino_shape):
    """proj_cuda.sd_f_proj_numba_g (proj_cuda.py:30-62)."""
    return sd_f_proj(phantom, sdlist, flat=True, sino_shape=sino_shape)
