"""Utilities over Siddon ray lists (sys_mat/siddon.py:12-633).

The reference keeps a system matrix as an object array of per-ray lists --
[ix, iy, iz, length] (unraveled) or [ravel_index, length] (raveled), None
for rays that miss -- or as the flat CSR triple (ind, len, off).  These
functions keep its names, arguments, in-place behaviour and output formats.

Index and format transforms (ravel/unravel, flips, shifts, trimming,
flattening, counting) are host-side array bookkeeping on the object array,
as in the reference.  The two reductions -- per-voxel sums of lengths
(list2array) and the merge of several rays into one averaged list
(rays_ave, rays_weight_ave) -- run on the GPU: the entries are sorted by
voxel with a stable radix sort and each voxel's lengths are summed in float32
in their original order (vir_csr_transpose + vir_csc_bp in sequential mode),
which reproduces the reference's float32 accumulation loops
(siddon.py:116-120, 381-383, 463-465) bit for bit.
"""
import numpy as np
import torch

from .. import _lib
from ..device import require_gpu, stream


def _is_none(ray):
    return ray is None


def _voxel_sums(ind, ln, nvox):
    """(per-voxel float32 sums of ln over entries ind, per-voxel counts),
    summed in entry order on the GPU."""
    require_gpu()
    from .siddon import csr_transpose
    ind = np.ascontiguousarray(ind, dtype=np.int32).reshape(-1)
    ln = np.ascontiguousarray(ln, dtype=np.float32).reshape(-1)
    if ind.size and (ind.min() < 0 or ind.max() >= nvox):
        raise ValueError("invalid entry in coordinates array")
    ind_d = torch.from_numpy(ind).cuda()
    ln_d = torch.from_numpy(ln).cuda()
    off_d = torch.tensor([0, ind.size], dtype=torch.int64, device="cuda")
    colptr, row, clen = csr_transpose(ind_d, ln_d, off_d, nvox)
    ones = torch.ones(1, dtype=torch.float32, device="cuda")
    out = torch.empty(nvox, dtype=torch.float32, device="cuda")
    _lib.check(_lib.lib().vir_csc_bp(_lib.ptr(colptr), _lib.ptr(row), _lib.ptr(clen), nvox,
                                     _lib.ptr(ones), _lib.ptr(out), _lib.VIR_SEQUENTIAL, stream()),
               "vir_csc_bp")
    counts = (colptr[1:] - colptr[:-1]).cpu().numpy()
    return out.cpu().numpy(), counts


# ------------------------------------------------------- single rays

def ray_ravel(ray, nPixels):
    """[ix, iy, iz, len] -> [ravel_index, len] (siddon.py:592-611)."""
    return [np.ravel_multi_index((ray[0], ray[1], ray[2]), nPixels), ray[3]]


def ray_unravel(ray, nPixels):
    """[ravel_index, len] -> [ix, iy, iz, len] (siddon.py:614-633)."""
    return list(np.unravel_index(ray[0], nPixels) + (ray[1],))


def ray_index_shift(ray, shift):
    """Adds shift (3) to the voxel indices (siddon.py:251-252)."""
    return [ray[0] + shift[0], ray[1] + shift[1], ray[2] + shift[2], ray[3]]


def ray_boundary(ray, nPixels):
    """Keeps the entries inside the grid (siddon.py:254-262)."""
    keep = ((ray[0] >= 0) & (ray[0] < nPixels[0]) & (ray[1] >= 0) & (ray[1] < nPixels[1])
            & (ray[2] >= 0) & (ray[2] < nPixels[2]))
    return [ray[0][keep], ray[1][keep], ray[2][keep], ray[3][keep]]


def ray_flip(ray, x=None, y=None, z=None, ravel=False, nPixels=None):
    """Mirrors the indices about x, y and/or z (index i -> 2*x - i); an axis
    value of 0 or None leaves that axis alone, as in siddon.py:214-248."""
    if ravel:
        ray = ray_unravel(ray, nPixels)
    if x:
        ray = [(2 * x - ray[0]).astype(int), ray[1], ray[2], ray[3]]
    if y:
        ray = [ray[0], (2 * y - ray[1]).astype(int), ray[2], ray[3]]
    if z:
        ray = [ray[0], ray[1], (2 * z - ray[2]).astype(int), ray[3]]
    if ravel:
        return [np.ravel_multi_index((ray[0], ray[1], ray[2]), nPixels), ray[3]]
    return ray


# ------------------------------------------------------- object arrays

def list_ravel(sdlist, nPixels):
    """In place (siddon.py:538-562)."""
    for idx, ray in np.ndenumerate(sdlist):
        if not _is_none(ray):
            sdlist[idx] = ray_ravel(ray, nPixels)
    return sdlist


def list_unravel(sdlist, nPixels):
    """In place (siddon.py:565-589)."""
    for idx, ray in np.ndenumerate(sdlist):
        if not _is_none(ray):
            sdlist[idx] = ray_unravel(ray, nPixels)
    return sdlist


def list_index_shift(sdlist, shift, ravel=False):
    """In place (siddon.py:57-93); unraveled lists only."""
    shift = np.array(shift, dtype=int)
    if ravel:
        raise ValueError('Cuurently unsuported, ravel must be False')
    for idx, ray in np.ndenumerate(sdlist):
        if not _is_none(ray):
            sdlist[idx] = ray_index_shift(ray, shift)
    return sdlist


def list_del_ind(sdlist, X_ind, Y_ind, Z_ind, ravel=False):
    """Keeps the entries whose x, y and z indices are all listed, in place
    (siddon.py:12-54); unraveled lists only."""
    if ravel:
        raise ValueError('Cuurently unsuported, ravel must be False')
    for idx, ray in np.ndenumerate(sdlist):
        if not _is_none(ray):
            keep = np.isin(ray[0], X_ind) & np.isin(ray[1], Y_ind) & np.isin(ray[2], Z_ind)
            sdlist[idx] = [ray[0][keep], ray[1][keep], ray[2][keep], ray[3][keep]]
    return sdlist


def list_flip(sdlist, x=None, y=None, z=None, ravel=False, nPixels=None):
    """In place (siddon.py:205-211)."""
    for idx, ray in np.ndenumerate(sdlist):
        if not _is_none(ray):
            sdlist[idx] = ray_flip(ray, x=x, y=y, z=z, ravel=ravel, nPixels=nPixels)
    return sdlist


def list_count_elem(sdlist):
    """Entries per ray, int32 array of the list's shape; a single ray gives
    its count (siddon.py:166-191)."""
    if not isinstance(sdlist, np.ndarray):
        return sdlist[0].size
    count = np.zeros(sdlist.shape, dtype=np.int32)
    for idx, ray in np.ndenumerate(sdlist):
        if not _is_none(ray):
            count[idx] = ray[0].size
    return count


def list_flatten(sdlist, ravel=False, nPixels=None):
    """Object array -> (ind int32, len float32, off int32) CSR
    (siddon.py:490-535)."""
    if ravel is False and nPixels is None:
        raise ValueError('nPixels must be provided if ravel=False')
    off = np.insert(np.cumsum(list_count_elem(sdlist)), 0, 0).astype(np.int32)
    ind = np.zeros(off[-1], dtype=np.int32)
    ln = np.zeros(off[-1], dtype=np.float32)
    for k, ray in enumerate(sdlist.flat):
        if not _is_none(ray):
            if not ravel:
                ray = ray_ravel(ray, nPixels)
            ind[off[k]:off[k + 1]] = ray[0]
            ln[off[k]:off[k + 1]] = ray[1]
    return ind, ln, off


def list2array_iter_ray(ray, array, ravel=False):
    """Adds one ray's lengths into `array` (duplicates accumulate,
    siddon.py:96-121)."""
    if ravel:
        np.add.at(array, np.unravel_index(np.asarray(ray[0]), array.shape),
                  np.asarray(ray[1], array.dtype))
    else:
        np.add.at(array, (ray[0], ray[1], ray[2]), np.asarray(ray[3], array.dtype))


def list2array(sdlist, nPixels, ravel=False, flat=False):
    """Per-voxel sum of intersection lengths, float32 array nPixels
    (siddon.py:124-163) -- the backprojection of ones, on the GPU."""
    nPixels = tuple(int(v) for v in np.atleast_1d(nPixels))
    nvox = int(np.prod(nPixels))
    if flat:
        ind, ln = np.asarray(sdlist[0]), np.asarray(sdlist[1])
    else:
        ind, ln, _ = list_flatten(sdlist, ravel=ravel, nPixels=nPixels)
    sums, _ = _voxel_sums(ind, ln, nvox)
    return sums.reshape(nPixels)


def _merged(ind, ln, nvox):
    sums, counts = _voxel_sums(ind, ln, nvox)
    uniq = np.nonzero(counts)[0].astype(np.int64)
    return uniq, sums[uniq]


def _nvox_for(flat_ind, nPixels):
    if nPixels is not None:
        return int(np.prod(nPixels))
    return int(flat_ind.max()) + 1 if flat_ind.size else 1


def rays_ave(sdlist, ravel=False, flat=False, nPixels=None, nRays=None):
    """Merges rays into one list: unique voxels, float32 length sums in
    entry order, divided by the number of rays (siddon.py:432-487)."""
    if ravel is False and nPixels is None:
        raise ValueError('nPixels must be provided if ravel=False')
    if flat is True and nRays is None:
        raise ValueError('nRays must be provided if flat=True')
    if flat is False and nPixels is None:
        raise ValueError('nPixels must be provided if flat=False')
    if flat:
        flat_ind, flat_len, _ = sdlist
        flat_ind = np.asarray(flat_ind)
    else:
        nRays = sdlist.size
        flat_ind, flat_len, _ = list_flatten(sdlist, nPixels=nPixels, ravel=ravel)
    if flat_ind.size == 0:
        return None
    uniq, sums = _merged(flat_ind, flat_len, _nvox_for(flat_ind, nPixels))
    sums = (sums / np.float32(nRays)).astype(np.float32)
    if ravel or flat:
        return [uniq.astype(flat_ind.dtype), sums]
    return ray_unravel([uniq, sums], nPixels)


def list_ave(sdlist, ravel=False, flat=False, nPixels=None, nRays=None, axis=None):
    """rays_ave over all rays, or over the given axes of the object array
    (siddon.py:388-429)."""
    if flat is True and axis is not None:
        raise ValueError('Axis cannot be provided when flat=True')
    if axis is None:
        return rays_ave(sdlist, ravel=ravel, flat=flat, nPixels=nPixels, nRays=nRays)
    out = np.empty(np.delete(sdlist.shape, axis), dtype=object)
    moved = np.moveaxis(sdlist, axis, np.arange(-1 * (np.array(axis).size), 0))
    for idx, _ in np.ndenumerate(out):
        out[idx] = rays_ave(moved[idx], ravel=ravel, flat=flat, nPixels=nPixels, nRays=nRays)
    return out


def rays_weight_ave(sdlist, weights, ravel=False, nPixels=None):
    """Merges rays with per-ray weights: unique voxels, float32 sums of
    weight * length (siddon.py:350-385)."""
    if ravel is False and nPixels is None:
        raise ValueError('nPixels must be provided if ravel=False')
    weighted = np.empty(sdlist.shape, dtype=object)
    for idx, ray in np.ndenumerate(sdlist):
        if _is_none(ray):
            weighted[idx] = None
        else:
            if not ravel:
                ray = ray_ravel(ray, nPixels)
            weighted[idx] = [ray[0], weights[idx] * ray[1]]
    flat_ind, flat_len, _ = list_flatten(weighted, nPixels=nPixels, ravel=True)
    if flat_ind.size == 0:
        return None
    uniq, sums = _merged(flat_ind, flat_len, _nvox_for(flat_ind, nPixels))
    if ravel:
        return [uniq.astype(flat_ind.dtype), sums]
    return ray_unravel([uniq, sums], nPixels)


def list_weight_ave(sdlist, weights, ravel=False, nPixels=None, axis=None):
    """rays_weight_ave over the given axes (siddon.py:328-347)."""
    if axis is None:
        return rays_weight_ave(sdlist, weights, ravel=ravel, nPixels=nPixels)
    if len(axis) != weights.ndim:
        raise ValueError('Number of axis must match weights dimensions')
    out = np.empty(np.delete(sdlist.shape, axis), dtype=object)
    moved = np.moveaxis(sdlist, axis, np.arange(-1 * (np.array(axis).size), 0))
    for idx, _ in np.ndenumerate(out):
        out[idx] = rays_weight_ave(moved[idx], weights, ravel=ravel, nPixels=nPixels)
    return out


def list_convole(sdlist, kernel, ravel=False, nPixels=None, axis=None):
    """Convolves the ray list with an odd-sized kernel along `axis` (a
    weighted merge of neighbouring rays, kernel renormalised at the edges;
    siddon.py:265-325)."""
    if len(axis) != kernel.ndim:
        raise ValueError('Number of axis must match kernel dimensions')
    for dim in kernel.shape:
        if dim % 2 == 0:
            raise ValueError('Kernel must have an odd numer of elements in each dimension')
    k_dims = kernel.shape
    sd_dims = sdlist.shape
    out = np.empty(sdlist.shape, dtype=object)
    for ray_idx, _ in np.ndenumerate(sdlist):
        ray_sl = list(ray_idx)
        ker_sl = []
        for i, (ax, k_dim) in enumerate(zip(axis, k_dims)):
            lo = ray_sl[ax] - int(k_dim / 2)
            k_lo = -lo if lo < 0 else 0
            lo = max(lo, 0)
            hi = ray_sl[ax] + int(k_dim / 2) + 1
            k_hi = k_dims[i] - (hi - sd_dims[ax]) if hi > sd_dims[ax] else k_dims[i]
            hi = min(hi, sd_dims[ax])
            ray_sl[ax] = slice(lo, hi)
            ker_sl.append(slice(k_lo, k_hi))
        ker_sl = tuple(ker_sl)
        kF = kernel.sum() / kernel[ker_sl].sum()
        out[ray_idx] = rays_weight_ave(sdlist[tuple(ray_sl)], kF * kernel[ker_sl], ravel=ravel,
                                       nPixels=nPixels)
    return out
