"""Siddon system matrix and projectors on the GPU.

Drop-in for the reference's ``sys_mat.siddon.siddons`` (sys_mat/siddon.py:
636-833): same arguments, same output formats, same (bit-identical) voxel
indices, lengths and counts -- computed by libvir_amd's HIP kernels.

``SiddonProjector`` is the device-resident operator behind it: on-the-fly
forward projection (no stored matrix) and its exact adjoint, the
brick-gather backprojection, for ray sets laid out per view as produced by
``geom_circular``; other ray sets fall back to an explicit stored matrix
transposed into voxel-major form.
"""
import numpy as np
import torch

from .. import _lib
from ..core import np_array
from ..device import as_device, as_device_f32, out_tensor, require_gpu, stream
from ..proj_geom import check_in_front, check_projections, view_projections


def grid_params(nPixels, dPixels, origin):
    """Grid3d(nPixels, dPixels, origin, float32) parameters
    (core/core_funcs.py:235-265)."""
    n = np_array(nPixels, 3, dtype=int).astype(np.int64)
    d = np_array(dPixels, 3, dtype=np.float32)
    o = np_array(origin, 3, dtype=np.float32)
    if n.size != 3 or d.size != 3 or o.size != 3:
        raise ValueError("nPixels/dPixels/origin must be scalars or 3 values")
    return n, d, o


def _prep_rays(src, trg):
    """float32 casts and the (...,3) layout of siddon.py:684-690."""
    if isinstance(trg, torch.Tensor):
        trg = trg.detach().to("cpu").numpy()
    if isinstance(src, torch.Tensor):
        src = src.detach().to("cpu").numpy()
    trg = np.array(trg, dtype=np.float32)
    if trg.ndim == 1:
        trg = trg[np.newaxis, :]
    src = np.array(src, dtype=np.float32)
    if src.ndim == 1:
        src = src[np.newaxis, :]
    if trg.shape[-1] != 3 or src.shape[-1] != 3:
        raise ValueError("src and trg must have a trailing dimension of 3")
    src = np.broadcast_to(src, trg.shape)
    return src, trg, trg.shape[:-1]


def _trace(src_d, trg_d, grid, nrays):
    """Count + scan + fill on the GPU.  Returns device tensors."""
    L = _lib.lib()
    st = stream()
    dev = src_d.device
    counts = torch.empty(max(nrays, 1), dtype=torch.int32, device=dev)
    flags = torch.empty(max(nrays, 1), dtype=torch.uint8, device=dev)
    _lib.check(L.vir_siddon_count(_lib.ptr(src_d), _lib.ptr(trg_d), nrays, grid,
                                  _lib.ptr(counts), _lib.ptr(flags), st), "vir_siddon_count")
    off = torch.zeros(nrays + 1, dtype=torch.int64, device=dev)
    if nrays:
        torch.cumsum(counts[:nrays].to(torch.int64), 0, out=off[1:])
    nnz = int(off[-1].item()) if nrays else 0
    rav = torch.empty(max(nnz, 1), dtype=torch.int32, device=dev)
    ix = torch.empty(max(nnz, 1), dtype=torch.int32, device=dev)
    iy = torch.empty(max(nnz, 1), dtype=torch.int32, device=dev)
    iz = torch.empty(max(nnz, 1), dtype=torch.int32, device=dev)
    ln = torch.empty(max(nnz, 1), dtype=torch.float32, device=dev)
    _lib.check(L.vir_siddon_fill(_lib.ptr(src_d), _lib.ptr(trg_d), nrays, grid, _lib.ptr(off),
                                 _lib.ptr(rav), _lib.ptr(ix), _lib.ptr(iy), _lib.ptr(iz),
                                 _lib.ptr(ln), st), "vir_siddon_fill")
    return off, flags[:nrays], rav[:nnz], ix[:nnz], iy[:nnz], iz[:nnz], ln[:nnz]


def siddons(src, trg, nPixels=128, dPixels=1.0, origin=0.0, ravel=False, flat=False,
            epsilon=1e-5):
    """Siddon's algorithm (Med. Phys. 12, 252-255) on the GPU, with the call
    surface and outputs of the reference's siddons() (sys_mat/siddon.py:636).

    Returns, like the reference:
      flat=True  -> (flat_ind int32[nnz], flat_len float32[nnz],
                     flat_off int32[nRays+1])
      ravel=True -> object array (ray shape) of [ravel_idx int32, len f32]
                    or None for rays that miss the grid
      otherwise  -> object array of [x int32, y int32, z int32, len f32]
    ValueError is raised in flat/ravel modes when an index falls outside the
    grid, as np.ravel_multi_index does in the reference (siddon.py:815,821).
    """
    require_gpu()
    n, d, o = grid_params(nPixels, dPixels, origin)
    grid = _lib.make_grid(n, d, o, epsilon)
    src, trg, shape = _prep_rays(src, trg)
    nrays = int(np.prod(shape))
    src_d = as_device_f32(src.reshape(-1, 3))
    trg_d = as_device_f32(trg.reshape(-1, 3))
    off, flags, rav, ix, iy, iz, ln = _trace(src_d, trg_d, grid, nrays)
    flags = flags.cpu().numpy()
    if (flat or ravel) and (flags & 1).any():
        raise ValueError("invalid entry in coordinates array")
    if flat:
        return (rav.cpu().numpy(), ln.cpu().numpy(), off.cpu().numpy().astype(np.int32))
    offh = off.cpu().numpy()
    lnh = ln.cpu().numpy()
    rays = np.empty(shape, dtype=object)
    flat_rays = rays.reshape(-1)
    if ravel:
        ravh = rav.cpu().numpy()
        for r in np.nonzero(flags & 2)[0]:
            s, e = offh[r], offh[r + 1]
            flat_rays[r] = [ravh[s:e], lnh[s:e]]
    else:
        xh, yh, zh = ix.cpu().numpy(), iy.cpu().numpy(), iz.cpu().numpy()
        for r in np.nonzero(flags & 2)[0]:
            s, e = offh[r], offh[r + 1]
            flat_rays[r] = [xh[s:e], yh[s:e], zh[s:e], lnh[s:e]]
    return rays


class SiddonProjector:
    """Device-resident Siddon projector pair A / A^T for one ray set.

    src, trg: (..., 3) ray endpoints (numpy or torch; cast to float32 as the
    reference does).  When the ray array is laid out per view --
    (nviews, nu) or (nviews, nu, nv) -- and every view is a point source or a
    parallel bundle onto an affine detector grid (what geom_circular
    produces), the backprojector is the on-the-fly brick gather; otherwise it
    is a stored matrix transposed to voxel-major order.  Both are the exact
    transpose of fp(): they visit the same (voxel, length) entries.
    """

    def __init__(self, src, trg, nPixels=128, dPixels=1.0, origin=0.0, epsilon=1e-5,
                 structured=None, proj=None, records=True):
        """records: build per-ray records once (80 bytes per ray;
        vir_siddon_ray_records) so fp() and the gather bp() read them instead
        of recomputing each ray's setup per call -- same bits.
        proj: optional per-view projection matrices (nviews, 12) for rays
        already on the device (float32 CUDA tensors of shape (nviews, nu[, nv],
        3)), e.g. from vir_amd.geometry.circular_rays; the rays then stay
        on the device."""
        require_gpu()
        n, d, o = grid_params(nPixels, dPixels, origin)
        self.nPixels = tuple(int(v) for v in n)
        self.grid = _lib.make_grid(n, d, o, epsilon)
        self.nvox = int(np.prod(self.nPixels))
        self._csc = None
        self.proj = None
        if proj is not None:
            if not (isinstance(src, torch.Tensor) and isinstance(trg, torch.Tensor)):
                raise ValueError("proj= needs src and trg as device tensors")
            shape = tuple(trg.shape[:-1])
            if len(shape) not in (2, 3) or tuple(src.shape) != tuple(trg.shape):
                raise ValueError("device rays must be (nviews, nu[, nv], 3)")
            self.ray_shape = shape
            self.nrays = int(np.prod(shape))
            self.src = as_device_f32(src).reshape(-1, 3)
            self.trg = as_device_f32(trg).reshape(-1, 3)
            mats = np.asarray(proj, np.float32).reshape(shape[0], 12)
            if not check_in_front(mats, n, d, o):
                raise ValueError("a view's source lies inside the volume's guard region")
            # the backprojector takes its candidate rays and colour strides
            # from these matrices: they must describe these rays
            nv_ = shape[2] if len(shape) == 3 else 1
            check_projections(self.src.reshape(shape[0], shape[1], nv_, 3),
                              self.trg.reshape(shape[0], shape[1], nv_, 3), mats)
            self.proj = torch.from_numpy(mats).to(self.src.device)
            self.layout = (shape[0], shape[1], shape[2] if len(shape) == 3 else 1)
        else:
            src, trg, shape = _prep_rays(src, trg)
            self.ray_shape = shape
            self.nrays = int(np.prod(shape))
            self.src = as_device_f32(src.reshape(-1, 3))
            self.trg = as_device_f32(trg.reshape(-1, 3))
            if structured is not False and len(shape) in (2, 3) and self.nrays > 0:
                nviews, nu = shape[0], shape[1]
                nv = shape[2] if len(shape) == 3 else 1
                try:
                    mats = view_projections(src.reshape(nviews, nu, nv, 3),
                                            trg.reshape(nviews, nu, nv, 3))
                    if not check_in_front(mats, n, d, o):
                        raise ValueError("a view's source lies inside the volume's guard region")
                    self.proj = torch.from_numpy(mats).to(self.src.device)
                    self.layout = (nviews, nu, nv)
                except ValueError:
                    if structured:
                        raise
                    self.proj = None
        self._err = torch.zeros(1, dtype=torch.int32, device=self.src.device)
        # rays whose tracing needs the range-checked division or a float64
        # alpha list; when there are none, FP/BP run kernels built without
        # those paths (same bits, fewer registers)
        guarded = torch.zeros(1, dtype=torch.int32, device=self.src.device)
        _lib.check(_lib.lib().vir_siddon_count_guarded(_lib.ptr(self.src), _lib.ptr(self.trg), self.nrays,
                                                       self.grid, _lib.ptr(guarded), stream()),
                   "vir_siddon_count_guarded")
        self.guarded_rays = int(guarded.item())
        self._ung = _lib.VIR_UNGUARDED if self.guarded_rays == 0 else 0
        ws = 0
        self.rec = None
        if self.proj is not None:
            ws = int(_lib.lib().vir_siddon_bp_workspace(self.grid, self.layout[0]))
            if records and self.nrays > 0:
                # the per-ray setup of the gather backprojector, built once
                # (80 bytes per ray; records=False recomputes it per call)
                self.rec = torch.empty(self.nrays * (_lib.SIDDON_RECORD_BYTES // 4), dtype=torch.float32,
                                       device=self.src.device)
                _lib.check(_lib.lib().vir_siddon_ray_records(_lib.ptr(self.src), _lib.ptr(self.trg),
                                                             self.nrays, self.grid, _lib.ptr(self.rec),
                                                             stream()), "vir_siddon_ray_records")
        # one extra volume lets VIR_ACCUMULATE stage through the workspace
        self._ws = torch.empty(max(ws, self.nvox * 4 if self.proj is not None else 0) // 4 + 1,
                               dtype=torch.float32, device=self.src.device)

    def device_bytes(self):
        """Device memory held by this projector (rays, matrices, workspace)."""
        b = (self.src.numel() + self.trg.numel() + self._ws.numel()) * 4
        b += self.rec.numel() * 4 if self.rec is not None else 0
        return b + (self.proj.numel() * 4 if self.proj is not None else 0)

    # -- forward ---------------------------------------------------------
    def fp(self, vol, out=None, accumulate=False, check=False):
        """sino = A vol (float32, ray shape).  check=True raises ValueError
        when a ray produced an out-of-grid index (the reference raises)."""
        vol = as_device_f32(vol)
        if vol.numel() != self.nvox:
            raise ValueError("volume size does not match the grid")
        out = out_tensor(out, self.ray_shape, vol.device)
        if check:
            self._err.zero_()
        fl = (_lib.VIR_ACCUMULATE if accumulate else 0) | self._ung
        err = _lib.ptr(self._err) if check else None
        if self.rec is not None:
            _lib.check(_lib.lib().vir_siddon_fp_rec(_lib.ptr(self.rec), self.nrays, self.grid, _lib.ptr(vol),
                                                    _lib.ptr(out), fl, err, stream()), "vir_siddon_fp_rec")
        else:
            _lib.check(_lib.lib().vir_siddon_fp(_lib.ptr(self.src), _lib.ptr(self.trg), self.nrays,
                                                self.grid, _lib.ptr(vol), _lib.ptr(out), fl, err, stream()),
                       "vir_siddon_fp")
        if check and int(self._err.item()):
            raise ValueError("invalid entry in coordinates array")
        return out

    # -- adjoint ---------------------------------------------------------
    def bp(self, sino, out=None, accumulate=False):
        """vol = A^T sino (float32, grid shape)."""
        sino = as_device_f32(sino)
        if sino.numel() != self.nrays:
            raise ValueError("sinogram size does not match the ray set")
        out = out_tensor(out, self.nPixels, sino.device)
        fl = _lib.VIR_ACCUMULATE if accumulate else 0
        if self.proj is not None:
            fl |= self._ung
            nviews, nu, nv = self.layout
            if self.rec is not None:
                _lib.check(_lib.lib().vir_siddon_bp_rec(_lib.ptr(self.rec), nviews, nu, nv, _lib.ptr(self.proj),
                                                        self.grid, _lib.ptr(sino), _lib.ptr(out), fl,
                                                        _lib.ptr(self._ws), stream()), "vir_siddon_bp_rec")
                return out
            _lib.check(_lib.lib().vir_siddon_bp(_lib.ptr(self.src), _lib.ptr(self.trg), nviews, nu,
                                                nv, _lib.ptr(self.proj), self.grid, _lib.ptr(sino),
                                                _lib.ptr(out), fl, _lib.ptr(self._ws), stream()),
                       "vir_siddon_bp")
            return out
        colptr, row, clen = self.stored_csc()
        _lib.check(_lib.lib().vir_csc_bp(_lib.ptr(colptr), _lib.ptr(row), _lib.ptr(clen), self.nvox,
                                         _lib.ptr(sino), _lib.ptr(out), fl, stream()), "vir_csc_bp")
        return out

    # -- SIRT steps fused into the projections ----------------------------
    @property
    def fused_sirt(self):
        """True when fp_residual / bp_update are available (the brick-gather
        backprojector from per-ray records)."""
        return self.rec is not None and self.proj is not None

    def fp_residual(self, vol, b, w_sino, out):
        """out = (b - A vol) / w_sino where w_sino != 0, else A vol: the SIRT
        residual written by the forward projection (same bits as fp()
        followed by the separate residual step)."""
        _lib.check(_lib.lib().vir_siddon_fp_rec_res(_lib.ptr(self.rec), self.nrays, self.grid, _lib.ptr(vol),
                                                    _lib.ptr(b), _lib.ptr(w_sino), _lib.ptr(out), self._ung,
                                                    stream()), "vir_siddon_fp_rec_res")
        return out

    def bp_update(self, r, w_img, gamma, x):
        """x += gamma * (A^T r) / w_img (where w_img != 0, else A^T r), in
        place: the SIRT update written by the backprojection."""
        nviews, nu, nv = self.layout
        _lib.check(_lib.lib().vir_siddon_bp_rec_upd(_lib.ptr(self.rec), nviews, nu, nv, _lib.ptr(self.proj),
                                                    self.grid, _lib.ptr(r), _lib.ptr(w_img), float(gamma),
                                                    _lib.ptr(x), self._ung, _lib.ptr(self._ws), stream()),
                   "vir_siddon_bp_rec_upd")
        return x

    # -- batches of volumes on these rays ---------------------------------
    def _batch(self, t, per_item, what):
        t = as_device_f32(t)
        if t.numel() == 0 or t.numel() % per_item:
            raise ValueError(f"{what} size is not a whole number of items")
        return t, t.numel() // per_item

    def _batch_ws(self, nbatch, accumulate):
        """Staging for a batched backprojection (one group of sinograms)."""
        nviews = self.layout[0]
        need = int(_lib.lib().vir_siddon_bp_batch_workspace(self.grid, nviews, nbatch))
        if accumulate:
            need = max(need, min(nbatch, 2) * self.nvox * 4)
        if self._ws.numel() * 4 >= need:
            return self._ws
        if getattr(self, "_wsb", None) is None or self._wsb.numel() * 4 < need:
            self._wsb = torch.empty(need // 4 + 1, dtype=torch.float32, device=self.src.device)
        return self._wsb

    def _require_batch_path(self):
        if self.rec is None or self.proj is None:
            raise ValueError("batched projection needs the brick-gather layout with ray records")

    def fp_batch(self, vols, out=None, accumulate=False, check=False):
        """sino[j] = A vols[j] for a batch of volumes (B, *grid) on these rays.
        The rays are traced once per group of four volumes; every item equals
        fp(vols[j]) bit for bit."""
        self._require_batch_path()
        vols, nb = self._batch(vols, self.nvox, "volume")
        out = out_tensor(out, (nb,) + tuple(self.ray_shape), vols.device)
        if check:
            self._err.zero_()
        fl = (_lib.VIR_ACCUMULATE if accumulate else 0) | self._ung
        err = _lib.ptr(self._err) if check else None
        _lib.check(_lib.lib().vir_siddon_fp_rec_batch(_lib.ptr(self.rec), self.nrays, self.grid, nb,
                                                      _lib.ptr(vols), _lib.ptr(out), fl, err, stream()),
                   "vir_siddon_fp_rec_batch")
        if check and int(self._err.item()):
            raise ValueError("invalid entry in coordinates array")
        return out

    def bp_batch(self, sinos, out=None, accumulate=False):
        """vols[j] = A^T sinos[j] for a batch of sinograms (B, *ray_shape).
        Each candidate ray is traced once per group of two sinograms; every
        item equals bp(sinos[j]) bit for bit."""
        self._require_batch_path()
        sinos, nb = self._batch(sinos, self.nrays, "sinogram")
        out = out_tensor(out, (nb,) + tuple(self.nPixels), sinos.device)
        fl = (_lib.VIR_ACCUMULATE if accumulate else 0) | self._ung
        nviews, nu, nv = self.layout
        ws = self._batch_ws(nb, accumulate)
        _lib.check(_lib.lib().vir_siddon_bp_rec_batch(_lib.ptr(self.rec), nviews, nu, nv, _lib.ptr(self.proj),
                                                      self.grid, nb, _lib.ptr(sinos), _lib.ptr(out), fl,
                                                      _lib.ptr(ws), stream()), "vir_siddon_bp_rec_batch")
        return out

    def fp_residual_batch(self, vols, b, w_sino, out):
        """fp_residual for a batch: out[j] = (b[j] - A vols[j]) / w_sino."""
        self._require_batch_path()
        vols, nb = self._batch(vols, self.nvox, "volume")
        b, w_sino = as_device_f32(b), as_device_f32(w_sino)
        out = out_tensor(out, (nb,) + tuple(self.ray_shape), vols.device)
        if b.numel() != nb * self.nrays or w_sino.numel() != self.nrays:
            raise ValueError("b must hold one sinogram per volume, w_sino one sinogram")
        _lib.check(_lib.lib().vir_siddon_fp_rec_res_batch(_lib.ptr(self.rec), self.nrays, self.grid, nb,
                                                          _lib.ptr(vols), _lib.ptr(b), _lib.ptr(w_sino),
                                                          _lib.ptr(out), self._ung, stream()),
                   "vir_siddon_fp_rec_res_batch")
        return out

    def bp_update_batch(self, r, w_img, gamma, x):
        """bp_update for a batch: x[j] += gamma * (A^T r[j]) / w_img, in place."""
        self._require_batch_path()
        r, nb = self._batch(r, self.nrays, "sinogram")
        w_img = as_device_f32(w_img)
        x = out_tensor(x, (nb,) + tuple(self.nPixels), r.device, "x")
        if w_img.numel() != self.nvox:
            raise ValueError("w_img must hold one volume")
        nviews, nu, nv = self.layout
        ws = self._batch_ws(nb, False)
        _lib.check(_lib.lib().vir_siddon_bp_rec_upd_batch(_lib.ptr(self.rec), nviews, nu, nv,
                                                          _lib.ptr(self.proj), self.grid, nb, _lib.ptr(r),
                                                          _lib.ptr(w_img), float(gamma), _lib.ptr(x), self._ung,
                                                          _lib.ptr(ws), stream()), "vir_siddon_bp_rec_upd_batch")
        return x

    def stored_csr(self):
        """(off int64, ravel_idx int32, len f32, flags) device tensors."""
        off, flags, rav, ix, iy, iz, ln = _trace(self.src, self.trg, self.grid, self.nrays)
        return off, rav, ln, flags

    def stored_csc(self):
        if self._csc is None:
            off, rav, ln, flags = self.stored_csr()
            self._csc = csr_transpose(rav, ln, off, self.nvox)
        return self._csc


def csr_transpose(ind, ln, off, nvox):
    """Voxel-major copy (colptr int64, row int32, len f32) of a CSR matrix,
    entries of each voxel in CSR order (deterministic)."""
    nrays = off.numel() - 1
    nnz = ind.numel()
    dev = ind.device
    colptr = torch.empty(nvox + 1, dtype=torch.int64, device=dev)
    row = torch.empty(max(nnz, 1), dtype=torch.int32, device=dev)
    clen = torch.empty(max(nnz, 1), dtype=torch.float32, device=dev)
    _lib.check(_lib.lib().vir_csr_transpose(_lib.ptr(ind), _lib.ptr(ln), _lib.ptr(off), nrays, nnz,
                                            nvox, _lib.ptr(colptr), _lib.ptr(row), _lib.ptr(clen),
                                            stream()), "vir_csr_transpose")
    return colptr, row[:nnz], clen[:nnz]


# list utilities of sys_mat/siddon.py:12-633, same module as in the reference
from .siddon_lists import (list2array, list2array_iter_ray, list_ave, list_convole,  # noqa: E402,F401
                           list_count_elem, list_del_ind, list_flatten, list_flip,
                           list_index_shift, list_ravel, list_unravel, list_weight_ave, ray_boundary,
                           ray_flip, ray_index_shift, ray_ravel, ray_unravel, rays_ave,
                           rays_weight_ave)
