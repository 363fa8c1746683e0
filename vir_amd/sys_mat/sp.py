"""Separable-footprint prototypes of the reference's sys_mat/sp.py on the GPU.

  sf_fp_fan_2d(img, angs, nu, DSO, DSD, du=1.0, d_pix=1.0)          sp.py:12-87
      fan beam, triangle (linear B-spline) footprint of half-width
      h = DSD d_pix / (DSO - yp) around each pixel centre's projection,
      sampled at the bin centres, output times d_pix
  sf_bp_fan_2d(sino, angs, img_shape, DSO, DSD, du=1.0, d_pix=1.0)
      its exact adjoint (no reference counterpart)
  footprint1D(r), CTGeometry, forward_project(volume, geom),
  back_project(proj, geom)                                          sp.py:93-211
      the separable toy model: proj[v, u] = sum vol f(u/Dd - x/Ds)
      f(v/Dd - y/Ds) f(z/Ds), f(r) = max(0, 1 - |r|), and its transpose

Kernels: vir_sp_fan_fp / vir_sp_fan_bp (a thread per bin walking the lines
of the grid, a thread per pixel walking the views) and vir_sep_fp /
vir_sep_bp (three contractions in double).  The reference sums in float32
per pixel (fan) or in float64 per detector pixel (toy); results agree to
float32 rounding.  (The reference module also runs a 128^3 example at import;
this one does not.)

numpy in -> numpy out; CUDA tensors in -> CUDA tensors out.
"""
import numpy as np
import torch

from .. import _lib
from ..device import as_device_f32, out_tensor, require_gpu, stream


class SPFanProjector:
    """Triangle-footprint fan-beam forward projector and its exact adjoint,
    both scaled by d_pix as sf_fp_fan_2d's output (sp.py:87)."""

    def __init__(self, img_shape, angs, nu, DSO, DSD, du=1.0, d_pix=1.0):
        require_gpu()
        ang = np.asarray(angs, np.float64).reshape(-1)
        self.img_shape = tuple(int(v) for v in img_shape)
        if len(self.img_shape) != 2:
            raise ValueError("sf_fp_fan_2d needs a 2-D image")
        self.na, self.nu = ang.size, int(nu)
        self.sino_shape = (self.na, self.nu)
        # float32 of the float64 np.cos / np.sin the reference evaluates (sp.py:52)
        self.cos = torch.from_numpy(np.cos(ang).astype(np.float32)).cuda()
        self.sin = torch.from_numpy(np.sin(ang).astype(np.float32)).cuda()
        g = _lib.VirSpGeom()
        g.nx, g.ny = self.img_shape
        g.nu, g.na = self.nu, self.na
        g.d_pix, g.du, g.DSO, g.DSD = (float(np.float32(v)) for v in (d_pix, du, DSO, DSD))
        self.g = g
        self.scale = float(np.float32(d_pix))

    def _run(self, name, src, dst, accumulate):
        _lib.check(getattr(_lib.lib(), name)(self.g, _lib.ptr(self.cos), _lib.ptr(self.sin), _lib.ptr(src),
                                             _lib.ptr(dst), self.scale,
                                             _lib.VIR_ACCUMULATE if accumulate else 0, stream()), name)
        return dst

    def fp(self, img, out=None, accumulate=False):
        img = as_device_f32(img)
        if img.numel() != int(np.prod(self.img_shape)):
            raise ValueError("image size does not match")
        return self._run("vir_sp_fan_fp", img, out_tensor(out, self.sino_shape, img.device), accumulate)

    def bp(self, sino, out=None, accumulate=False):
        sino = as_device_f32(sino)
        if sino.numel() != self.na * self.nu:
            raise ValueError("sinogram size does not match")
        return self._run("vir_sp_fan_bp", sino, out_tensor(out, self.img_shape, sino.device), accumulate)


def _wrap(x, torch_in):
    return x if torch_in else x.cpu().numpy()


def sf_fp_fan_2d(img, angs, nu, DSO, DSD, du=1.0, d_pix=1.0):
    torch_in = isinstance(img, torch.Tensor)
    P = SPFanProjector(tuple(img.shape), angs, nu, DSO, DSD, du=du, d_pix=d_pix)
    return _wrap(P.fp(img), torch_in)


def sf_bp_fan_2d(sino, angs, img_shape, DSO, DSD, du=1.0, d_pix=1.0):
    torch_in = isinstance(sino, torch.Tensor)
    P = SPFanProjector(img_shape, angs, sino.shape[1], DSO, DSD, du=du, d_pix=d_pix)
    return _wrap(P.bp(sino), torch_in)


# ---------------------------------------------------------- the toy model

def footprint1D(r):
    """max(0, 1 - |r|) (sp.py:122-128)."""
    return np.maximum(1 - np.abs(r), 0)


class CTGeometry:
    """Geometry of the separable toy model (sp.py:93-115): volume and
    detector grids, Ds = source-isocentre, Dd = source-detector distance."""

    def __init__(self, vol_shape, voxel_size, det_shape, det_spacing, source_to_iso, iso_to_det):
        self.nx, self.ny, self.nz = (int(v) for v in vol_shape)
        self.dx, self.dy, self.dz = voxel_size
        self.nu, self.nv = (int(v) for v in det_shape)
        self.su, self.sv = det_spacing
        self.Ds = source_to_iso
        self.Dd = source_to_iso + iso_to_det

        def centres(n, d):
            return (np.arange(n) - n / 2 + 0.5) * d
        self.xs, self.ys, self.zs = centres(self.nx, self.dx), centres(self.ny, self.dy), centres(self.nz, self.dz)
        self.u, self.v = centres(self.nu, self.su), centres(self.nv, self.sv)

    def c_geom(self):
        g = _lib.VirSepGeom()
        g.nx, g.ny, g.nz, g.nu, g.nv = self.nx, self.ny, self.nz, self.nu, self.nv
        g.dx, g.dy, g.dz, g.su, g.sv = (float(v) for v in (self.dx, self.dy, self.dz, self.su, self.sv))
        g.Ds, g.Dd = float(self.Ds), float(self.Dd)
        return g


def _sep(name, src, geom, out_shape):
    require_gpu()
    torch_in = isinstance(src, torch.Tensor)
    x = as_device_f32(src)
    g = geom.c_geom()
    L = _lib.lib()
    work = torch.empty(max(int(L.vir_sep_workspace(g)), 8) // 8, dtype=torch.float64, device=x.device)
    out = torch.empty(out_shape, dtype=torch.float32, device=x.device)
    _lib.check(getattr(L, name)(g, _lib.ptr(x), _lib.ptr(out), _lib.ptr(work), stream()), name)
    return _wrap(out, torch_in)


def forward_project(volume, geom):
    """(nv, nu) float32 projection of a (nx, ny, nz) volume (sp.py:135-171)."""
    if int(np.prod(volume.shape)) != geom.nx * geom.ny * geom.nz:
        raise ValueError("volume does not match the geometry")
    return _sep("vir_sep_fp", volume, geom, (geom.nv, geom.nu))


def back_project(proj, geom):
    """(nx, ny, nz) float32 transpose of forward_project (sp.py:177-211)."""
    if int(np.prod(proj.shape)) != geom.nv * geom.nu:
        raise ValueError("projection does not match the geometry")
    return _sep("vir_sep_bp", proj, geom, (geom.nx, geom.ny, geom.nz))
