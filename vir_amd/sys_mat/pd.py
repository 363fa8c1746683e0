"""Pixel-driven separable-footprint projectors on the GPU.

Drop-ins for the reference's sys_mat/pd.py drivers, same arguments and
output scaling:

  pd_fp_par_2d(img, ang_arr, nu, du=1.0, su=0.0, d_pix=1.0)                 pd.py:129-134
  pd_bp_par_2d(sino, ang_arr, img_shape, du=1.0, su=0.0, d_pix=1.0)         pd.py:137-144
  pd_fp_fan_2d(img, ang_arr, nu, DSO, DSD, du=1.0, su=0.0, d_pix=1.0)       pd.py:237-242
  pd_bp_fan_2d(sino, ang_arr, img_shape, DSO, DSD, du=1.0, su=0.0, d_pix=1.0)  pd.py:245-252
  pd_fp_cone_3d(img, ang_arr, nu, nv, DSO, DSD, du=1.0, dv=1.0, su=0.0, sv=0.0,
                d_pix=1.0)                                                   pd.py:353-358
  pd_bp_cone_3d(sino, ang_arr, img_shape, DSO, DSD, du=1.0, dv=1.0, su=0.0,
                sv=0.0, d_pix=1.0)                                           pd.py:362-369

Forward projections are d_pix * W img; the reference's backprojections are
du/(na*d_pix) * W^T sino (2-D) and du*dv/(na*d_pix) * W^T sino (cone).

SFProjector is the device-resident operator pair (fp = A, bp = A^T exactly)
with a batch dimension, and adds geometry "par3d": 3-D parallel beam with
the detector rows along z (each z-slice projected like pd_p_par_2d,
weighted by the overlap of the slab with the rows of height dv).

The fan forward pass evaluates every pixel's footprint from its own
boundaries.  The reference's loop skips zero pixels without advancing one
of the boundaries it uses for the next pixel's amplitude (pd.py:321-325),
so its results differ on images with zero runs in a column; for strictly
nonzero images the two agree.  pd_fp_fan_2d(..., compat=True) (and
SFProjector(geom="fan", compat=True).fp) reproduces the reference's loop.

numpy in -> numpy out; CUDA tensors in -> CUDA tensors out.
"""
import numpy as np
import torch

from .. import _lib
from ..device import as_device_f32, out_tensor, require_gpu, stream

_MODES = {"par": 0, "fan": 1, "cone": 2, "par3d": 3}


class SFProjector:
    """Separable-footprint forward projector A and its exact adjoint A^T.

    img_shape: (nx, ny) for "par"/"fan", (nx, ny, nz) for "cone"/"par3d".
    batch: number of volumes projected per call (leading dimension when
    > 1).  Angles in radians; cos/sin are rounded to float32."""

    def __init__(self, img_shape, ang_arr, nu, geom="par", nv=1, DSO=0.0, DSD=0.0, du=1.0,
                 dv=1.0, su=0.0, sv=0.0, d_pix=1.0, batch=1, unit_amplitude=False,
                 compat=False):
        """unit_amplitude (parallel geometries): footprint amplitude 1 instead
        of 1/(|sin| + |cos|), the rectangle footprint of fp_2d_par.py:67-125.
        compat (fan): the forward pass takes the amplitude boundary after a
        run of zero pixels as the reference's loop does (pd.py:321-325); the
        backward pass is unchanged (the reference's has no such skip)."""
        require_gpu()
        if unit_amplitude and geom not in ("par", "par3d"):
            raise ValueError("unit_amplitude applies to the parallel geometries")
        if compat and geom != "fan":
            raise ValueError("compat applies to the fan geometry")
        self._flags = _lib.VIR_SF_UNIT_AMPLITUDE if unit_amplitude else 0
        self._fp_flags = _lib.VIR_COMPAT if compat else 0
        if geom not in _MODES:
            raise ValueError(f"unknown geometry {geom!r}")
        ang64 = np.asarray(ang_arr, dtype=np.float64).reshape(-1)
        self.na = ang64.size
        self.geom = geom
        self.is3 = geom in ("cone", "par3d")
        self.img_shape = tuple(int(v) for v in img_shape)
        if len(self.img_shape) != (3 if self.is3 else 2):
            raise ValueError(f"{geom} needs a {3 if self.is3 else 2}-D image shape")
        self.batch = int(batch)
        self.ang = torch.from_numpy(ang64.astype(np.float32)).cuda()
        self.cos = torch.from_numpy(np.cos(ang64).astype(np.float32)).cuda()
        self.sin = torch.from_numpy(np.sin(ang64).astype(np.float32)).cuda()
        g = _lib.VirSFGeom()
        g.nx, g.ny = self.img_shape[0], self.img_shape[1]
        g.nz = self.img_shape[2] if self.is3 else 1
        g.nu, g.nv, g.na = int(nu), int(nv) if self.is3 else 1, self.na
        g.nbatch, g.mode = self.batch, _MODES[geom]
        for k, v in dict(d_pix=d_pix, du=du, dv=dv, su=su, sv=sv, DSO=DSO, DSD=DSD).items():
            setattr(g, k, float(np.float32(v)))
        self.g = g
        self.d_pix, self.du, self.dv = np.float32(d_pix), np.float32(du), np.float32(dv)
        det = (self.na, int(nu), int(nv)) if self.is3 else (self.na, int(nu))
        lead = (self.batch,) if self.batch > 1 else ()
        self.sino_shape = lead + det
        self.vol_shape = lead + self.img_shape
        self.fp_scale = float(self.d_pix)
        s = np.float32(self.du) / (np.float32(max(self.na, 1)) * self.d_pix)
        self.ref_bp_scale = float(s * self.dv if geom == "cone" else s)

    def _call(self, name, src, dst, scale, accumulate):
        fl = (_lib.VIR_ACCUMULATE if accumulate else 0) | self._flags
        if name == "vir_sf_fp":
            fl |= self._fp_flags
        _lib.check(getattr(_lib.lib(), name)(self.g, _lib.ptr(self.ang), _lib.ptr(self.cos),
                                             _lib.ptr(self.sin), _lib.ptr(src), _lib.ptr(dst),
                                             scale, fl, stream()), name)

    def fp(self, img, out=None, accumulate=False):
        img = as_device_f32(img)
        if img.numel() != int(np.prod(self.vol_shape)):
            raise ValueError("image size does not match")
        out = out_tensor(out, self.sino_shape, img.device)
        self._call("vir_sf_fp", img, out, self.fp_scale, accumulate)
        return out

    def bp(self, sino, out=None, accumulate=False, reference_scale=False):
        """A^T sino; reference_scale=True gives the reference's scaled
        backprojection (pd.py:231-232, 347-348, 482-483)."""
        sino = as_device_f32(sino)
        if sino.numel() != int(np.prod(self.sino_shape)):
            raise ValueError("sinogram size does not match")
        out = out_tensor(out, self.vol_shape, sino.device)
        scale = self.ref_bp_scale if reference_scale else self.fp_scale
        self._call("vir_sf_bp", sino, out, scale, accumulate)
        return out


def _wrap(x, torch_in):
    return x if torch_in else x.cpu().numpy()


def pd_fp_par_2d(img, ang_arr, nu, du=1.0, su=0.0, d_pix=1.0):
    torch_in = isinstance(img, torch.Tensor)
    P = SFProjector(tuple(img.shape), ang_arr, nu, "par", du=du, su=su, d_pix=d_pix)
    return _wrap(P.fp(img), torch_in)


def pd_bp_par_2d(sino, ang_arr, img_shape, du=1.0, su=0.0, d_pix=1.0):
    torch_in = isinstance(sino, torch.Tensor)
    P = SFProjector(img_shape, ang_arr, sino.shape[1], "par", du=du, su=su, d_pix=d_pix)
    return _wrap(P.bp(sino, reference_scale=True), torch_in)


def pd_fp_fan_2d(img, ang_arr, nu, DSO, DSD, du=1.0, su=0.0, d_pix=1.0, compat=False):
    torch_in = isinstance(img, torch.Tensor)
    P = SFProjector(tuple(img.shape), ang_arr, nu, "fan", DSO=DSO, DSD=DSD, du=du, su=su,
                    d_pix=d_pix, compat=compat)
    return _wrap(P.fp(img), torch_in)


def pd_bp_fan_2d(sino, ang_arr, img_shape, DSO, DSD, du=1.0, su=0.0, d_pix=1.0):
    torch_in = isinstance(sino, torch.Tensor)
    P = SFProjector(img_shape, ang_arr, sino.shape[1], "fan", DSO=DSO, DSD=DSD, du=du, su=su,
                    d_pix=d_pix)
    return _wrap(P.bp(sino, reference_scale=True), torch_in)


def pd_fp_cone_3d(img, ang_arr, nu, nv, DSO, DSD, du=1.0, dv=1.0, su=0.0, sv=0.0, d_pix=1.0):
    torch_in = isinstance(img, torch.Tensor)
    P = SFProjector(tuple(img.shape), ang_arr, nu, "cone", nv=nv, DSO=DSO, DSD=DSD, du=du, dv=dv,
                    su=su, sv=sv, d_pix=d_pix)
    return _wrap(P.fp(img), torch_in)


def pd_bp_cone_3d(sino, ang_arr, img_shape, DSO, DSD, du=1.0, dv=1.0, su=0.0, sv=0.0, d_pix=1.0):
    torch_in = isinstance(sino, torch.Tensor)
    P = SFProjector(img_shape, ang_arr, sino.shape[1], "cone", nv=sino.shape[2], DSO=DSO,
                    DSD=DSD, du=du, dv=dv, su=su, sv=sv, d_pix=d_pix)
    return _wrap(P.bp(sino, reference_scale=True), torch_in)
