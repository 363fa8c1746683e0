"""Ray-driven projectors on the GPU: Amanatides-Woo path lengths and Joseph
interpolation.

Drop-ins for the reference's sys_mat/rd.py drivers, same arguments and
scaling (joseph=False: Amanatides-Woo, joseph=True: Joseph):

  aw_fp_par_2d(img, ang_arr, nu, du=1.0, su=0.0, d_pix=1.0, joseph=False)        rd.py:450-457
  aw_bp_par_2d(sino, ang_arr, img_shape, du=1.0, su=0.0, d_pix=1.0, joseph=False) rd.py:460-469
  aw_fp_fan_2d(img, ang_arr, nu, DSO, DSD, du=1.0, su=0.0, d_pix=1.0, joseph=False)    rd.py:544-552
  aw_bp_fan_2d(sino, ang_arr, img_shape, DSO, DSD, du=1.0, su=0.0, d_pix=1.0,
               joseph=False)                                                     rd.py:556-563
  aw_fp_cone_3d(img, ang_arr, nu, nv, DSO, DSD, du=1.0, dv=1.0, su=0.0, sv=0.0,
                d_pix=1.0, joseph=False)                                         rd.py:652-662
  aw_bp_cone_3d(sino, ang_arr, img_shape, DSO, DSD, du=1.0, dv=1.0, su=0.0, sv=0.0,
                d_pix=1.0, joseph=False)                                         rd.py:665-674

Forward projections are unscaled ray sums (the AW cone sum times
DSD/sqrt(DSD^2+u^2+v^2)); backprojections are the reference's splats scaled
by du/(na*d_pix^2) (2-D) or du*dv/na (cone).  The AW backprojector is the
transpose of the AW forward projector (cone: without the ray scale); the
Joseph pair is the reference's, which are transposes only away from the
image border (the 2-D forward clamps samples, the backward splats with
bounds checks; the 3-D forward truncates indices toward zero).

Rays parallel to an axis that lie outside the image miss here; the
reference clamps them onto the border row (rd.py:66-73), so at views along
an axis with a detector wider than the image its outer bins repeat the
border row's sums.  compat=True reproduces that (forward and backward).
Cone rows are centred with nu as in the reference (rd.py:691).

numpy in -> numpy out; CUDA tensors in -> CUDA tensors out.
"""
import numpy as np
import torch

from .. import _lib
from ..device import as_device_f32, out_tensor, require_gpu, stream

_MODES = {"par": 0, "fan": 1, "cone": 2}


class RDProjector:
    """Ray-driven forward projector and the reference's backprojector."""

    def __init__(self, img_shape, ang_arr, nu, geom="par", nv=1, DSO=0.0, DSD=0.0, du=1.0,
                 dv=1.0, su=0.0, sv=0.0, d_pix=1.0, joseph=False, compat=False):
        """compat=True: rays parallel to an axis outside the image are
        clamped onto the border row as in the reference (rd.py:66-73)."""
        require_gpu()
        self._flags = _lib.VIR_COMPAT if compat else 0
        if geom not in _MODES:
            raise ValueError(f"unknown geometry {geom!r}")
        ang = np.asarray(ang_arr, dtype=np.float32).reshape(-1)
        self.na = ang.size
        self.geom = geom
        self.joseph = bool(joseph)
        self.img_shape = tuple(int(v) for v in img_shape)
        if len(self.img_shape) != (3 if geom == "cone" else 2):
            raise ValueError(f"{geom} needs a {3 if geom == 'cone' else 2}-D image shape")
        # float32 cos/sin of float32 angles, as np.cos(ang_arr) after as_float32
        self.cos = torch.from_numpy(np.cos(ang)).cuda()
        self.sin = torch.from_numpy(np.sin(ang)).cuda()
        g = _lib.VirRDGeom()
        g.nx, g.ny = self.img_shape[0], self.img_shape[1]
        g.nz = self.img_shape[2] if geom == "cone" else 1
        g.nu, g.nv, g.na = int(nu), int(nv) if geom == "cone" else 1, self.na
        g.mode, g.joseph = _MODES[geom], int(self.joseph)
        for k, v in dict(d_pix=d_pix, du=du, dv=dv, su=su, sv=sv, DSO=DSO, DSD=DSD).items():
            setattr(g, k, float(np.float32(v)))
        self.g = g
        d, du32, dv32 = np.float32(d_pix), np.float32(du), np.float32(dv)
        self.sino_shape = (self.na, int(nu), int(nv)) if geom == "cone" else (self.na, int(nu))
        na = np.float32(max(self.na, 1))
        if geom == "cone":
            self.ref_bp_scale = float(du32 * dv32 / na)
        else:
            self.ref_bp_scale = float(du32 / (na * d * d))

    def _call(self, name, src, dst, scale, accumulate):
        fl = (_lib.VIR_ACCUMULATE if accumulate else 0) | self._flags
        _lib.check(getattr(_lib.lib(), name)(self.g, _lib.ptr(self.cos), _lib.ptr(self.sin),
                                             _lib.ptr(src), _lib.ptr(dst), scale, fl, stream()), name)

    def fp(self, img, out=None, accumulate=False):
        img = as_device_f32(img)
        if img.numel() != int(np.prod(self.img_shape)):
            raise ValueError("image size does not match")
        out = out_tensor(out, self.sino_shape, img.device)
        self._call("vir_rd_fp", img, out, 1.0, accumulate)
        return out

    def bp(self, sino, out=None, accumulate=False, reference_scale=False):
        """The backprojection splat; reference_scale=True applies the
        reference driver's scaling (rd.py:538-541, 646-649, 787-788)."""
        sino = as_device_f32(sino)
        if sino.numel() != int(np.prod(self.sino_shape)):
            raise ValueError("sinogram size does not match")
        out = out_tensor(out, self.img_shape, sino.device)
        self._call("vir_rd_bp", sino, out, self.ref_bp_scale if reference_scale else 1.0,
                   accumulate)
        return out


def _wrap(x, torch_in):
    return x if torch_in else x.cpu().numpy()


def aw_fp_par_2d(img, ang_arr, nu, du=1.0, su=0.0, d_pix=1.0, joseph=False, compat=False):
    torch_in = isinstance(img, torch.Tensor)
    P = RDProjector(tuple(img.shape), ang_arr, nu, "par", du=du, su=su, d_pix=d_pix,
                    joseph=joseph, compat=compat)
    return _wrap(P.fp(img), torch_in)


def aw_bp_par_2d(sino, ang_arr, img_shape, du=1.0, su=0.0, d_pix=1.0, joseph=False, compat=False):
    torch_in = isinstance(sino, torch.Tensor)
    P = RDProjector(img_shape, ang_arr, sino.shape[1], "par", du=du, su=su, d_pix=d_pix,
                    joseph=joseph, compat=compat)
    return _wrap(P.bp(sino, reference_scale=True), torch_in)


def aw_fp_fan_2d(img, ang_arr, nu, DSO, DSD, du=1.0, su=0.0, d_pix=1.0, joseph=False, compat=False):
    torch_in = isinstance(img, torch.Tensor)
    P = RDProjector(tuple(img.shape), ang_arr, nu, "fan", DSO=DSO, DSD=DSD, du=du, su=su,
                    d_pix=d_pix, joseph=joseph, compat=compat)
    return _wrap(P.fp(img), torch_in)


def aw_bp_fan_2d(sino, ang_arr, img_shape, DSO, DSD, du=1.0, su=0.0, d_pix=1.0, joseph=False,
                 compat=False):
    torch_in = isinstance(sino, torch.Tensor)
    P = RDProjector(img_shape, ang_arr, sino.shape[1], "fan", DSO=DSO, DSD=DSD, du=du, su=su,
                    d_pix=d_pix, joseph=joseph, compat=compat)
    return _wrap(P.bp(sino, reference_scale=True), torch_in)


def aw_fp_cone_3d(img, ang_arr, nu, nv, DSO, DSD, du=1.0, dv=1.0, su=0.0, sv=0.0, d_pix=1.0,
                  joseph=False, compat=False):
    torch_in = isinstance(img, torch.Tensor)
    P = RDProjector(tuple(img.shape), ang_arr, nu, "cone", nv=nv, DSO=DSO, DSD=DSD, du=du, dv=dv,
                    su=su, sv=sv, d_pix=d_pix, joseph=joseph, compat=compat)
    return _wrap(P.fp(img), torch_in)


def aw_bp_cone_3d(sino, ang_arr, img_shape, DSO, DSD, du=1.0, dv=1.0, su=0.0, sv=0.0, d_pix=1.0,
                  joseph=False, compat=False):
    torch_in = isinstance(sino, torch.Tensor)
    P = RDProjector(img_shape, ang_arr, sino.shape[1], "cone", nv=sino.shape[2], DSO=DSO,
                    DSD=DSD, du=du, dv=dv, su=su, sv=sv, d_pix=d_pix, joseph=joseph, compat=compat)
    return _wrap(P.bp(sino, reference_scale=True), torch_in)
