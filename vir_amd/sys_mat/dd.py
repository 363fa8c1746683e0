"""Distance-driven projectors on the GPU.

Drop-ins for the reference's sys_mat/dd.py drivers, same arguments and
output scaling:

  dd_fp_par_2d(img, ang_arr, nu, du=1.0, su=0.0, d_pix=1.0, sample=1)     dd.py:651-661
  dd_bp_par_2d(sino, ang_arr, img_shape, du=1.0, su=0.0, d_pix=1.0)       dd.py:666-672
  dd_fp_fan_2d(img, ang_arr, nu, DSO, DSD, du=1.0, su=0.0, d_pix=1.0)     dd.py:730-736
  dd_bp_fan_2d(sino, ang_arr, img_shape, DSO, DSD, du=1.0, su=0.0, d_pix=1.0)   dd.py:739-745
  dd_fp_cone_3d(img, ang_arr, nu, nv, DSO, DSD, du=1.0, dv=1.0, su=0.0, sv=1.0,
                d_pix=1.0, sample=1)                                       dd.py:815-825
  dd_bp_cone_3d(sino, ang_arr, img_shape, DSO, DSD, du=1.0, dv=1.0, su=0.0,
                sv=0.0, d_pix=1.0)                                         dd.py:829-835
  dd_fp_square(img, nu, nv, nsrc_p, nsrc_z, DSO, DSD, du=1.0, dv=1.0, dsrc_p=1.0,
               dsrc_z=1.0, su=0.0, sv=0.0, ssrc_p=0.0, ssrc_z=0.0, d_pix=1.0)  dd.py:903-930
  dd_bp_square(sino, img_shape, DSO, DSD, du=1.0, dv=1.0, dsrc_p=1.0, dsrc_z=1.0,
               su=0.0, sv=0.0, ssrc_p=0.0, ssrc_z=0.0, d_pix=1.0)          dd.py:1104-1135

The reference's backprojectors are A^T scaled by du/(na*d_pix^2) (2-D) and
du*dv/(na*d_pix^3) (cone); these functions keep that scaling.  DDProjector is
the device-resident operator pair with the exact adjoint (bp = A^T), plus an
equiangular-detector fan variant (detector="arc").

numpy in -> numpy out; CUDA tensors in -> CUDA tensors out.
"""
import numpy as np
import torch

from .. import _lib
from ..device import as_device_f32, out_tensor, require_gpu, stream

_MODES = {"par": 0, "flat": 1, "arc": 2}


def as_float32(*args):
    """Scalar -> np.float32, array -> float32 array (dd.py:29-36)."""
    out = []
    for x in args:
        if isinstance(x, torch.Tensor):
            out.append(x)
        elif np.isscalar(x):
            out.append(np.float32(x))
        else:
            out.append(np.asarray(x, dtype=np.float32))
    return out


class DDProjector:
    """Distance-driven forward projector A and its exact adjoint A^T.

    geom: "par", "fan" or "cone"; detector: "flat" (the reference's) or
    "arc" (equiangular fan, du = arc length at radius DSD).  Angles in
    radians; cos/sin are evaluated in float32 like the reference
    (np.cos of a float32 array)."""

    def __init__(self, img_shape, ang_arr, nu, geom="par", nv=1, DSO=0.0, DSD=0.0, du=1.0,
                 dv=1.0, su=0.0, sv=0.0, d_pix=1.0, detector="flat", compat=False):
        """compat (cone): the forward pass re-zeroes its u buffer only over
        each line's footprint, as the reference's sweep does (dd.py:317, 374);
        the backward pass is unchanged."""
        require_gpu()
        if compat and geom != "cone":
            raise ValueError("compat applies to the cone geometry")
        self._fp_flags = _lib.VIR_COMPAT if compat else 0
        ang = np.asarray(ang_arr, dtype=np.float32).reshape(-1)
        self.na = ang.size
        self.geom = geom
        self.img_shape = tuple(int(v) for v in img_shape)
        self.cos = torch.from_numpy(np.cos(ang).astype(np.float32)).cuda()
        self.sin = torch.from_numpy(np.sin(ang).astype(np.float32)).cuda()
        g = _lib.VirDDGeom()
        g.nx, g.ny = self.img_shape[0], self.img_shape[1]
        g.nz = self.img_shape[2] if geom == "cone" else 1
        g.nu, g.nv, g.na = int(nu), int(nv) if geom == "cone" else 1, self.na
        if geom == "par":
            g.mode = 0
        elif geom in ("fan", "cone"):
            g.mode = _MODES[detector] if geom == "fan" else 1
        else:
            raise ValueError(f"unknown geometry {geom!r}")
        for k, v in dict(d_pix=d_pix, du=du, dv=dv, su=su, sv=sv, DSO=DSO, DSD=DSD).items():
            setattr(g, k, float(np.float32(v)))
        self.g = g
        self.d_pix, self.du, self.dv = np.float32(d_pix), np.float32(du), np.float32(dv)
        if geom == "cone":
            self.sino_shape = (self.na, int(nu), int(nv))
        else:
            self.sino_shape = (self.na, int(nu))
        # A includes the reference's forward scaling
        if geom == "cone":
            self.fp_scale = float(self.d_pix * self.d_pix / (self.du * self.dv))
            self.ref_bp_scale = float(np.float32(1.0) / (np.float32(self.na) * self.d_pix))
        else:
            self.fp_scale = float(self.d_pix / self.du)
            self.ref_bp_scale = float(np.float32(1.0) / (np.float32(self.na) * self.d_pix))

    def _call(self, name, src, dst, scale, accumulate):
        fl = _lib.VIR_ACCUMULATE if accumulate else 0
        if name == "vir_ddcone_fp":
            fl |= self._fp_flags
        _lib.check(getattr(_lib.lib(), name)(self.g, _lib.ptr(self.cos), _lib.ptr(self.sin),
                                             _lib.ptr(src), _lib.ptr(dst), scale, fl, stream()), name)

    def fp(self, img, out=None, accumulate=False):
        img = as_device_f32(img)
        if img.numel() != int(np.prod(self.img_shape)):
            raise ValueError("image size does not match")
        out = out_tensor(out, self.sino_shape, img.device)
        self._call("vir_ddcone_fp" if self.geom == "cone" else "vir_dd2d_fp", img, out,
                   self.fp_scale, accumulate)
        return out

    def bp(self, sino, out=None, accumulate=False, reference_scale=False):
        """A^T sino; reference_scale=True returns the reference's scaled
        backprojection (dd.py:721-722, 799-800, 886-887)."""
        sino = as_device_f32(sino)
        if sino.numel() != int(np.prod(self.sino_shape)):
            raise ValueError("sinogram size does not match")
        out = out_tensor(out, self.img_shape, sino.device)
        scale = self.ref_bp_scale if reference_scale else self.fp_scale
        self._call("vir_ddcone_bp" if self.geom == "cone" else "vir_dd2d_bp", sino, out, scale,
                   accumulate)
        return out


    # -- SIRT steps fused into the projections (2-D geometries) ------------
    @property
    def fused_sirt(self):
        return self.geom != "cone"

    def fp_residual(self, img, b, w_sino, out):
        """out = (b - A img) / w_sino where w_sino != 0, else A img (the SIRT
        residual written by the forward projection; same bits as fp()
        followed by the separate residual step)."""
        _lib.check(_lib.lib().vir_dd2d_fp_res(self.g, _lib.ptr(self.cos), _lib.ptr(self.sin), _lib.ptr(img),
                                              _lib.ptr(b), _lib.ptr(w_sino), _lib.ptr(out), self.fp_scale, 0,
                                              stream()), "vir_dd2d_fp_res")
        return out

    def bp_update(self, r, w_img, gamma, x):
        """x += gamma * (A^T r) / w_img (where w_img != 0, else A^T r), in
        place: the SIRT update written by the backprojection."""
        _lib.check(_lib.lib().vir_dd2d_bp_upd(self.g, _lib.ptr(self.cos), _lib.ptr(self.sin), _lib.ptr(r),
                                              _lib.ptr(w_img), float(gamma), _lib.ptr(x), self.fp_scale, 0,
                                              stream()), "vir_dd2d_bp_upd")
        return x


def _wrap(x, torch_in):
    return x if torch_in else x.cpu().numpy()


def _kron_sample(img, sample):
    if sample == 1:
        return img
    if isinstance(img, torch.Tensor):
        img = img.cpu().numpy()
    return np.kron(img, np.ones([sample] * img.ndim)).astype(np.float32)


def dd_fp_par_2d(img, ang_arr, nu, du=1.0, su=0.0, d_pix=1.0, sample=1):
    torch_in = isinstance(img, torch.Tensor)
    img, ang_arr, du, su, d_pix = as_float32(img, ang_arr, du, su, d_pix)
    P = DDProjector(tuple(img.shape), ang_arr, nu, "par", du=du, su=su, d_pix=d_pix)
    return _wrap(P.fp(img), torch_in)


def dd_bp_par_2d(sino, ang_arr, img_shape, du=1.0, su=0.0, d_pix=1.0):
    torch_in = isinstance(sino, torch.Tensor)
    sino, ang_arr, du, su, d_pix = as_float32(sino, ang_arr, du, su, d_pix)
    P = DDProjector(img_shape, ang_arr, sino.shape[1], "par", du=du, su=su, d_pix=d_pix)
    return _wrap(P.bp(sino, reference_scale=True), torch_in)


def dd_fp_fan_2d(img, ang_arr, nu, DSO, DSD, du=1.0, su=0.0, d_pix=1.0, detector="flat"):
    torch_in = isinstance(img, torch.Tensor)
    img, ang_arr, DSO, DSD, du, su, d_pix = as_float32(img, ang_arr, DSO, DSD, du, su, d_pix)
    P = DDProjector(tuple(img.shape), ang_arr, nu, "fan", DSO=DSO, DSD=DSD, du=du, su=su,
                    d_pix=d_pix, detector=detector)
    return _wrap(P.fp(img), torch_in)


def dd_bp_fan_2d(sino, ang_arr, img_shape, DSO, DSD, du=1.0, su=0.0, d_pix=1.0, detector="flat"):
    torch_in = isinstance(sino, torch.Tensor)
    sino, ang_arr, DSO, DSD, du, su, d_pix = as_float32(sino, ang_arr, DSO, DSD, du, su, d_pix)
    P = DDProjector(img_shape, ang_arr, sino.shape[1], "fan", DSO=DSO, DSD=DSD, du=du, su=su,
                    d_pix=d_pix, detector=detector)
    return _wrap(P.bp(sino, reference_scale=True), torch_in)


def dd_fp_cone_3d(img, ang_arr, nu, nv, DSO, DSD, du=1.0, dv=1.0, su=0.0, sv=1.0, d_pix=1.0,
                  sample=1, compat=False):
    """Cone DD forward projection.  The u buffer is cleared fully per slab;
    the reference clears only the current footprint (dd.py:317) and adds the
    whole buffer (dd.py:374), which carries leftovers of earlier lines into
    bins outside the current footprint when the object touches the volume
    border (identical results otherwise).  compat=True reproduces the
    reference."""
    torch_in = isinstance(img, torch.Tensor)
    img, ang_arr, DSO, DSD, du, dv, su, sv, d_pix = as_float32(img, ang_arr, DSO, DSD, du, dv, su,
                                                               sv, d_pix)
    if sample != 1:
        img = _kron_sample(img, sample)
        d_pix = np.float32(d_pix / sample)
    P = DDProjector(tuple(img.shape), ang_arr, nu, "cone", nv=nv, DSO=DSO, DSD=DSD, du=du, dv=dv,
                    su=su, sv=sv, d_pix=d_pix, compat=compat)
    return _wrap(P.fp(img), torch_in)


def dd_bp_cone_3d(sino, ang_arr, img_shape, DSO, DSD, du=1.0, dv=1.0, su=0.0, sv=0.0, d_pix=1.0):
    torch_in = isinstance(sino, torch.Tensor)
    sino, ang_arr, DSO, DSD, du, dv, su, sv, d_pix = as_float32(sino, ang_arr, DSO, DSD, du, dv,
                                                                su, sv, d_pix)
    P = DDProjector(img_shape, ang_arr, sino.shape[1], "cone", nv=sino.shape[2], DSO=DSO, DSD=DSD,
                    du=du, dv=dv, su=su, sv=sv, d_pix=d_pix)
    return _wrap(P.bp(sino, reference_scale=True), torch_in)


class SquareDDProjector:
    """Distance-driven projector of the translational four-sided geometry
    (dd.py:903-1313): fp = the reference's dd_fp_square (scaled by
    d_pix^2/(du*dv)), bp = the reference's dd_bp_square (unscaled), which is
    the transpose of the unscaled forward sum.  Volume (n, n, nz); sinogram
    (nsrc_p, nsrc_z, nv, nu, 4)."""

    def __init__(self, img_shape, nu, nv, nsrc_p, nsrc_z, DSO, DSD, du=1.0, dv=1.0, dsrc_p=1.0,
                 dsrc_z=1.0, su=0.0, sv=0.0, ssrc_p=0.0, ssrc_z=0.0, d_pix=1.0):
        require_gpu()
        nx, ny, nz = (int(v) for v in img_shape)
        if nx != ny:
            raise ValueError("nx must equal ny")
        self.img_shape = (nx, ny, nz)
        self.sino_shape = (int(nsrc_p), int(nsrc_z), int(nv), int(nu), 4)
        g = _lib.VirDDSqGeom()
        g.n, g.nz, g.nu, g.nv, g.nsrc_p, g.nsrc_z = nx, nz, int(nu), int(nv), int(nsrc_p), int(nsrc_z)
        for k, v in dict(d_pix=d_pix, du=du, dv=dv, dsrc_p=dsrc_p, dsrc_z=dsrc_z, su=su, sv=sv,
                         ssrc_p=ssrc_p, ssrc_z=ssrc_z, DSO=DSO, DSD=DSD).items():
            setattr(g, k, float(np.float32(v)))
        self.g = g
        d, du32, dv32 = np.float32(d_pix), np.float32(du), np.float32(dv)
        self.fp_scale = float(d * d / (du32 * dv32))

    def fp(self, img, out=None, accumulate=False):
        img = as_device_f32(img)
        if img.numel() != int(np.prod(self.img_shape)):
            raise ValueError("image size does not match")
        out = out_tensor(out, self.sino_shape, img.device)
        fl = _lib.VIR_ACCUMULATE if accumulate else 0
        _lib.check(_lib.lib().vir_ddsq_fp(self.g, _lib.ptr(img), _lib.ptr(out), self.fp_scale, fl,
                                          stream()), "vir_ddsq_fp")
        return out

    def bp(self, sino, out=None, accumulate=False):
        sino = as_device_f32(sino)
        if sino.numel() != int(np.prod(self.sino_shape)):
            raise ValueError("sinogram size does not match")
        out = out_tensor(out, self.img_shape, sino.device)
        fl = _lib.VIR_ACCUMULATE if accumulate else 0
        _lib.check(_lib.lib().vir_ddsq_bp(self.g, _lib.ptr(sino), _lib.ptr(out), 1.0, fl, stream()),
                   "vir_ddsq_bp")
        return out


def dd_fp_square(img, nu, nv, nsrc_p, nsrc_z, DSO, DSD, du=1.0, dv=1.0, dsrc_p=1.0, dsrc_z=1.0,
                 su=0.0, sv=0.0, ssrc_p=0.0, ssrc_z=0.0, d_pix=1.0):
    """dd.py:903-930: sinogram (nsrc_p, nsrc_z, nv, nu, 4) * d_pix^2/(du*dv)."""
    torch_in = isinstance(img, torch.Tensor)
    P = SquareDDProjector(tuple(img.shape), nu, nv, nsrc_p, nsrc_z, DSO, DSD, du, dv, dsrc_p,
                          dsrc_z, su, sv, ssrc_p, ssrc_z, d_pix)
    return _wrap(P.fp(img), torch_in)


def dd_bp_square(sino, img_shape, DSO, DSD, du=1.0, dv=1.0, dsrc_p=1.0, dsrc_z=1.0, su=0.0, sv=0.0,
                 ssrc_p=0.0, ssrc_z=0.0, d_pix=1.0):
    """dd.py:1104-1135: the unscaled transpose."""
    torch_in = isinstance(sino, torch.Tensor)
    nsrc_p, nsrc_z, nv, nu = (int(v) for v in sino.shape[:4])
    P = SquareDDProjector(img_shape, nu, nv, nsrc_p, nsrc_z, DSO, DSD, du, dv, dsrc_p, dsrc_z, su,
                          sv, ssrc_p, ssrc_z, d_pix)
    return _wrap(P.bp(sino), torch_in)
