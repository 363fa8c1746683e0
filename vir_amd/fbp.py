"""Filtered backprojection on the GPU (libvir_amd: vir_fbp_filter, vir_fbp_bp).

Drop-ins for filtration.py (filtration.py:4-329):

  ramp_filter(nu, du=1.0, cutoff=0.5, zero_pad=True, half_spectrum=True,
              real_space=True, return_freq=False)          filtration.py:4-115
  window_filter(H, f, du=1.0, filter_type="ramp", cutoff=0.5, sigma=0.5)
                                                            filtration.py:118-264
  filter_sino(sino, du=1.0, filter_type="ramp", cutoff=0.5, real_space=True)
                                                            filtration.py:267-329

ramp_filter / window_filter design the filter on the host (float64, once per
geometry, like the projection matrices of the Siddon backprojector);
filter_sino runs on the device as the equivalent linear convolution with the
filter's taps.

FBP(img_shape, ang_arr, nu, geom="par"|"fan", DSO, DSD, du, su, d_pix,
detector="flat"|"arc", filter_type, cutoff) reconstructs in the geometry of
the DD projector (vir_amd.sys_mat.dd.DDProjector): the reference's fan FBP
steps (processing/reconstruction.py:7-65 -- pre-weighting, filter_sino,
pixel-driven backprojection, pi/(2 n_views)) with Kak & Slaney's weights for a
flat or an equiangular detector, and parallel beam.  fbp(...) is the
one-call form.  Used as the initial image of SIRT it cuts the iterations
needed for a given error.
"""
import numpy as np
import torch

from . import _lib
from .device import as_device_f32, out_tensor, require_gpu, stream

_MODES = {"flat": 1, "arc": 2}


def ramp_filter(nu, du=1.0, cutoff=0.5, zero_pad=True, half_spectrum=True, real_space=True,
                return_freq=False):
    """Ramp (Ram-Lak) filter H(f) on the rFFT (or FFT) grid of nu samples
    zero-padded to the next power of two >= 2 nu; real_space=True transforms
    the real-space Ram-Lak kernel (scaled by du), otherwise H = |f|.  Zero
    above cutoff/du."""
    n = int(2 ** np.ceil(np.log2(2 * nu))) if zero_pad else int(nu)
    f = np.fft.rfftfreq(n, d=du) if half_spectrum else np.fft.fftfreq(n, d=du)
    if real_space:
        k = np.fft.fftfreq(n, d=1.0 / n)
        h = np.zeros(n)
        h[0] = 1.0 / (4.0 * du ** 2)
        h[1::2] = -1.0 / (np.pi ** 2 * du ** 2 * k[1::2] ** 2)
        spec = np.fft.rfft(h) if half_spectrum else np.fft.fft(h)
        H = du * np.real(spec)
    else:
        H = np.abs(f)
    H[np.abs(f) > cutoff * (1.0 / du)] = 0.0
    return (H, f) if return_freq else H


def window_filter(H, f, du=1.0, filter_type="ramp", cutoff=0.5, sigma=0.5):
    """H(f) W(f) for the windows of filtration.py:118-264 ("ramp"/"ram-lak",
    "hann", "hamming", "blackman", "shepp-logan", "cosine", "gaussian")."""
    f_c = cutoff * (1.0 / du)
    f_norm = f / f_c
    valid = np.abs(f) <= f_c
    W = np.zeros_like(f)
    if filter_type in ("ram-lak", "ramp"):
        return H
    if filter_type == "hann":
        W[valid] = 0.5 + 0.5 * np.cos(np.pi * f_norm[valid])
    elif filter_type == "hamming":
        W[valid] = 0.54 + 0.46 * np.cos(np.pi * f_norm[valid])
    elif filter_type == "blackman":
        W[valid] = 0.42 + 0.5 * np.cos(np.pi * f_norm[valid]) + 0.08 * np.cos(2 * np.pi * f_norm[valid])
    elif filter_type == "shepp-logan":
        W[0] = 1.0
        m = valid & (f != 0)
        W[m] = np.sin(np.pi * f[m] / (2 * f_c)) / (np.pi * f[m] / (2 * f_c))
    elif filter_type == "cosine":
        W[valid] = np.cos(np.pi * f_norm[valid] / 2.0)
    elif filter_type == "gaussian":
        W = np.exp(-0.5 * (f / sigma) ** 2)
    else:
        raise ValueError(f"Unknown filter type: {filter_type}")
    return H * W


def filter_taps(nu, du=1.0, filter_type="ramp", cutoff=0.5, real_space=True):
    """The kernel g[0..nu-1] (float64) with filter_sino(x)[i] =
    sum_j x[j] g[|i - j|]: g = irfft(2 H) on the zero-padded grid, even, and
    the padding (>= 2 nu) keeps the circular product free of wrap-around for
    the first nu outputs."""
    z = int(2 ** np.ceil(np.log2(2 * nu)))
    H, f = ramp_filter(nu, du=du, cutoff=cutoff, zero_pad=True, half_spectrum=True, real_space=real_space,
                       return_freq=True)
    H = window_filter(H, f, du=du, filter_type=filter_type, cutoff=cutoff)
    return np.fft.irfft(2.0 * H, n=z)[:nu]


def _filter_device(sino, taps, weight=None):
    nu = sino.shape[-1]
    rows = sino.numel() // nu if nu else 0
    out = torch.empty_like(sino)
    t = torch.from_numpy(np.asarray(taps, np.float32)).to(sino.device)
    w = torch.from_numpy(np.asarray(weight, np.float32)).to(sino.device) if weight is not None else None
    _lib.check(_lib.lib().vir_fbp_filter(_lib.ptr(sino), rows, nu, _lib.ptr(w) if w is not None else None,
                                         _lib.ptr(t), _lib.ptr(out), stream()), "vir_fbp_filter")
    return out


def filter_sino(sino, du=1.0, filter_type="ramp", cutoff=0.5, real_space=True):
    """filtration.filter_sino: filter along the last axis with the windowed
    ramp (times 2, zero-padded), on the device.  None -> unfiltered.  Torch
    input stays on the device; numpy input returns numpy (float32)."""
    if filter_type is None:
        return sino
    require_gpu()
    torch_in = isinstance(sino, torch.Tensor)
    x = as_device_f32(sino).contiguous()
    out = _filter_device(x, filter_taps(x.shape[-1], du, filter_type, cutoff, real_space))
    return out if torch_in else out.cpu().numpy()


class FBP:
    """Filtered backprojection for the DD projector's 2-D geometries.

    geom "par": detector coordinate u = -sin(a) x + cos(a) y; "fan": source
    at distance DSO, detector at DSD, detector="flat" (u on the flat
    detector) or "arc" (u = arc length at radius DSD).  Bins at
    du (j + 0.5 - nu/2 + su).  Views uniformly spaced over a multiple of pi."""

    def __init__(self, img_shape, ang_arr, nu, geom="par", DSO=0.0, DSD=0.0, du=1.0, su=0.0, d_pix=1.0,
                 detector="flat", filter_type="ramp", cutoff=0.5):
        require_gpu()
        if geom not in ("par", "fan"):
            raise ValueError(f"unknown geometry {geom!r}")
        if geom == "fan" and detector not in _MODES:
            raise ValueError(f"unknown detector {detector!r}")
        ang = np.asarray(ang_arr, dtype=np.float32).reshape(-1)
        self.na, self.nu = ang.size, int(nu)
        self.img_shape = tuple(int(v) for v in img_shape)
        self.cos = torch.from_numpy(np.cos(ang).astype(np.float32)).cuda()
        self.sin = torch.from_numpy(np.sin(ang).astype(np.float32)).cuda()
        g = _lib.VirDDGeom()
        g.nx, g.ny, g.nz = self.img_shape[0], self.img_shape[1], 1
        g.nu, g.nv, g.na = self.nu, 1, self.na
        g.mode = 0 if geom == "par" else _MODES[detector]
        for k, v in dict(d_pix=d_pix, du=du, dv=1.0, su=su, sv=0.0, DSO=DSO, DSD=DSD).items():
            setattr(g, k, float(np.float32(v)))
        self.g = g
        u = float(du) * (np.arange(self.nu) + 0.5 - self.nu / 2.0 + float(su))
        if geom == "par":
            spacing, weight = float(du), None
        elif detector == "flat":
            s = u * DSO / DSD
            spacing, weight = float(du) * DSO / DSD, DSO / np.sqrt(DSO ** 2 + s ** 2)
        else:
            gam = u / DSD
            spacing, weight = float(du) / DSD, DSO * np.cos(gam)
        taps = filter_taps(self.nu, spacing, filter_type, cutoff)
        if geom == "fan" and detector == "arc":
            k = np.arange(1, self.nu) * spacing
            taps[1:] *= (k / np.sin(k)) ** 2
        self.taps = torch.from_numpy(taps.astype(np.float32)).cuda()
        self.weight = torch.from_numpy(weight.astype(np.float32)).cuda() if weight is not None else None
        self.scale = float(np.pi / (2.0 * max(self.na, 1)))

    def filtered(self, sino):
        """The pre-weighted, filtered sinogram (device)."""
        sino = as_device_f32(sino).contiguous()
        if sino.numel() != self.na * self.nu:
            raise ValueError("sinogram size does not match")
        out = torch.empty((self.na, self.nu), dtype=torch.float32, device=sino.device)
        if self.na:
            _lib.check(_lib.lib().vir_fbp_filter(_lib.ptr(sino), self.na, self.nu,
                                                 _lib.ptr(self.weight) if self.weight is not None else None,
                                                 _lib.ptr(self.taps), _lib.ptr(out), stream()), "vir_fbp_filter")
        return out

    def __call__(self, sino, out=None):
        q = self.filtered(sino)
        out = out_tensor(out, self.img_shape, q.device)
        _lib.check(_lib.lib().vir_fbp_bp(self.g, _lib.ptr(self.cos), _lib.ptr(self.sin), _lib.ptr(q),
                                         _lib.ptr(out), self.scale, 0, stream()), "vir_fbp_bp")
        return out


def fbp(sino, ang_arr, img_shape, geom="par", DSO=0.0, DSD=0.0, du=1.0, su=0.0, d_pix=1.0, detector="flat",
        filter_type="ramp", cutoff=0.5):
    """One-call FBP; torch input -> device image, numpy input -> numpy."""
    torch_in = isinstance(sino, torch.Tensor)
    nu = sino.shape[-1]
    img = FBP(img_shape, ang_arr, nu, geom, DSO, DSD, du, su, d_pix, detector, filter_type, cutoff)(sino)
    return img if torch_in else img.cpu().numpy()
