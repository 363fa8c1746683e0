"""ORACLE: filtered backprojection in float64 -- test infrastructure only.

* ramp_filter / window_filter / filter_sino restate filtration.py:4-329:
  the ramp built from the real-space Ram-Lak kernel h[0] = 1/(4 du^2),
  h[odd k] = -1/(pi^2 du^2 k^2) on the zero-padded grid z = 2^ceil(log2 2 nu)
  (filtration.py:60-100) and transformed with the rFFT, scaled by du, cut
  off above cutoff/du; the window of window_filter; filtering by rFFT of the
  zero-padded rows, times 2 H, inverse rFFT, first nu samples
  (filtration.py:316-329).
* fbp() is the Kak & Slaney filtered backprojection for the geometries of
  the DD projector (sys_mat/dd.py conventions: p = -s x + c y, o = c x + s y,
  source at o = DSO): parallel; fan with a flat detector at DSD (pre-weight
  DSO/sqrt(DSO^2 + s^2) on the virtual detector s = u DSO/DSD, backprojection
  weight (DSO/(DSO - o))^2); equiangular fan (pre-weight DSO cos(gamma),
  kernel times (gamma/sin gamma)^2, weight 1/L^2); linear interpolation
  between bin centres, zero outside; scale pi/(2 na) (the reference's fan FBP
  normalisation, processing/reconstruction.py:60, with filter_sino's 2 H).
  The reference's own fan_fbp passes keyword names filter_sino does not
  accept (reconstruction.py:51) and backprojects in the coordinates of an
  older projector module (processing/proj2.py:580-614); this restates its
  steps in the path's geometry.

Pins (tests/test_fbp_oracle.py): the ramp's kernel against the closed-form
Ram-Lak taps, filter_sino against direct convolution with its taps, window
values at 0 and at the cutoff, and reconstructions of analytic disc
sinograms (parallel, flat fan, equiangular fan) against the disc.
"""
import numpy as np


def ramp_filter(nu, du=1.0, cutoff=0.5, zero_pad=True, half_spectrum=True, real_space=True,
                return_freq=False):
    n = int(2 ** np.ceil(np.log2(2 * nu))) if zero_pad else nu
    f = np.fft.rfftfreq(n, d=du) if half_spectrum else np.fft.fftfreq(n, d=du)
    if real_space:
        k = np.fft.fftfreq(n, d=1.0 / n)
        h = np.zeros(n)
        h[0] = 1.0 / (4.0 * du ** 2)
        h[1::2] = -1.0 / (np.pi ** 2 * du ** 2 * k[1::2] ** 2)
        H = du * np.real(np.fft.rfft(h) if half_spectrum else np.fft.fft(h))
    else:
        H = np.abs(f)
    H = H.copy()
    H[np.abs(f) > cutoff / du] = 0.0
    return (H, f) if return_freq else H


def window_filter(H, f, du=1.0, filter_type="ramp", cutoff=0.5, sigma=0.5):
    fc = cutoff / du
    fn = f / fc
    valid = np.abs(f) <= fc
    W = np.zeros_like(f)
    if filter_type in ("ramp", "ram-lak"):
        return H
    if filter_type == "hann":
        W[valid] = 0.5 + 0.5 * np.cos(np.pi * fn[valid])
    elif filter_type == "hamming":
        W[valid] = 0.54 + 0.46 * np.cos(np.pi * fn[valid])
    elif filter_type == "blackman":
        W[valid] = 0.42 + 0.5 * np.cos(np.pi * fn[valid]) + 0.08 * np.cos(2 * np.pi * fn[valid])
    elif filter_type == "shepp-logan":
        W[0] = 1.0
        m = valid & (f != 0)
        W[m] = np.sin(np.pi * f[m] / (2 * fc)) / (np.pi * f[m] / (2 * fc))
    elif filter_type == "cosine":
        W[valid] = np.cos(np.pi * fn[valid] / 2.0)
    elif filter_type == "gaussian":
        W = np.exp(-0.5 * (f / sigma) ** 2)
    else:
        raise ValueError(f"Unknown filter type: {filter_type}")
    return H * W


def filter_sino(sino, du=1.0, filter_type="ramp", cutoff=0.5, real_space=True):
    if filter_type is None:
        return sino
    sino = np.asarray(sino, np.float64)
    nu = sino.shape[-1]
    z = int(2 ** np.ceil(np.log2(2 * nu)))
    H, f = ramp_filter(nu, du=du, cutoff=cutoff, zero_pad=True, half_spectrum=True, real_space=real_space,
                       return_freq=True)
    H = window_filter(H, f, du=du, filter_type=filter_type, cutoff=cutoff)
    out = np.fft.irfft(np.fft.rfft(sino, axis=-1, n=z) * (2.0 * H), n=z, axis=-1)
    return out[..., :nu]


def taps(nu, du=1.0, filter_type="ramp", cutoff=0.5, real_space=True):
    """The convolution kernel g[0..nu-1] of filter_sino (g = irfft(2 H), even)."""
    z = int(2 ** np.ceil(np.log2(2 * nu)))
    H, f = ramp_filter(nu, du=du, cutoff=cutoff, zero_pad=True, half_spectrum=True, real_space=real_space,
                       return_freq=True)
    H = window_filter(H, f, du=du, filter_type=filter_type, cutoff=cutoff)
    return np.fft.irfft(2.0 * H, n=z)[:nu]


def _bins(nu, du, su):
    return du * (np.arange(nu) + 0.5 - nu / 2.0 + su)


def fbp(sino, ang, img_shape, geom="par", DSO=None, DSD=None, du=1.0, su=0.0, d_pix=1.0, detector="flat",
        filter_type="ramp", cutoff=0.5):
    sino = np.asarray(sino, np.float64)
    na, nu = sino.shape
    u = _bins(nu, du, su)
    if geom == "par":
        coord, spacing, w_in = u, du, np.ones(nu)
    elif detector == "flat":
        coord = u * DSO / DSD
        spacing = du * DSO / DSD
        w_in = DSO / np.sqrt(DSO ** 2 + coord ** 2)
    else:
        coord = u / DSD
        spacing = du / DSD
        w_in = DSO * np.cos(coord)
    g = taps(nu, spacing, filter_type, cutoff)
    if geom == "fan" and detector == "arc":
        k = np.arange(1, nu) * spacing
        g = g.copy()
        g[1:] *= (k / np.sin(k)) ** 2
    # linear convolution with the even kernel (filter_sino's result for the ramp)
    z = int(2 ** np.ceil(np.log2(2 * nu)))
    kern = np.zeros(z)
    kern[:nu] = g
    kern[z - nu + 1:] = g[1:][::-1]
    q = np.fft.irfft(np.fft.rfft(sino * w_in, n=z, axis=-1) * np.fft.rfft(kern), n=z, axis=-1)[:, :nu]
    nx, ny = img_shape
    x = d_pix * (np.arange(nx) + 0.5 - nx / 2.0)
    y = d_pix * (np.arange(ny) + 0.5 - ny / 2.0)
    X, Y = np.meshgrid(x, y, indexing="ij")
    img = np.zeros((nx, ny))
    ang = np.asarray(ang)
    c_all, s_all = np.cos(ang.astype(np.float32)).astype(np.float64), np.sin(ang.astype(np.float32)).astype(np.float64)
    for ia in range(na):
        c, s = c_all[ia], s_all[ia]
        p = -s * X + c * Y
        if geom == "par":
            t, w = p, np.ones_like(p)
        else:
            den = DSO - (c * X + s * Y)
            if detector == "flat":
                t, w = DSO * p / den, (DSO / den) ** 2
            else:
                t, w = np.arctan2(p, den), 1.0 / (den ** 2 + p ** 2)
        tf = (t - coord[0]) / spacing
        j = np.floor(tf).astype(np.int64)
        fr = tf - j
        a = np.where((j >= 0) & (j < nu), q[ia, np.clip(j, 0, nu - 1)], 0.0)
        b = np.where((j + 1 >= 0) & (j + 1 < nu), q[ia, np.clip(j + 1, 0, nu - 1)], 0.0)
        img += w * ((1.0 - fr) * a + fr * b)
    return img * np.pi / (2.0 * na)


def disc_sinogram(ang, nu, geom="par", DSO=None, DSD=None, du=1.0, su=0.0, detector="flat", R=1.0, cx=0.0,
                  cy=0.0):
    """Line integrals of a unit-density disc along the rays through the bin
    centres, in the DD projector's conventions."""
    u = _bins(nu, du, su)
    ang = np.asarray(ang, np.float64)
    c, s = np.cos(ang)[:, None], np.sin(ang)[:, None]
    if geom == "par":
        dist = np.abs((-s * cx + c * cy) - u[None, :])
    else:
        t = (u / DSD if detector == "flat" else np.tan(u / DSD))[None, :]
        dist = np.abs((-s + t * c) * cx + (c + t * s) * cy - t * DSO) / np.sqrt(1.0 + t * t)
    return 2.0 * np.sqrt(np.clip(R * R - dist * dist, 0.0, None))
