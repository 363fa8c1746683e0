"""Pins of the filtered-backprojection oracle (oracle/fbp_ref.py).

* the ramp filter's convolution kernel irfft(2 H) is twice the closed-form
  Ram-Lak kernel du * h[k], h[0] = 1/(4 du^2), h[odd k] = -1/(pi^2 du^2 k^2),
  h[even k] = 0 (filtration.py:86-94);
* filter_sino (rFFT route) equals direct linear convolution with its taps,
  for every window;
* windows: 1 at f = 0 (cosine, Hann, Shepp-Logan), 0 at the cutoff (Hann,
  cosine), 0.08 = 0.54 - 0.46 at the cutoff (Hamming);
* reconstructions of analytic disc sinograms -- parallel, fan with a flat
  detector, fan with an equiangular detector -- equal the disc (density 1
  inside, 0 outside) away from its edge.
"""
import numpy as np
import pytest

from oracle import fbp_ref


def test_ramp_kernel_is_closed_form_ram_lak():
    for nu, du in ((37, 1.0), (64, 0.7), (101, 2.5)):
        g = fbp_ref.taps(nu, du)
        k = np.arange(nu)
        h = np.zeros(nu)
        h[0] = 1.0 / (4.0 * du * du)
        h[1::2] = -1.0 / (np.pi ** 2 * du * du * k[1::2] ** 2)
        assert np.allclose(g, 2.0 * du * h, rtol=1e-10, atol=1e-13)


@pytest.mark.parametrize("ft", ["ramp", "hann", "hamming", "blackman", "shepp-logan", "cosine", "gaussian"])
def test_filter_sino_is_linear_convolution_with_taps(ft):
    rng = np.random.default_rng(0)
    nu, du = 53, 0.8
    x = rng.normal(size=(4, nu))
    out = fbp_ref.filter_sino(x, du=du, filter_type=ft, cutoff=0.45)
    g = fbp_ref.taps(nu, du, ft, 0.45)
    i, j = np.meshgrid(np.arange(nu), np.arange(nu), indexing="ij")
    direct = x @ g[np.abs(i - j)].T
    assert np.allclose(out, direct, rtol=1e-10, atol=1e-12)


def test_window_values():
    du, cut = 1.0, 0.5
    H, f = fbp_ref.ramp_filter(40, du=du, cutoff=cut, return_freq=True, real_space=False)
    ones = np.ones_like(H)
    for ft in ("cosine", "hann", "shepp-logan"):
        assert fbp_ref.window_filter(ones, f, du, ft, cut)[0] == pytest.approx(1.0)
    fc = np.argmin(np.abs(f - cut / du))
    assert fbp_ref.window_filter(ones, f, du, "hann", cut)[fc] == pytest.approx(0.0, abs=1e-12)
    assert fbp_ref.window_filter(ones, f, du, "cosine", cut)[fc] == pytest.approx(0.0, abs=1e-12)
    assert fbp_ref.window_filter(ones, f, du, "hamming", cut)[fc] == pytest.approx(0.08)
    assert fbp_ref.filter_sino(np.ones((2, 5)), filter_type=None) is not None


@pytest.mark.parametrize("geom,detector,na,span", [
    ("par", "flat", 180, np.pi),
    ("par", "flat", 240, 2 * np.pi),
    ("fan", "flat", 360, 2 * np.pi),
    ("fan", "arc", 360, 2 * np.pi),
])
def test_disc_reconstruction(geom, detector, na, span):
    n, d, R, cx, cy = 96, 1.0, 25.0, 8.0, -5.0
    ang = np.arange(na) * span / na
    kw = dict(DSO=150.0, DSD=260.0) if geom == "fan" else {}
    du, nu = (1.0, 141) if geom == "par" else (1.6, 161)
    sino = fbp_ref.disc_sinogram(ang, nu, geom, du=du, detector=detector, R=R, cx=cx, cy=cy, **kw)
    img = fbp_ref.fbp(sino, ang, (n, n), geom, du=du, d_pix=d, detector=detector, **kw)
    x = d * (np.arange(n) + 0.5 - n / 2.0)
    X, Y = np.meshgrid(x, x, indexing="ij")
    r = np.hypot(X - cx, Y - cy)
    inside, outside = r < R - 3.0, (r > R + 3.0) & (np.hypot(X, Y) < 40.0)
    assert abs(img[inside].mean() - 1.0) < 0.01
    assert img[inside].std() < 0.03
    assert abs(img[outside].mean()) < 0.01
