"""Pins of oracle/sp_ref.py (the separable-footprint prototypes of
sys_mat/sp.py) by closed forms."""
import numpy as np

from oracle import sp_ref


def test_triangle_mass_on_the_central_line():
    """DSD = 2 DSO and du = d: a pixel on the line through the origin
    perpendicular to the view (yp = 0) has h = 2 bins, so its weights sum to
    exactly 2 (sum_k max(0, 1 - |t - k|/m) = m for integer m)."""
    nx = ny = 15                       # odd: pixel rows through the origin
    nu = 64
    for ang, pick in ((0.0, lambda W: W[:, 7::15]), (np.pi / 2, lambda W: W[:, 7 * 15:8 * 15])):
        W = sp_ref.fan_matrix(nx, ny, [ang], nu, 50.0, 100.0, du=1.0, d_pix=1.0).toarray()
        sums = pick(W).sum(0)
        assert np.allclose(sums, 2.0, rtol=0, atol=1e-12)


def test_parallel_limit():
    nx, ny, nu, d, du = 9, 7, 30, 0.8, 0.6
    angs = np.array([0.0, 0.4, 1.3, 2.9])
    big = 1e9
    W = sp_ref.fan_matrix(nx, ny, angs, nu, big, big, du=du, d_pix=d).toarray()
    x = d * (np.arange(nx) - nx / 2 + 0.5)[:, None]
    y = d * (np.arange(ny) - ny / 2 + 0.5)[None, :]
    ub = du * (np.arange(nu) - nu / 2 + 0.5)
    for a, th in enumerate(angs):
        p = (np.cos(th) * x + np.sin(th) * y).reshape(-1)
        want = np.maximum(0.0, 1.0 - np.abs(ub[:, None] - p[None, :]) / d)
        assert np.allclose(W[a * nu:(a + 1) * nu], want, rtol=0, atol=1e-6)


def _geom():
    return (5, 4, 3, 12, 10, 1.0, 1.2, 0.9, 0.8, 0.7, 6.0, 10.0)


def test_single_voxel_toy_projection():
    g = _geom()
    nx, ny, nz, nu, nv, dx, dy, dz, su, sv, Ds, Dd = g
    vol = np.zeros((nx, ny, nz))
    vol[3, 1, 2] = 2.5
    proj = sp_ref.sep_fp(vol, *g)
    u = (np.arange(nu) - nu / 2 + 0.5) * su
    v = (np.arange(nv) - nv / 2 + 0.5) * sv
    xs = (3 - nx / 2 + 0.5) * dx
    ys = (1 - ny / 2 + 0.5) * dy
    zs = (2 - nz / 2 + 0.5) * dz
    f = lambda r: np.maximum(1 - np.abs(r), 0)  # noqa: E731
    want = 2.5 * np.outer(f(v / Dd - ys / Ds), f(u / Dd - xs / Ds)) * f(zs / Ds)
    assert np.abs(want).max() > 0
    assert np.allclose(proj, want, rtol=0, atol=1e-14)


def test_toy_backward_is_the_transpose():
    g = _geom()
    rng = np.random.default_rng(0)
    x = rng.normal(size=g[:3])
    y = rng.normal(size=(g[4], g[3]))
    lhs = float((sp_ref.sep_fp(x, *g) * y).sum())
    rhs = float((x * sp_ref.sep_bp(y, *g)).sum())
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)
