"""Pins of the ray-driven oracle (oracle/rd_ref.py).

* AW forward = path lengths of the Siddon oracle (a sort-based C restatement
  of siddons(), a different algorithm) on the same rays, par/fan/cone;
* Joseph forward of a constant image = value * 0.5 * number of samples, with
  the sample count from the closed-form chord;
* Joseph forward of a linear image: bilinear/trilinear interpolation is exact
  for linear functions, so the sum is the linear function at the (clamped /
  valid) sample positions -- catches transposed or swapped corners;
* backprojection = transpose of the forward projection where the reference
  pairs are adjoint (AW everywhere, cone without its ray scale; Joseph for
  images away from the border).
"""
import math

import numpy as np
import pytest

from oracle import projection as op
from oracle import rd_ref
from oracle import siddon as osd

GEOMS = {
    "par": dict(mode="par", ang=np.linspace(0.05, np.pi, 9, endpoint=False), nu=17, du=0.9,
                su=0.3),
    "fan": dict(mode="fan", ang=np.linspace(0.1, 2 * np.pi, 8, endpoint=False), nu=21, du=1.1,
                su=-0.2, DSO=30.0, DSD=50.0),
    "cone": dict(mode="cone", ang=np.linspace(0.2, 2 * np.pi, 6, endpoint=False), nu=11,
                 du=1.3, su=0.1, DSO=25.0, DSD=40.0, nv=11, dv=1.2, sv=0.2),
}


def _img(shape, seed):
    return np.random.default_rng(seed).uniform(0.2, 1.0, shape)


@pytest.mark.parametrize("name", ["par", "fan", "cone"])
def test_aw_equals_siddon_path_lengths(name):
    g = dict(GEOMS[name])
    mode = g.pop("mode")
    shape = (7, 6, 5) if mode == "cone" else (8, 6)
    img = _img(shape, 0)
    d = 0.8
    ref = rd_ref.fp(img, mode, d_pix=d, **g)
    if mode == "cone":
        ref = ref / np.array([[[g["DSD"] / math.sqrt(g["DSD"] ** 2 + u * u + v * v)
                                for v in g["dv"] * (np.arange(g["nv"]) - (g["nu"] / 2 - 0.5) + g["sv"])]
                               for u in g["du"] * (np.arange(g["nu"]) - (g["nu"] / 2 - 0.5) + g["su"])]])
    # short rays: siddons() drops rays whose alpha interval is below its
    # epsilon gate, so keep the parametrisation tight around the box
    src, trg = rd_ref.ray_endpoints(mode, reach=20.0 if mode == "par" else 2 * g["DSD"], **g)
    nP = shape if mode == "cone" else shape + (1,)
    ind, ln, off = osd.siddons(src.astype(np.float32), trg.astype(np.float32), nP, d, 0.0, flat=True)
    sid = op.fp64(img.reshape(nP).astype(np.float32), ind, ln, off).reshape(ref.shape)
    np.testing.assert_allclose(ref, sid, rtol=2e-4, atol=2e-4 * np.abs(sid).max())


def _chord(o, h, shape, d):
    t0, t1 = -math.inf, math.inf
    for a in range(len(shape)):
        lo, hi = -d * shape[a] / 2, d * shape[a] / 2
        if abs(h[a]) > 1e-6:
            ta, tb = (lo - o[a]) / h[a], (hi - o[a]) / h[a]
            t0, t1 = max(t0, min(ta, tb)), min(t1, max(ta, tb))
    return t0, t1


@pytest.mark.parametrize("name", ["par", "fan"])
def test_joseph_constant_and_linear_2d(name):
    g = dict(GEOMS[name])
    mode = g.pop("mode")
    shape, d = (9, 7), 1.1
    ii, jj = np.meshgrid(np.arange(9), np.arange(7), indexing="ij")
    a, b, c = 0.7, -0.3, 2.0
    for img, linear in ((np.full(shape, 3.0), False), (a * ii + b * jj + c, True)):
        got = rd_ref.fp(img, mode, d_pix=d, joseph=True, **g).reshape(-1)
        want = []
        for _, _, _, o, h, _, _ in rd_ref.rays(mode, **g):
            t0, t1 = _chord(o, h, shape, d)
            if not t1 > t0:
                want.append(0.0)
                continue
            K = int(math.floor((t1 - t0) / 0.5 + 1e-12)) + 1
            if not linear:
                want.append(3.0 * 0.5 * K)
                continue
            s = 0.0
            for k in range(K):
                t = t0 + 0.5 * k
                fx = (o[0] + t * h[0] + d * (shape[0] / 2 - 0.5)) / d
                fy = (o[1] + t * h[1] + d * (shape[1] / 2 - 0.5)) / d
                fx, fy = min(max(fx, 0), shape[0] - 2), min(max(fy, 0), shape[1] - 2)
                s += (a * fx + b * fy + c) * 0.5
            want.append(s)
        np.testing.assert_allclose(got, want, rtol=1e-9, atol=1e-9)


def test_joseph_linear_3d_truncation_rule():
    g = dict(GEOMS["cone"])
    g.pop("mode")
    shape, d = (6, 7, 5), 1.0
    ii, jj, kk = np.meshgrid(*[np.arange(n) for n in shape], indexing="ij")
    img = 0.5 * ii - 0.2 * jj + 0.3 * kk + 1.0
    got = rd_ref.fp(img, "cone", d_pix=d, joseph=True, **g).reshape(-1)
    want = []
    for _, _, _, o, h, _, _ in rd_ref.rays("cone", **g):
        t0, t1 = _chord(o, h, shape, d)
        s = 0.0
        if t1 > t0:
            K = int(math.floor((t1 - t0) / 0.5 + 1e-12)) + 1
            for k in range(K):
                t = t0 + 0.5 * k
                f = [(o[a] + t * h[a] + d * (shape[a] / 2 - 0.5)) / d for a in range(3)]
                if all(0 <= int(f[a]) < shape[a] - 1 for a in range(3)):
                    s += (0.5 * f[0] - 0.2 * f[1] + 0.3 * f[2] + 1.0) * 0.5
        want.append(s)
    np.testing.assert_allclose(got, want, rtol=1e-9, atol=1e-9)


@pytest.mark.parametrize("name", ["par", "fan", "cone"])
@pytest.mark.parametrize("joseph", [False, True])
def test_backprojection_is_transpose(name, joseph):
    g = dict(GEOMS[name])
    mode = g.pop("mode")
    shape = (7, 6, 6) if mode == "cone" else (9, 8)
    rng = np.random.default_rng(1)
    x = np.zeros(shape)
    inner = tuple(slice(2, n - 2) for n in shape)       # away from the border (Joseph)
    x[inner] = rng.normal(size=x[inner].shape)
    y_shape = (len(g["ang"]), g["nu"], g["nv"]) if mode == "cone" else (len(g["ang"]), g["nu"])
    y = rng.normal(size=y_shape)
    Ax = rd_ref.fp(x, mode, d_pix=0.9, joseph=joseph, **g)
    if mode == "cone" and not joseph:
        u = g["du"] * (np.arange(g["nu"]) - (g["nu"] / 2 - 0.5) + g["su"])
        v = g["dv"] * (np.arange(g["nv"]) - (g["nu"] / 2 - 0.5) + g["sv"])
        Ax = Ax / (g["DSD"] / np.sqrt(g["DSD"] ** 2 + u[:, None] ** 2 + v[None, :] ** 2))[None]
    kw = {k: v for k, v in g.items() if k not in ("nu", "nv")}
    ATy = rd_ref.bp(y, mode, img_shape=shape, d_pix=0.9, joseph=joseph, reference_scale=False,
                    **kw)
    lhs, rhs = (Ax * y).sum(), (x * ATy).sum()
    assert abs(lhs - rhs) <= 1e-10 * max(1.0, abs(lhs))


def test_reference_bp_scaling():
    g = dict(GEOMS["par"])
    g.pop("mode")
    y = np.random.default_rng(2).normal(size=(9, 17))
    kw = {k: v for k, v in g.items() if k != "nu"}
    a = rd_ref.bp(y, "par", img_shape=(8, 6), d_pix=0.8, reference_scale=True, **kw)
    b = rd_ref.bp(y, "par", img_shape=(8, 6), d_pix=0.8, reference_scale=False, **kw)
    np.testing.assert_allclose(a, b / 9 / 0.8 / 0.8 * 0.9, rtol=1e-12)


def test_clamp_parallel_closed_form_and_transpose():
    """clamp_parallel=True (rd.py:66-73): a parallel-beam ray along x that
    passes outside the image is clamped onto the nearest border row, so its
    AW sum is that row's sum times d; at 90 degrees the same holds for the
    columns.  Inside rays are unchanged, and the clamped BP stays the
    transpose of the clamped FP."""
    rng = np.random.default_rng(21)
    d, shape = 0.9, (6, 5)
    img = rng.uniform(0.2, 1.0, shape)
    ang = np.array([0.0, np.pi / 2])
    nu, du, su = 17, 0.9, 0.3
    fp0 = rd_ref.fp(img, "par", ang, nu, du, su, d)
    fp1 = rd_ref.fp(img, "par", ang, nu, du, su, d, clamp_parallel=True)
    u = du * (np.arange(nu) - (nu / 2 - 0.5) + su)
    below, above = u < -shape[1] * d / 2, u >= shape[1] * d / 2
    assert below.any() and above.any()
    np.testing.assert_allclose(fp1[0, below], img[:, 0].sum() * d, rtol=1e-12)
    np.testing.assert_allclose(fp1[0, above], img[:, -1].sum() * d, rtol=1e-12)
    assert np.all(fp0[0, below | above] == 0)
    inside = ~(below | above)
    np.testing.assert_array_equal(fp1[0, inside], fp0[0, inside])
    # 90 degrees: rays along y at x = -u
    x = -u
    left, right = x < -shape[0] * d / 2, x >= shape[0] * d / 2
    assert left.any() and right.any()
    np.testing.assert_allclose(fp1[1, left], img[0, :].sum() * d, rtol=1e-12)
    np.testing.assert_allclose(fp1[1, right], img[-1, :].sum() * d, rtol=1e-12)
    y = rng.uniform(0, 1, fp1.shape)
    b = rd_ref.bp(y, "par", ang, shape, du, su, d, clamp_parallel=True, reference_scale=False)
    x_ = rng.uniform(0, 1, shape)
    f = rd_ref.fp(x_, "par", ang, nu, du, su, d, clamp_parallel=True)
    assert abs(np.vdot(f, y) - np.vdot(x_, b)) <= 1e-12 * abs(np.vdot(f, y))
