"""Pins of the separable-footprint oracle (oracle/pd_ref.py) by closed forms.

Each check is chosen so that a plausible slip in the restatement fails it:
corners vs centre, the amplitude, the fan division, the cone v
magnification (fm vs fm/cos g), the bin origin su, the output scalings.
"""
import numpy as np
import pytest

from oracle import pd_ref


def test_par_view0_and_view90_are_row_and_column_sums():
    rng = np.random.default_rng(0)
    nx, ny, d = 7, 9, 1.0
    img = rng.uniform(0, 1, (nx, ny))
    ang = np.array([0.0, np.pi / 2])
    W = pd_ref.matrix("par", (nx, ny), ang, nu=11, du=1.0, su=0.0, d_pix=d)
    sino = pd_ref.fp(W, img, d).reshape(2, 11)
    # view 0: p = y, bins aligned with the ny pixel rows, offset (11-9)/2 = 1
    np.testing.assert_allclose(sino[0, 1:10], img.sum(0), rtol=1e-13)
    assert np.all(sino[0, [0, 10]] == 0)
    # view 90: p = -x, footprint of column ix is [-x_r, -x_l]
    np.testing.assert_allclose(sino[1, 2:9], img.sum(1)[::-1], rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("ang", [0.1, 0.7, np.pi / 4, 2.0, 3.0])
def test_par_mass_conservation(ang):
    rng = np.random.default_rng(1)
    nx, ny, d, du = 8, 6, 0.7, 0.45
    img = rng.uniform(0, 1, (nx, ny))
    W = pd_ref.matrix("par", (nx, ny), [ang], nu=60, du=du, su=0.3, d_pix=d)
    sino = pd_ref.fp(W, img, d)
    assert abs(sino.sum() * du - d * d * img.sum()) < 1e-12 * img.sum()


def test_par_45_degree_pixel_footprint():
    # a single pixel at 45 degrees: rectangle of width d*sqrt(2) centred on
    # its centre projection, height 1/sqrt(2)
    nx = ny = 3
    img = np.zeros((3, 3))
    img[1, 1] = 1.0
    W = pd_ref.matrix("par", (3, 3), [np.pi / 4], nu=400, du=0.01, d_pix=1.0)
    prof = pd_ref.fp(W, img, 1.0)
    ub = pd_ref.det_bnd(400, 0.01, 0.0)
    cen = 0.5 * (ub[1:] + ub[:-1])
    inside = np.abs(cen) < np.sqrt(2) / 2 - 0.01
    outside = np.abs(cen) > np.sqrt(2) / 2 + 0.01
    np.testing.assert_allclose(prof[inside], 1 / np.sqrt(2), rtol=1e-12)
    assert np.all(prof[outside] == 0)


def test_fan_centre_pixel_closed_form():
    DSO, DSD, d = 50.0, 80.0, 1.0
    u0, u1, amp, _ = pd_ref.footprints("fan", 3, 3, d, 0.0, DSO, DSD)
    # pixel (1,1) sits at the origin: corners (+-d/2, +-d/2), p = y, o = x
    w = DSD * (d / 2) / (DSO - d / 2)
    assert abs(u0[1, 1] + w) < 1e-13 and abs(u1[1, 1] - w) < 1e-13
    assert abs(amp[1, 1] - 1.0) < 1e-14
    # pixel (2, 1): centre (d, 0) -> o = d, still on the central ray
    w2 = DSD * (d / 2) / (DSO - 1.5 * d)
    assert abs(u1[2, 1] - w2) < 1e-13 and abs(u0[2, 1] + w2) < 1e-13
    # pixel (1, 2): centre (0, d), gamma = atan(u_c/DSD), u_c = DSD*d/DSO
    g = np.arctan(d / DSO)
    assert abs(amp[1, 2] - 1 / (abs(np.cos(g)) + abs(np.sin(g)))) < 1e-14


@pytest.mark.parametrize("mode", ["fan", "cone"])
def test_fan_and_cone_parallel_limit(mode):
    big = 1e7
    nx, ny = 6, 5
    for a in (0.3, 1.2, 2.5):
        p = pd_ref.footprints("par", nx, ny, 1.0, a)
        f = pd_ref.footprints(mode, nx, ny, 1.0, a, DSO=big, DSD=big)
        for k in range(3):
            np.testing.assert_allclose(f[k], p[k], rtol=0, atol=2e-6)


def test_fan_mass_on_central_ray():
    # centre pixel at angle 0: total deposit = d_pix * amp * width
    DSO, DSD = 40.0, 70.0
    img = np.zeros((3, 3))
    img[1, 1] = 2.0
    W = pd_ref.matrix("fan", (3, 3), [0.0], nu=200, du=0.05, DSO=DSO, DSD=DSD)
    s = pd_ref.fp(W, img, 1.0).sum() * 0.05
    width = 2 * DSD * 0.5 / (DSO - 0.5)
    assert abs(s - 2.0 * width) < 1e-12


def test_cone_off_axis_v_magnification():
    DSO, DSD, d = 30.0, 45.0, 1.0
    nx, ny, nz = 5, 5, 2
    u0, u1, amp, cm = pd_ref.footprints("cone", nx, ny, d, 0.0, DSO, DSD)
    # voxel column (ix=3, iy=4): centre x0 = 1, y0 = 2; at angle 0 o = x, p = y
    x0, y0 = 1.0, 2.0
    fm = DSD / (DSO - x0)
    uc = fm * y0
    assert abs(cm[3, 4] - fm * np.sqrt(1 + (uc / DSD) ** 2)) < 1e-12
    assert abs(u0[3, 4] - fm * 1.5) < 1e-12 and abs(u1[3, 4] - fm * 2.5) < 1e-12
    # its sinogram v-profile spans cm * [z_l, z_r] with total d_pix*fm*d*cm*d
    img = np.zeros((nx, ny, nz))
    img[3, 4, 1] = 1.0
    W = pd_ref.matrix("cone", (nx, ny, nz), [0.0], nu=40, du=0.25, d_pix=d, DSO=DSO,
                      DSD=DSD, nv=40, dv=0.1, sv=0.0)
    sino = pd_ref.fp(W, img, d).reshape(40, 40)
    assert abs(sino.sum() * 0.25 * 0.1 - d * (fm * d) * 1.0 * (cm[3, 4] * d)) < 1e-12
    vprof = sino.sum(0)
    vb = pd_ref.det_bnd(40, 0.1, 0.0)
    nz_rows = np.nonzero(vprof)[0]
    assert vb[nz_rows[0]] < 1e-12 and vb[nz_rows[-1] + 1] >= cm[3, 4] * d - 1e-12


def test_reference_bp_scaling_and_transpose():
    rng = np.random.default_rng(3)
    ang = np.linspace(0, np.pi, 9, endpoint=False)
    W = pd_ref.matrix("par", (6, 6), ang, nu=10, du=0.8, su=0.2, d_pix=0.9)
    x, y = rng.normal(size=36), rng.normal(size=90)
    lhs = pd_ref.fp(W, x, 0.9) @ y
    rhs = x @ pd_ref.bp_reference(W, y, 9, 0.9, 0.8) * (9 * 0.9 / 0.8) * 0.9
    assert abs(lhs - rhs) < 1e-12 * abs(lhs)


def test_par3d_is_slicewise_par_when_rows_match_slices():
    rng = np.random.default_rng(4)
    img = rng.uniform(0, 1, (5, 4, 3))
    ang = [0.4, 1.9]
    W3 = pd_ref.matrix("par3d", img.shape, ang, nu=9, du=1.0, d_pix=1.0, nv=3, dv=1.0)
    W2 = pd_ref.matrix("par", img.shape[:2], ang, nu=9, du=1.0, d_pix=1.0)
    s3 = pd_ref.fp(W3, img, 1.0).reshape(2, 9, 3)
    for z in range(3):
        np.testing.assert_allclose(s3[:, :, z].reshape(-1), pd_ref.fp(W2, img[:, :, z], 1.0),
                                   rtol=1e-13, atol=1e-14)


def test_fan_loop_matches_matrix_without_zeros_and_flags_the_stale_skip():
    rng = np.random.default_rng(5)
    ang = np.linspace(0, 2 * np.pi, 7, endpoint=False)
    kw = dict(nu=30, DSO=40.0, DSD=60.0, du=0.9, su=0.1, d_pix=1.1)
    img = rng.uniform(0.1, 1, (8, 7))
    W = pd_ref.matrix("fan", (8, 7), ang, **{k: kw[k] for k in ("nu", "du", "su", "d_pix", "DSO", "DSD")})
    ref = pd_ref.fp(W, img, 1.1)
    for stale in (True, False):
        got = pd_ref.fan_fp_loop(img, ang, **kw, stale_oy=stale).reshape(-1)
        np.testing.assert_allclose(got, ref, rtol=1e-12, atol=1e-12)
    img[:, :3] = 0.0   # zero runs at the start of each column
    fixed = pd_ref.fan_fp_loop(img, ang, **kw, stale_oy=False).reshape(-1)
    stale = pd_ref.fan_fp_loop(img, ang, **kw, stale_oy=True).reshape(-1)
    np.testing.assert_allclose(fixed, pd_ref.fp(W, img, 1.1), rtol=1e-12, atol=1e-12)
    assert np.abs(stale - fixed).max() > 1e-4 * np.abs(fixed).max()


def test_par3d_operators_equal_the_dense_matrix():
    """par3d_ops (used at workload size) applies matrix("par3d") and its
    transpose."""
    rng = np.random.default_rng(12)
    shape, nu, nv = (9, 7, 5), 20, 8
    ang = np.linspace(0, np.pi, 6, endpoint=False)
    W = pd_ref.matrix("par3d", shape, ang, nu, du=0.95, su=0.1, d_pix=1.3, nv=nv, dv=0.75, sv=0.25)
    fp_, wt = pd_ref.par3d_ops(shape, ang, nu, nv, du=0.95, dv=0.75, su=0.1, sv=0.25, d_pix=1.3)
    x = rng.uniform(0, 1, shape)
    y = rng.uniform(0, 1, (6, nu, nv))
    assert np.allclose(fp_(x).reshape(-1), pd_ref.fp(W, x, 1.3), rtol=0, atol=1e-13)
    assert np.allclose(wt(y).reshape(-1), W.T @ y.reshape(-1), rtol=0, atol=1e-13)
