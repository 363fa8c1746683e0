"""CPU: pin the distance-driven oracle (oracle/dd_ref.py).

The reference's numba kernels (fastmath) cannot run here (numba is absent),
so the float64 restatement is pinned by what the mathematics fixes:
  * mass conservation: sum_u sino*du = d_pix^2 * sum img (parallel, object
    fully on the detector) -- dd.py:120,725;
  * impulse response of one pixel at 0 and 90 degrees: a rectangle of width
    d_pix, height 1/|sin| (resp. 1/|cos|) times d_pix/du;
  * a brute-force system matrix (each pixel's projected segment against each
    bin, no sweep) equals the sweep FP column by column (par and fan);
  * the BP equals A^T * du / (na * d_pix^2) (the reference's BP scaling,
    dd.py:721-725, 799-803) -- the sweeps are transposes of each other;
  * a planted mistake (dropping the ray scale) breaks the brute-force match.
"""
import numpy as np
import pytest

from oracle import dd_ref


def test_mass_conservation_parallel():
    rng = np.random.default_rng(0)
    img = rng.uniform(0, 1, (16, 12))
    ang = np.linspace(0, np.pi, 13, endpoint=False) + 0.05
    du, d_pix = 0.7, 1.1
    sino = dd_ref.par_2d(img, ang, 64, du=du, d_pix=d_pix)
    tot = d_pix ** 2 * img.sum()
    for ia in range(ang.size):
        assert abs(sino[ia].sum() * du - tot) <= 1e-10 * tot


def test_impulse_response():
    img = np.zeros((8, 8))
    img[3, 5] = 1.0
    for a, (cs, sn) in [(0.0, (1.0, 0.0)), (np.pi / 2, (0.0, 1.0))]:
        sino = dd_ref.par_2d(img, np.array([a]), 40, du=0.25, d_pix=1.0)[0]
        # one pixel covers d_pix of detector length; height d_pix/du * scale.
        # cos/sin come from the float32 angle (cos(float32(pi/2)) = -4.4e-8),
        # so a sliver of order 1e-7 may fall into a fifth bin.
        assert abs(sino.sum() * 0.25 - 1.0) < 1e-6
        nz = sino[sino > 1e-5]
        assert len(nz) == 4 and np.allclose(nz, 1.0 / 0.25 * 0.25, atol=1e-6)


@pytest.mark.parametrize("geom", ["par", "fan"])
def test_sweep_equals_brute_force(geom):
    rng = np.random.default_rng(1)
    nx, ny, nu = 6, 5, 23
    # random views plus exact diagonals, where the float32 |sin| == |cos|
    # tie picks the driving axis
    ang = np.concatenate([rng.uniform(0, 2 * np.pi, 7), np.arange(1, 8, 2) * np.pi / 4])
    du, su, d_pix = 0.9, 0.25, 1.0
    kw = dict(DSO=30.0, DSD=45.0) if geom == "fan" else {}
    A = dd_ref.overlap_brute_2d(nx, ny, ang, nu, du, su, d_pix, geom=geom, **kw)
    for j in range(nx * ny):
        e = np.zeros(nx * ny)
        e[j] = 1.0
        img = e.reshape(nx, ny)
        if geom == "par":
            col = dd_ref.par_2d(img, ang, nu, du=du, su=su, d_pix=d_pix)
        else:
            col = dd_ref.fan_2d(img, ang, nu, 30.0, 45.0, du=du, su=su, d_pix=d_pix)
        assert np.allclose(col.reshape(-1), A[:, j], rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("geom", ["par", "fan"])
def test_bp_is_scaled_transpose(geom):
    rng = np.random.default_rng(2)
    nx, ny, nu = 7, 6, 21
    ang = rng.uniform(0, 2 * np.pi, 9)
    du, su, d_pix = 1.1, 0.0, 0.8
    kw = dict(DSO=30.0, DSD=45.0) if geom == "fan" else {}
    A = dd_ref.overlap_brute_2d(nx, ny, ang, nu, du, su, d_pix, geom=geom, **kw)
    y = rng.uniform(0, 1, (ang.size, nu))
    if geom == "par":
        b = dd_ref.par_2d(y, ang, (nx, ny), du=du, su=su, d_pix=d_pix, bp=True)
    else:
        b = dd_ref.fan_2d(y, ang, (nx, ny), 30.0, 45.0, du=du, su=su, d_pix=d_pix, bp=True)
    ref = (A.T @ y.reshape(-1)) * du / (ang.size * d_pix ** 2)
    assert np.allclose(b.reshape(-1), ref, rtol=1e-12, atol=1e-12)


def test_planted_mistake_is_caught():
    """Dropping the 1/|sin| ray scale (a plausible slip) must break the
    brute-force match, so the comparison above has teeth."""
    nx, ny, nu = 4, 4, 11
    ang = np.array([0.3, 1.2])
    A = dd_ref.overlap_brute_2d(nx, ny, ang, nu, 1.0, 0.0, 1.0)
    img = np.ones((nx, ny))
    good = dd_ref.par_2d(img, ang, nu)
    orig = dd_ref.par_geom

    def bad_geom(*a):
        xd, pdrv, porth, scale, rev = orig(*a)
        return xd, pdrv, porth, 1.0, rev
    dd_ref.par_geom = bad_geom
    try:
        bad = dd_ref.par_2d(img, ang, nu)
    finally:
        dd_ref.par_geom = orig
    assert np.allclose(good.reshape(-1), A @ img.reshape(-1))
    assert not np.allclose(bad.reshape(-1), A @ img.reshape(-1))


def test_cone_stale_buffer_only_matters_at_the_border():
    """Cone DD: the reference's partial clear of its u buffer (dd.py:317)
    leaves the previous orthogonal line's edge bins in place, and those hold
    the contributions of that line's end pixels.  For an object that does
    not touch the volume border they are zero and the literal sweep equals
    the fully-cleared one; for a dense volume the literal sweep adds them."""
    rng = np.random.default_rng(3)
    img = np.zeros((8, 8, 4))
    img[2:6, 2:6, 1:3] = rng.uniform(0.5, 1, (4, 4, 2))
    ang = np.array([0.2, 1.0, 2.3])
    DSO, DSD = 40.0, 60.0
    fixed = dd_ref.cone_fp(img, ang, 40, 16, DSO, DSD, du=1.0, dv=1.0, sv=0.0)
    stale = dd_ref.cone_fp(img, ang, 40, 16, DSO, DSD, du=1.0, dv=1.0, sv=0.0, stale_buffer=True)
    assert np.all(fixed >= 0)
    assert np.array_equal(fixed, stale)
    dense = rng.uniform(0.5, 1, (8, 8, 4))
    f2 = dd_ref.cone_fp(dense, ang, 40, 16, DSO, DSD, sv=0.0)
    s2 = dd_ref.cone_fp(dense, ang, 40, 16, DSO, DSD, sv=0.0, stale_buffer=True)
    assert np.abs(f2 - s2).sum() > 0.01 * np.abs(f2).sum()
    # cone BP is the transpose of the fixed FP up to du*dv/(na*d_pix^3)
    y = rng.uniform(0, 1, fixed.shape)
    b = dd_ref.cone_bp(y, ang, img.shape, DSO, DSD, du=1.0, dv=1.0, sv=0.0)
    x = rng.uniform(0, 1, img.shape)
    fx = dd_ref.cone_fp(x, ang, 40, 16, DSO, DSD, du=1.0, dv=1.0, sv=0.0)
    lhs = np.dot(fx.reshape(-1), y.reshape(-1))
    rhs = np.dot(x.reshape(-1), b.reshape(-1)) * ang.size * 1.0  # BP scale: du*dv/(na*d_pix^3)
    assert abs(lhs - rhs) <= 1e-9 * abs(lhs)


@pytest.mark.parametrize("geom", ["par", "fan", "arc"])
def test_sparse_overlap_matrix_equals_the_brute_force_one(geom):
    """overlap_sparse_2d (used at workload size) is overlap_brute_2d."""
    rng = np.random.default_rng(11)
    ang = np.concatenate([rng.uniform(0, 2 * np.pi, 7), np.arange(8) * np.pi / 4]).astype(np.float32)
    kw = {} if geom == "par" else dict(DSO=60.0, DSD=95.0)
    A = dd_ref.overlap_brute_2d(13, 11, ang, 30, 1.1, 0.2, 0.9, geom=geom, **kw)
    S = dd_ref.overlap_sparse_2d(13, 11, ang, 30, 1.1, 0.2, 0.9, geom=geom, **kw).toarray()
    assert np.abs(A).max() > 0
    assert np.allclose(A, S, rtol=0, atol=1e-14)
