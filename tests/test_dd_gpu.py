"""GPU parity of the distance-driven projectors (libvir_amd via its C ABI).

Against the float64 oracle (oracle/dd_ref.py, itself pinned by
tests/test_dd_oracle.py): forward projections and the reference-scaled
backprojections within 1e-5 relative L2 error (float32 arithmetic against
float64), the equiangular-fan variant against its brute-force system
matrix, and the dot-product adjoint test for every geometry.
"""
import numpy as np
import pytest

from oracle import dd_ref

pytestmark = pytest.mark.gpu
TOL = 1e-5  # relative L2 error, float32 vs float64


def rel(a, b):
    a = np.asarray(a, np.float64).reshape(-1)
    b = np.asarray(b, np.float64).reshape(-1)
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30)


def _phantom(shape, rng, border=True):
    img = rng.uniform(0.2, 1.0, shape).astype(np.float32)
    if not border:
        m = np.zeros(shape, bool)
        sl = tuple(slice(1, n - 1) if n > 2 else slice(0, n) for n in shape)
        m[sl] = True
        img = img * m
    return img


@pytest.mark.parametrize("case", [
    dict(nx=24, ny=20, nu=41, na=17, du=1.0, su=0.0, d_pix=1.0),
    dict(nx=17, ny=31, nu=64, na=12, du=0.7, su=0.3, d_pix=1.3),
])
def test_dd_par_2d(case):
    from vir_amd.sys_mat import dd
    rng = np.random.default_rng(0)
    img = _phantom((case["nx"], case["ny"]), rng)
    ang = np.sort(rng.uniform(0, np.pi, case["na"])).astype(np.float32)
    kw = dict(du=case["du"], su=case["su"], d_pix=case["d_pix"])
    g = dd.dd_fp_par_2d(img, ang, case["nu"], **kw)
    r = dd_ref.par_2d(img, ang, case["nu"], **kw)
    assert rel(g, r) <= TOL
    y = rng.uniform(0, 1, r.shape).astype(np.float32)
    gb = dd.dd_bp_par_2d(y, ang, img.shape, **kw)
    rb = dd_ref.par_2d(y, ang, img.shape, bp=True, **kw)
    assert rel(gb, rb) <= TOL


@pytest.mark.parametrize("detector", ["flat", "arc"])
def test_dd_fan_2d(detector):
    from vir_amd.sys_mat import dd
    rng = np.random.default_rng(1)
    nx, ny, nu, na = 32, 28, 60, 24
    DSO, DSD = 60.0, 95.0
    img = _phantom((nx, ny), rng)
    ang = np.linspace(0, 2 * np.pi, na, endpoint=False).astype(np.float32) + np.float32(0.013)
    g = dd.dd_fp_fan_2d(img, ang, nu, DSO, DSD, du=1.1, su=0.2, d_pix=0.9, detector=detector)
    if detector == "flat":
        r = dd_ref.fan_2d(img, ang, nu, DSO, DSD, du=1.1, su=0.2, d_pix=0.9)
    else:
        A = dd_ref.overlap_brute_2d(nx, ny, ang, nu, 1.1, 0.2, 0.9, geom="arc",
                                    DSO=DSO, DSD=DSD)
        r = (A @ img.reshape(-1).astype(np.float64)).reshape(na, nu)
    assert rel(g, r) <= TOL
    if detector == "flat":
        y = rng.uniform(0, 1, r.shape).astype(np.float32)
        gb = dd.dd_bp_fan_2d(y, ang, (nx, ny), DSO, DSD, du=1.1, su=0.2, d_pix=0.9)
        rb = dd_ref.fan_2d(y, ang, (nx, ny), DSO, DSD, du=1.1, su=0.2, d_pix=0.9, bp=True)
        assert rel(gb, rb) <= TOL


def test_dd_fan_2d_diagonal_views_narrow_detector():
    """Views at exact multiples of 45 degrees (float32 angles, where
    |sin| == |cos| decides the driving axis as in the reference's float32
    np.cos) and a detector narrower than the image's shadow."""
    from vir_amd.sys_mat import dd
    rng = np.random.default_rng(5)
    n, nu, na, du, DSO, DSD = 20, 24, 24, 1.2, 50.0, 80.0
    img = _phantom((n, n), rng)
    ang = np.linspace(0, 2 * np.pi, na, endpoint=False).astype(np.float32)
    g = dd.dd_fp_fan_2d(img, ang, nu, DSO, DSD, du=du)
    r = dd_ref.fan_2d(img, ang, nu, DSO, DSD, du=du)
    A = dd_ref.overlap_brute_2d(n, n, ang, nu, du, 0.0, 1.0, geom="fan", DSO=DSO, DSD=DSD)
    assert rel(r, A @ img.reshape(-1).astype(np.float64)) < 1e-12
    assert rel(g, r) <= TOL
    y = rng.uniform(0, 1, r.shape).astype(np.float32)
    gb = dd.dd_bp_fan_2d(y, ang, (n, n), DSO, DSD, du=du)
    rb = dd_ref.fan_2d(y, ang, (n, n), DSO, DSD, du=du, bp=True)
    assert rel(gb, rb) <= TOL


@pytest.mark.parametrize("geom,detector,shape,nu,du,d_pix", [
    ("par", "flat", (300, 260), 600, 0.9, 0.7),
    ("par", "flat", (1200, 40), 300, 1.5, 0.5),    # a bin tile's pixel band spans three chunks
    ("fan", "flat", (300, 260), 700, 0.6, 0.7),
    ("fan", "arc", (300, 260), 700, 0.6, 0.7),
])
def test_dd_2d_tiled_forward_equals_per_bin(geom, detector, shape, nu, du, d_pix):
    """The LDS-tiled forward kernel (projected pixel boundaries of each line
    evaluated once per 256-bin tile) gives the same bits as the
    one-thread-per-bin gather (VIR_PER_OUTPUT), 45-degree views included."""
    import torch
    from vir_amd import _lib
    from vir_amd.device import stream
    from vir_amd.sys_mat.dd import DDProjector
    rng = np.random.default_rng(9)
    ang = np.concatenate([rng.uniform(0, 2 * np.pi, 11), np.arange(8) * np.pi / 4]).astype(np.float32)
    kw = dict(DSO=400.0, DSD=700.0) if geom == "fan" else {}
    P = DDProjector(shape, ang, nu, geom, du=du, su=0.3, d_pix=d_pix, detector=detector, **kw)
    x = torch.from_numpy(rng.uniform(0, 1, shape).astype(np.float32)).cuda()
    tiled = P.fp(x)
    ref = torch.empty_like(tiled)
    _lib.check(_lib.lib().vir_dd2d_fp(P.g, _lib.ptr(P.cos), _lib.ptr(P.sin), _lib.ptr(x), _lib.ptr(ref),
                                      P.fp_scale, _lib.VIR_PER_OUTPUT, stream()), "fp")
    a, b = tiled.cpu().numpy(), ref.cpu().numpy()
    assert np.abs(b).max() > 0
    assert np.array_equal(a.view(np.int32), b.view(np.int32))


def test_dd_cone_3d():
    from vir_amd.sys_mat import dd
    rng = np.random.default_rng(2)
    shape = (12, 10, 8)
    img = _phantom(shape, rng, border=False)
    ang = np.linspace(0, 2 * np.pi, 9, endpoint=False).astype(np.float32) + np.float32(0.1)
    DSO, DSD = 40.0, 70.0
    kw = dict(du=1.2, dv=1.1, su=0.1, d_pix=1.0)
    g = dd.dd_fp_cone_3d(img, ang, 30, 20, DSO, DSD, sv=0.0, **kw)
    r = dd_ref.cone_fp(img, ang, 30, 20, DSO, DSD, sv=0.0, **kw)
    assert rel(g, r) <= TOL
    # the reference's default detector offsets differ between FP (sv=1) and BP (sv=0)
    g1 = dd.dd_fp_cone_3d(img, ang, 30, 20, DSO, DSD, **kw)
    r1 = dd_ref.cone_fp(img, ang, 30, 20, DSO, DSD, **kw)
    assert rel(g1, r1) <= TOL
    y = rng.uniform(0, 1, r.shape).astype(np.float32)
    gb = dd.dd_bp_cone_3d(y, ang, shape, DSO, DSD, **kw)
    rb = dd_ref.cone_bp(y, ang, shape, DSO, DSD, **kw)
    assert rel(gb, rb) <= TOL


@pytest.mark.parametrize("geom,detector", [("par", "flat"), ("fan", "flat"), ("fan", "arc"),
                                           ("cone", "flat")])
def test_dd_exact_adjoint(geom, detector):
    import torch
    from vir_amd.sys_mat.dd import DDProjector
    rng = np.random.default_rng(3)
    if geom == "cone":
        shape, nu, nv = (20, 18, 12), 40, 24
    else:
        shape, nu, nv = (40, 36), 70, 1
    ang = rng.uniform(0, 2 * np.pi, 15)
    P = DDProjector(shape, ang, nu, geom, nv=nv, DSO=80.0, DSD=120.0, du=1.0, dv=1.0, su=0.15,
                    sv=-0.2, d_pix=1.0, detector=detector)
    x = torch.from_numpy(rng.uniform(0, 1, shape).astype(np.float32)).cuda()
    y = torch.from_numpy(rng.uniform(0, 1, P.sino_shape).astype(np.float32)).cuda()
    ax = P.fp(x).double()
    aty = P.bp(y).double()
    lhs = float((ax * y.double()).sum())
    rhs = float((x.double() * aty).sum())
    assert abs(lhs - rhs) <= 1e-5 * abs(lhs)
    # deterministic
    assert torch.equal(P.bp(y), P.bp(y)) and torch.equal(P.fp(x), P.fp(x))
