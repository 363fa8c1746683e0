"""GPU square-geometry DD projectors and the sirt_dd driver against the
float64 oracle (oracle/ddsq_ref.py): forward and backward within 1e-5
relative L2, exact transpose, SIRT iterate within 1e-5 of the float64
recurrence."""
import numpy as np
import pytest

from oracle import ddsq_ref, sirt_ref

pytestmark = pytest.mark.gpu
TOL = 1e-5
KW = dict(nu=17, nv=9, nsrc_p=4, nsrc_z=2, DSO=14.0, DSD=24.0, du=1.1, dv=0.9, dsrc_p=0.7,
          dsrc_z=0.6, su=0.2, sv=-0.3, ssrc_p=0.1, ssrc_z=-0.2, d_pix=0.8)
SHAPE = (10, 10, 6)


def rel(a, b):
    a = np.asarray(a, np.float64).reshape(-1)
    b = np.asarray(b, np.float64).reshape(-1)
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def test_square_fp_bp_match_oracle_and_transpose():
    from vir_amd.sys_mat import dd
    rng = np.random.default_rng(0)
    img = rng.uniform(0, 1, SHAPE).astype(np.float32)
    W = ddsq_ref.matrix(SHAPE, **KW)
    sino_shape = (KW["nsrc_p"], KW["nsrc_z"], KW["nv"], KW["nu"], 4)
    sino = dd.dd_fp_square(img, **KW)
    assert sino.shape == sino_shape and sino.dtype == np.float32
    assert rel(sino, ddsq_ref.fp(W, img, KW["d_pix"], KW["du"], KW["dv"], sino_shape)) <= TOL
    y = rng.normal(size=sino_shape).astype(np.float32)
    kw = {k: v for k, v in KW.items() if k not in ("nu", "nv", "nsrc_p", "nsrc_z")}
    back = dd.dd_bp_square(y, SHAPE, **kw)
    assert back.shape == SHAPE
    assert rel(back, ddsq_ref.bp(W, y, SHAPE)) <= TOL
    c = KW["d_pix"] ** 2 / (KW["du"] * KW["dv"])
    lhs = float((sino.astype(np.float64) * y).sum()) / c
    rhs = float((img.astype(np.float64) * back).sum())
    assert abs(lhs - rhs) <= TOL * abs(lhs)


def test_square_rejects_non_square():
    from vir_amd.sys_mat import dd
    with pytest.raises(ValueError):
        dd.dd_fp_square(np.zeros((6, 5, 3), np.float32), **KW)


def test_sirt_dd_dropin_matches_fp64():
    from vir_amd.sys_mat import dd
    from vir_amd.sys_mat.sirt_dd import sirt
    rng = np.random.default_rng(1)
    x_true = np.zeros(SHAPE, np.float32)
    x_true[3:7, 2:8, 1:5] = rng.uniform(0.5, 1.0, (4, 6, 4))
    kw = {k: v for k, v in KW.items() if k not in ("nu", "nv", "nsrc_p", "nsrc_z")}
    sino = dd.dd_fp_square(x_true, **KW)
    W_sino = dd.dd_fp_square(np.ones(SHAPE, np.float32), **KW)
    W_img = dd.dd_bp_square(np.ones(sino.shape, np.float32), SHAPE, **kw)
    gem = (KW["DSO"], KW["DSD"])
    img_p = SHAPE + (KW["d_pix"],)
    src = (KW["nsrc_p"], KW["nsrc_z"], KW["dsrc_p"], KW["dsrc_z"], KW["ssrc_p"], KW["ssrc_z"])
    det = (KW["nu"], KW["nv"], KW["du"], KW["dv"], KW["su"], KW["sv"])
    x = sirt(sino, W_sino, W_img, gem, img_p, src, det, (1.0, 12))
    c = KW["d_pix"] ** 2 / (KW["du"] * KW["dv"])
    A = c * ddsq_ref.matrix(SHAPE, **KW)
    ref = sirt_ref.sirt(A, sino, 11, x0=np.full(int(np.prod(SHAPE)), np.float32(1e-5)))
    assert rel(x, ref) <= TOL
    assert rel(x, x_true) < rel(np.full(SHAPE, 1e-5), x_true)
