"""Pins of the square-geometry DD oracle (oracle/ddsq_ref.py).

* the matrix W (sides from the index map) equals the literal restatement of
  the reference's forward loop over its transposed copies imgX / imgY;
* W^T equals the literal restatement of the reference's backward loop (an
  independent reading of its side mapping and final transposes);
* a single voxel's footprint has total weight (M d_pix)^2 per source, M the
  magnification of its slice's lower boundary, when it lies on the detector;
* rotating the volume by 90 degrees permutes the sides.
"""
import numpy as np

from oracle import ddsq_ref

KW = dict(nu=11, nv=7, nsrc_p=3, nsrc_z=2, DSO=12.0, DSD=20.0, du=1.1, dv=0.9, dsrc_p=0.7,
          dsrc_z=0.6, su=0.2, sv=-0.3, ssrc_p=0.1, ssrc_z=-0.2, d_pix=0.8)
SHAPE = (5, 5, 3)


def _W():
    return ddsq_ref.matrix(SHAPE, **KW)


def test_matrix_equals_literal_forward_loop():
    rng = np.random.default_rng(0)
    img = rng.uniform(0, 1, SHAPE)
    W = _W()
    sino_shape = (KW["nsrc_p"], KW["nsrc_z"], KW["nv"], KW["nu"], 4)
    a = ddsq_ref.fp(W, img, KW["d_pix"], KW["du"], KW["dv"], sino_shape)
    b = ddsq_ref.fp_loop(img, **KW)
    np.testing.assert_allclose(a, b, rtol=1e-12, atol=1e-12)


def test_transpose_equals_literal_backward_loop():
    rng = np.random.default_rng(1)
    sino = rng.normal(size=(KW["nsrc_p"], KW["nsrc_z"], KW["nv"], KW["nu"], 4))
    W = _W()
    kw = {k: v for k, v in KW.items() if k not in ("nu", "nv", "nsrc_p", "nsrc_z")}
    np.testing.assert_allclose(ddsq_ref.bp(W, sino, SHAPE), ddsq_ref.bp_loop(sino, SHAPE, **kw),
                               rtol=1e-12, atol=1e-12)


def test_single_voxel_footprint_area():
    kw = dict(KW, nu=60, nv=40, du=0.5, dv=0.5, su=0.0, sv=0.0, ssrc_p=0.0, ssrc_z=0.0,
              nsrc_p=1, nsrc_z=1)
    W = ddsq_ref.matrix(SHAPE, **kw)
    img = np.zeros(SHAPE)
    img[2, 1, 1] = 1.0
    raw = (W @ img.reshape(-1)).reshape(1, 1, 40, 60, 4)
    pb = ddsq_ref.boundspace(5, 0.8)
    # side 0: io = ix = 2; side 1: io = iy = 1; side 2: io = 5-1-2; side 3: io = 5-1-1
    for side, io in ((0, 2), (1, 1), (2, 2), (3, 3)):
        M = kw["DSD"] / (kw["DSO"] - pb[io])
        assert abs(raw[..., side].sum() - (M * 0.8) ** 2) < 1e-10


def test_rotation_permutes_sides():
    rng = np.random.default_rng(2)
    img = rng.uniform(0, 1, SHAPE)
    sino_shape = (KW["nsrc_p"], KW["nsrc_z"], KW["nv"], KW["nu"], 4)
    W = _W()
    a = ddsq_ref.fp(W, img, KW["d_pix"], KW["du"], KW["dv"], sino_shape)
    # side 1 of img reads img[n-1-ip, io]: side 0 of the image rotated so that
    # rot[io, ip] = img[n-1-ip, io]
    rot = np.stack([np.rot90(img[:, :, k], -1) for k in range(SHAPE[2])], -1)
    assert np.allclose(rot[1, 3, 0], img[5 - 1 - 3, 1, 0])
    b = ddsq_ref.fp(W, rot, KW["d_pix"], KW["du"], KW["dv"], sino_shape)
    np.testing.assert_allclose(b[..., 0], a[..., 1], rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(b[..., 1], a[..., 2], rtol=1e-12, atol=1e-12)
