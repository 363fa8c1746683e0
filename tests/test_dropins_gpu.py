"""GPU drop-ins for the reference's ctypes and multiprocessing projection
calls (projection/proj_ctype.py:27-90, projection/proj_mp.py:43-96).

  sd_f_proj_c / sd_b_proj_c: bit for bit the compiled reference
      projection_c.c (oracle/_ref) called with the same ray records and
      buffers -- ravel and unravel records, numpy buffers and raw pointers
      (C=True), adding into nonzero buffers; a float64 sino is converted to
      a copy as mpct.ctypes_vars does, so the caller's array is unchanged;
  mp_sd_b_proj: bit for bit the reference's numpy backprojection loop in ray
      order (oracle.projection.sd_b_proj_numpy), repeated voxels included;
  mp_sd_f_proj: float64 products within 1e-6, with the reference's dtypes.
"""
import ctypes

import numpy as np
import pytest

from oracle import projection as op
from oracle import siddon as osd

pytestmark = pytest.mark.gpu


def _workload():
    from vir_amd.proj_geom import geom_circular
    v = np.arange(10) * np.pi / 10
    du = np.arange(24) - 11.5
    dv = np.arange(6) - 2.5
    s, t = geom_circular(du, v, "cone", src_iso=40.0, det_iso=30.0, DetsZ=dv)
    t[0, 0, 0] = t[0, 0, 0] + [0.0, 0.0, 500.0]          # one ray that misses the grid
    return s, t, (16, 14, 6)


def _ref_c():
    try:
        return op.RefProjection().so
    except FileNotFoundError:
        pytest.skip("oracle/_ref not built")


@pytest.mark.parametrize("ravel", [True, False])
@pytest.mark.parametrize("C", [False, True])
def test_ctypes_dropins_equal_the_compiled_reference(ravel, C):
    from vir_amd import mpct
    from vir_amd.projection.proj_ctype import sd_b_proj_c, sd_f_proj_c
    so = _ref_c()
    s, t, nP = _workload()
    sd = osd.siddons(s, t, nP, 1.0, 0.0, ravel=ravel)
    assert sd.reshape(-1)[0] is None
    recs = mpct.list_ctypes_object(sd, ravel=ravel, flat=True)
    nr = sd.size
    rng = np.random.default_rng(1)
    vol = rng.uniform(0, 1, nP).astype(np.float32)
    sino0 = rng.uniform(0, 1, sd.shape).astype(np.float32)     # the calls add into it
    vol0 = rng.uniform(0, 1, nP).astype(np.float32)
    fp_name = "forward_proj_ravel_rays_struct" if ravel else "forward_proj_unravel_rays_struct"
    bp_name = "back_proj_ravel_rays_struct" if ravel else "back_proj_unravel_rays_struct"
    dims = mpct.ctypes_coord(*nP)
    for name in (fp_name, bp_name):
        getattr(so, name).argtypes = None
    fptr = lambda a: a.ctypes.data_as(mpct.c_float_p)  # noqa: E731
    # the reference
    s_ref = sino0.copy()
    v_ref = vol0.copy()
    args = (recs, ctypes.c_int(nr)) if ravel else (recs, dims, ctypes.c_int(nr))
    getattr(so, fp_name)(fptr(vol), fptr(s_ref), *args)
    getattr(so, bp_name)(fptr(v_ref), fptr(s_ref), *args)
    # the drop-ins
    s_got = sino0.copy()
    v_got = vol0.copy()
    if C:
        sd_f_proj_c(fptr(vol), fptr(s_got), recs, ravel=ravel, C=True, dims=nP, nRays=nr)
        sd_b_proj_c(fptr(v_got), fptr(s_got), recs, ravel=ravel, C=True, dims=nP, nRays=nr)
    else:
        assert sd_f_proj_c(vol, s_got, recs, ravel=ravel) is None
        sd_b_proj_c(v_got, s_got, recs, ravel=ravel)
    assert np.array_equal(s_got.view(np.int32), s_ref.view(np.int32))
    assert np.array_equal(v_got.view(np.int32), v_ref.view(np.int32))
    assert not np.array_equal(s_got, sino0)


def test_float64_sino_is_converted_to_a_copy():
    from vir_amd import mpct
    from vir_amd.projection.proj_ctype import sd_f_proj_c
    s, t, nP = _workload()
    sd = osd.siddons(s, t, nP, 1.0, 0.0, ravel=True)
    recs = mpct.list_ctypes_object(sd, ravel=True, flat=True)
    sino = np.zeros(sd.shape, np.float64)
    sd_f_proj_c(np.ones(nP, np.float32), sino, recs)
    assert not sino.any()       # as in the reference: the result went to the float32 copy


def test_mp_dropins():
    from vir_amd.projection.proj_mp import mp_sd_b_proj, mp_sd_f_proj
    s, t, nP = _workload()
    sd = osd.siddons(s, t, nP, 1.0, 0.0)
    ind, ln, off = osd.siddons(s, t, nP, 1.0, 0.0, flat=True)
    rng = np.random.default_rng(2)
    vol = rng.uniform(0, 1, nP).astype(np.float32)
    f = mp_sd_f_proj(vol, sd, cpus=3)
    assert f.shape == sd.shape and f.dtype == np.float64          # a ray is None
    f64 = op.fp64(vol, ind, ln, off).reshape(sd.shape)
    assert np.max(np.abs(f - f64)) <= 1e-6 * np.abs(f64).max()
    hit = sd[1:]
    assert mp_sd_f_proj(vol, hit).dtype == np.float32             # every ray hits
    y = rng.uniform(0, 1, sd.shape).astype(np.float32)
    b = mp_sd_b_proj(y, sd, nP)
    ref = op.sd_b_proj_numpy(y, ind, ln, off, nP)
    assert b.dtype == np.float32 and b.shape == nP
    assert np.array_equal(b.view(np.int32), ref.view(np.int32))
