"""CPU: vir_amd.mpct, the host marshalling of the ctypes drop-ins.

The conversions of ctypes_vars (mpct.py:76-151), the ray records of
list_ctypes_object (mpct.py:206-263) and the shared-array helpers; the
records are layout-compatible with the reference's compiled projection_c.c
(oracle/_ref), whose own forward and backprojection read them correctly.
"""
import ctypes
import gc

import numpy as np
import pytest

from oracle import projection as op
from oracle import siddon as osd
from vir_amd import mpct


def test_ctypes_vars_conversions():
    for dt in (np.float16, np.float64):
        a = np.arange(4, dtype=dt)
        b, p = mpct.ctypes_vars(a)
        assert b.dtype == np.float32 and b is not a and isinstance(p, mpct.c_float_p)
    for dt in (np.int8, np.int16, np.int64, np.uint8, np.uint16, np.uint32, np.uint64):
        a = np.arange(4, dtype=dt)
        b, p = mpct.ctypes_vars(a)
        assert b.dtype == np.int32 and isinstance(p, mpct.c_int_p)
    a = np.arange(4, dtype=np.float32)
    b, p = mpct.ctypes_vars(a)
    assert b is a and p[2] == 2.0                       # no copy
    p[2] = 7.0
    assert a[2] == 7.0
    a = np.arange(4, dtype=np.int32)
    assert mpct.ctypes_vars(a)[0] is a
    assert mpct.ctypes_vars(2.5)[1].value == 2.5
    assert mpct.ctypes_vars(3)[1].value == 3
    assert mpct.ctypes_vars(np.int64(5)) [0] == 5
    with pytest.raises(ValueError):
        mpct.ctypes_vars(np.zeros(2, np.complex64))
    with pytest.raises(ValueError):
        mpct.ctypes_vars("x")


def _sdlist(ravel):
    s = np.array([[-20.0, 0.3, 0.2], [-20.0, 1.7, 0.6], [-20.0, 30.0, 0.1]], np.float32)
    t = np.array([[20.0, -0.4, 0.1], [20.0, 2.2, -0.3], [20.0, 30.0, 0.1]], np.float32)
    return osd.siddons(s, t, (6, 5, 2), 1.0, 0.0, ravel=ravel), osd.siddons(s, t, (6, 5, 2), 1.0, 0.0,
                                                                           flat=True)


@pytest.mark.parametrize("ravel", [True, False])
def test_list_ctypes_object_flat(ravel):
    sd, (ind, ln, off) = _sdlist(ravel)
    assert sd[2] is None                                  # the third ray misses
    arr = mpct.list_ctypes_object(sd, ravel=ravel, flat=True)._obj
    assert len(arr) == 3
    gc.collect()
    for r in range(3):
        n = arr[r].n
        assert n == off[r + 1] - off[r]
        if n == 0:
            assert not bool(arr[r].L)
            continue
        L = np.ctypeslib.as_array(arr[r].L, shape=(n,))
        assert np.array_equal(L, ln[off[r]:off[r + 1]])
        if ravel:
            R = np.ctypeslib.as_array(arr[r].R, shape=(n,))
        else:
            X, Y, Z = (np.ctypeslib.as_array(getattr(arr[r], k), shape=(n,)) for k in "XYZ")
            R = (X * 5 + Y) * 2 + Z
        assert np.array_equal(R, ind[off[r]:off[r + 1]])


def test_converted_copies_stay_alive():
    sd = np.empty(1, dtype=object)
    sd[0] = [np.array([3, 1], np.int64), np.array([0.5, 0.25], np.float64)]
    arr = mpct.list_ctypes_object(sd, ravel=True, flat=True)._obj
    del sd
    gc.collect()
    junk = [np.full(64, 99, np.int64) for _ in range(100)]   # reuse freed memory if any
    assert np.ctypeslib.as_array(arr[0].R, shape=(2,)).tolist() == [3, 1]
    assert np.ctypeslib.as_array(arr[0].L, shape=(2,)).tolist() == [0.5, 0.25]
    del junk


def test_list_ctypes_object_per_ray():
    sd, (ind, ln, off) = _sdlist(True)
    out = mpct.list_ctypes_object(sd, ravel=True, flat=False)
    assert out.shape == sd.shape and out[2] is None
    rec = out[0]._obj
    assert isinstance(rec, mpct.RaveledRay) and rec.n == off[1]


def test_shared_arrays():
    arr, view, lock = mpct.create_shared_array((3, 4), lock=True, dtype=np.float32)
    view[:] = 2.0
    again = mpct.shared_to_numpy(arr.get_obj(), (3, 4), np.float32)
    assert again.sum() == 24.0 and hasattr(lock, "acquire")
    arr2, view2 = mpct.create_shared_array((5,), lock=False, dtype=np.int32)
    view2[:] = 7
    mpct.init_shared_array(arr2, (5,), np.int32)
    assert mpct.shared_to_numpy(mpct.shared_arr["arr"], (5,), np.int32).sum() == 35


def test_records_drive_the_reference_c():
    """The compiled projection_c.c reads these records: its forward
    projection equals the float64 product to float32 rounding."""
    try:
        ref = op.RefProjection()
    except FileNotFoundError:
        pytest.skip("oracle/_ref not built (reference sources absent)")
    sd, (ind, ln, off) = _sdlist(True)
    arr = mpct.list_ctypes_object(sd, ravel=True, flat=True)
    rng = np.random.default_rng(0)
    vol = rng.uniform(0, 1, (6, 5, 2)).astype(np.float32)
    sino = np.zeros(3, np.float32)
    fn = ref.so.forward_proj_ravel_rays_struct
    fn.argtypes = None
    fn(vol.ctypes.data_as(mpct.c_float_p), sino.ctypes.data_as(mpct.c_float_p), arr, ctypes.c_int(3))
    assert np.allclose(sino, op.fp64(vol, ind, ln, off), rtol=1e-6, atol=0)
    assert sino[2] == 0.0 and sino[0] > 0
