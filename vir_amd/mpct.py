"""Host marshalling between numpy / siddons() output and C ray structs.

Same call surface as the reference's mpct.py (mpct.py:76-349), which the
ctypes drop-ins (vir_amd.projection.proj_ctype) and their callers use:

  ctypes_vars(var)            numpy array -> (array, C pointer): float16 /
                              float64 become a float32 copy, int8/16/64 and
                              unsigned ints an int32 copy, float32 / int32
                              are passed without a copy; Python float / int
                              and numpy integers become c_float / c_int
                              (mpct.py:76-151)
  RaveledRay, UnraveledRay, Coord
                              the ray records of projection/projection_c.c:
                              52-72 (mpct.py:154-172)
  ctypes_coord, ctypes_raveled_ray, ctypes_unraveled_ray, list_ctypes_object
                              siddons() object arrays -> arrays of those
                              records (mpct.py:174-263)
  init_shared_array, shared_to_numpy, create_shared_array
                              the shared-memory arrays of the multiprocessing
                              path (mpct.py:267-349)

The records only point at the numpy arrays of the sdlist.  Where the
reference converts a ray's arrays to int32/float32 and keeps only the
pointer (its copy is then freed while the pointer lives on), the arrays
built here keep the converted copies alive for as long as they exist.
"""
import ctypes
import multiprocessing as mp

import numpy as np

from .sys_mat.siddon_lists import list_count_elem

c_int_p = ctypes.POINTER(ctypes.c_int)
c_float_p = ctypes.POINTER(ctypes.c_float)

shared_arr = {}


def c_int(var):
    """ctypes int of an integer (mpct.py:41-54)."""
    return ctypes.c_int(var)


def c_float(var):
    """ctypes float of a number (mpct.py:57-73)."""
    return ctypes.c_float(var)


_AS_F32 = (np.float16, np.float64)
_AS_I32 = (np.int8, np.int16, np.int64, np.uint8, np.uint16, np.uint32, np.uint64)


def ctypes_vars(var):
    """(var, C pointer or value) with the reference's conversions; the
    element type is judged from the first element, as there."""
    if isinstance(var, np.ndarray):
        first = var.flat[0]
        if isinstance(first, _AS_F32):
            var = var.astype(np.float32)
            return var, var.ctypes.data_as(c_float_p)
        if isinstance(first, _AS_I32):
            var = var.astype(np.int32)
            return var, var.ctypes.data_as(c_int_p)
        if isinstance(first, np.float32):
            return var, var.ctypes.data_as(c_float_p)
        if isinstance(first, np.int32):
            return var, var.ctypes.data_as(c_int_p)
        raise ValueError(var.dtype, " is not yet supported in mpct")
    if isinstance(var, float):
        return var, ctypes.c_float(var)
    if isinstance(var, int):
        return var, ctypes.c_int(var)
    if isinstance(var, np.integer):
        var = int(var)
        return var, ctypes.c_int(var)
    raise ValueError(type(var), " is not yet supported in mpct")


class UnraveledRay(ctypes.Structure):
    """projection_c.c:52-58: n entries at (X, Y, Z) with lengths L."""
    _fields_ = [("n", ctypes.c_int), ("X", c_int_p), ("Y", c_int_p), ("Z", c_int_p),
                ("L", c_float_p)]


class RaveledRay(ctypes.Structure):
    """projection_c.c:61-65: n entries at flat indices R with lengths L."""
    _fields_ = [("n", ctypes.c_int), ("R", c_int_p), ("L", c_float_p)]


class Coord(ctypes.Structure):
    """projection_c.c:68-72: grid dimensions."""
    _fields_ = [("x", ctypes.c_int), ("y", ctypes.c_int), ("z", ctypes.c_int)]


def ctypes_coord(x, y, z):
    return ctypes.byref(Coord(ctypes.c_int(x), ctypes.c_int(y), ctypes.c_int(z)))


def _i32(a, keep):
    a, p = ctypes_vars(np.asarray(a) if not isinstance(a, np.ndarray) else a)
    keep.append(a)
    return ctypes.cast(p, c_int_p)


def _f32(a, keep):
    a, p = ctypes_vars(np.asarray(a) if not isinstance(a, np.ndarray) else a)
    keep.append(a)
    return ctypes.cast(p, c_float_p)


def ctypes_unraveled_ray(n, X, Y, Z, L, byref=True):
    keep = []
    r = UnraveledRay(ctypes.c_int(n), _i32(X, keep), _i32(Y, keep), _i32(Z, keep), _f32(L, keep))
    r._keep = keep
    return ctypes.byref(r) if byref else r


def ctypes_raveled_ray(n, R, L, byref=True):
    keep = []
    r = RaveledRay(ctypes.c_int(n), _i32(R, keep), _f32(L, keep))
    r._keep = keep
    return ctypes.byref(r) if byref else r


def list_ctypes_object(sdlist, ravel=False, flat=True):
    """siddons() object array -> C ray records (mpct.py:206-263).

    flat=True: byref of one contiguous array of RaveledRay (ravel=True) or
    UnraveledRay records in the sdlist's flat order, counts from
    list_count_elem and null pointers for rays that miss -- what
    forward_proj_*_rays_struct take.  flat=False: an object array of the
    sdlist's shape holding byref records (None for misses)."""
    sdlist = np.asarray(sdlist, dtype=object)
    if flat:
        count = np.asarray(list_count_elem(sdlist)).reshape(-1)
        rec = RaveledRay if ravel else UnraveledRay
        arr = (rec * int(np.prod(sdlist.shape)))()
        keep = []
        for k, ray in enumerate(sdlist.reshape(-1)):
            arr[k].n = int(count[k])
            if ray is None:
                continue
            if ravel:
                arr[k].R = _i32(ray[0], keep)
                arr[k].L = _f32(ray[1], keep)
            else:
                arr[k].X = _i32(ray[0], keep)
                arr[k].Y = _i32(ray[1], keep)
                arr[k].Z = _i32(ray[2], keep)
                arr[k].L = _f32(ray[3], keep)
        arr._keep = keep
        return ctypes.byref(arr)
    count = list_count_elem(sdlist)
    out = np.empty(sdlist.shape, dtype=object)
    for idx, ray in np.ndenumerate(sdlist):
        if ray is None:
            continue
        if ravel:
            out[idx] = ctypes_raveled_ray(int(count[idx]), ray[0], ray[1])
        else:
            out[idx] = ctypes_unraveled_ray(int(count[idx]), ray[0], ray[1], ray[2], ray[3])
    return out


def init_shared_array(arr, shape, dtype=float, l=None):  # noqa: E741 (the reference's name)
    """Pool initializer: publish a shared array to the worker (mpct.py:267-274)."""
    global shared_arr
    shared_arr = {"arr": arr, "shape": shape, "dtype": dtype, "lock": l}


def shared_to_numpy(arr_shared, shape, dtype=float):
    """numpy view of a shared ctypes buffer, no copy (mpct.py:277-299)."""
    return np.frombuffer(arr_shared, dtype=dtype).reshape(shape)


def create_shared_array(shape, lock=True, dtype=float):
    """A shared-memory array and its numpy view (and its lock when
    lock=True); values are not initialised (mpct.py:302-349)."""
    dtype = np.dtype(dtype)
    ctype = np.ctypeslib.as_ctypes_type(dtype)
    size = int(np.prod(shape))
    if lock:
        lk = mp.Lock()
        arr = mp.Array(ctype, size, lock=lk)
        return arr, shared_to_numpy(arr.get_obj(), shape, dtype), lk
    arr = mp.Array(ctype, size, lock=False)
    return arr, shared_to_numpy(arr, shape, dtype)
