"""ctypes binding of libvir_amd.so (include/vir_amd.h).

The library is built in-tree by vir_amd.build.  There is no fallback: if the
shared object is missing or a GPU is absent, calls fail loudly.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libvir_amd.so")

VIR_OK = 0
VIR_EINVAL = -1
VIR_EHIP = -2
VIR_ACCUMULATE = 1
VIR_SEQUENTIAL = 2
VIR_UNGUARDED = 4
VIR_PER_OUTPUT = 8
VIR_SF_UNIT_AMPLITUDE = 16
VIR_COMPAT = 32
SIDDON_RECORD_BYTES = 80  # include/vir_amd.h: VIR_SIDDON_RECORD_BYTES

P = ctypes.c_void_p
I64 = ctypes.c_int64
I32 = ctypes.c_int32
U32 = ctypes.c_uint32
F32 = ctypes.c_float


class VirGrid(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int32 * 3), ("d", ctypes.c_float * 3),
                ("o", ctypes.c_float * 3), ("eps", ctypes.c_double)]


G = ctypes.POINTER(VirGrid)


class VirDDGeom(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int32), ("ny", ctypes.c_int32), ("nz", ctypes.c_int32),
                ("nu", ctypes.c_int32), ("nv", ctypes.c_int32), ("na", ctypes.c_int32),
                ("mode", ctypes.c_int32),
                ("d_pix", ctypes.c_float), ("du", ctypes.c_float), ("dv", ctypes.c_float),
                ("su", ctypes.c_float), ("sv", ctypes.c_float), ("DSO", ctypes.c_float),
                ("DSD", ctypes.c_float)]


DG = ctypes.POINTER(VirDDGeom)


class VirSFGeom(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int32), ("ny", ctypes.c_int32), ("nz", ctypes.c_int32),
                ("nu", ctypes.c_int32), ("nv", ctypes.c_int32), ("na", ctypes.c_int32),
                ("nbatch", ctypes.c_int32), ("mode", ctypes.c_int32),
                ("d_pix", ctypes.c_float), ("du", ctypes.c_float), ("dv", ctypes.c_float),
                ("su", ctypes.c_float), ("sv", ctypes.c_float), ("DSO", ctypes.c_float),
                ("DSD", ctypes.c_float)]


SG = ctypes.POINTER(VirSFGeom)


class VirSpGeom(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int32), ("ny", ctypes.c_int32), ("nu", ctypes.c_int32),
                ("na", ctypes.c_int32), ("d_pix", ctypes.c_float), ("du", ctypes.c_float),
                ("DSO", ctypes.c_float), ("DSD", ctypes.c_float)]


class VirSepGeom(ctypes.Structure):
    _fields_ = [(k, ctypes.c_int32) for k in ("nx", "ny", "nz", "nu", "nv")] + \
               [(k, ctypes.c_double) for k in ("dx", "dy", "dz", "su", "sv", "Ds", "Dd")]


class VirScGrid(ctypes.Structure):
    _fields_ = [(k, ctypes.c_int32) for k in ("nx", "ny", "nz")] + \
               [(k, ctypes.c_float) for k in ("dx", "dy", "dz")]


SCG = ctypes.POINTER(VirScGrid)
SPG = ctypes.POINTER(VirSpGeom)
SEPG = ctypes.POINTER(VirSepGeom)


class VirRDGeom(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int32), ("ny", ctypes.c_int32), ("nz", ctypes.c_int32),
                ("nu", ctypes.c_int32), ("nv", ctypes.c_int32), ("na", ctypes.c_int32),
                ("mode", ctypes.c_int32), ("joseph", ctypes.c_int32),
                ("d_pix", ctypes.c_float), ("du", ctypes.c_float), ("dv", ctypes.c_float),
                ("su", ctypes.c_float), ("sv", ctypes.c_float), ("DSO", ctypes.c_float),
                ("DSD", ctypes.c_float)]


RG = ctypes.POINTER(VirRDGeom)


class VirDDSqGeom(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int32), ("nz", ctypes.c_int32), ("nu", ctypes.c_int32),
                ("nv", ctypes.c_int32), ("nsrc_p", ctypes.c_int32), ("nsrc_z", ctypes.c_int32)] + \
               [(k, ctypes.c_float) for k in ("d_pix", "du", "dv", "dsrc_p", "dsrc_z", "su", "sv",
                                              "ssrc_p", "ssrc_z", "DSO", "DSD")]


QG = ctypes.POINTER(VirDDSqGeom)

# name -> (restype, argtypes); the exported set of include/vir_amd.h
SIGNATURES = {
    "vir_version": (ctypes.c_int, []),
    "vir_last_error": (ctypes.c_char_p, []),
    "vir_siddon_count": (ctypes.c_int, [P, P, I64, G, P, P, P]),
    "vir_siddon_fill": (ctypes.c_int, [P, P, I64, G, P, P, P, P, P, P, P]),
    "vir_siddon_fp": (ctypes.c_int, [P, P, I64, G, P, P, U32, P, P]),
    "vir_fbp_filter": (ctypes.c_int, [P, I64, I32, P, P, P, P]),
    "vir_fbp_bp": (ctypes.c_int, [DG, P, P, P, P, F32, U32, P]),
    "vir_siddon_count_guarded": (ctypes.c_int, [P, P, I64, G, P, P]),
    "vir_siddon_bp_workspace": (I64, [G, I32]),
    "vir_siddon_bp": (ctypes.c_int, [P, P, I32, I32, I32, P, G, P, P, U32, P, P]),
    "vir_siddon_ray_records": (ctypes.c_int, [P, P, I64, G, P, P]),
    "vir_siddon_fp_rec": (ctypes.c_int, [P, I64, G, P, P, U32, P, P]),
    "vir_siddon_bp_rec": (ctypes.c_int, [P, I32, I32, I32, P, G, P, P, U32, P, P]),
    "vir_siddon_fp_rec_res": (ctypes.c_int, [P, I64, G, P, P, P, P, U32, P]),
    "vir_siddon_bp_rec_upd": (ctypes.c_int, [P, I32, I32, I32, P, G, P, P, F32, P, U32, P, P]),
    "vir_siddon_fp_rec_batch": (ctypes.c_int, [P, I64, G, I32, P, P, U32, P, P]),
    "vir_siddon_bp_batch_workspace": (I64, [G, I32, I32]),
    "vir_siddon_bp_rec_batch": (ctypes.c_int, [P, I32, I32, I32, P, G, I32, P, P, U32, P, P]),
    "vir_siddon_fp_rec_res_batch": (ctypes.c_int, [P, I64, G, I32, P, P, P, P, U32, P]),
    "vir_siddon_bp_rec_upd_batch": (ctypes.c_int, [P, I32, I32, I32, P, G, I32, P, P, F32, P, U32, P, P]),
    "vir_siddon_bp_rec_noalloc": (ctypes.c_int, [P, I32, I32, I32, P, G, P, P, U32, P]),
    "vir_siddons_c_count": (ctypes.c_int, [P, P, I64, SCG, P, P, P, P, P]),
    "vir_siddons_c_fill": (ctypes.c_int, [P, P, I64, SCG, P, P, P, P, P, P, P]),
    "vir_siddons_c_proj": (ctypes.c_int, [P, P, I64, SCG, P, P, P, P, P, P, P]),
    "vir_csr_fp": (ctypes.c_int, [P, P, P, I64, P, P, U32, P]),
    "vir_csr_transpose": (ctypes.c_int, [P, P, P, I64, I64, I64, P, P, P, P]),
    "vir_csc_bp": (ctypes.c_int, [P, P, P, I64, P, P, U32, P]),
    "vir_dd2d_fp": (ctypes.c_int, [DG, P, P, P, P, F32, U32, P]),
    "vir_dd2d_bp": (ctypes.c_int, [DG, P, P, P, P, F32, U32, P]),
    "vir_dd2d_fp_res": (ctypes.c_int, [DG, P, P, P, P, P, P, F32, U32, P]),
    "vir_dd2d_bp_upd": (ctypes.c_int, [DG, P, P, P, P, F32, P, F32, U32, P]),
    "vir_ddcone_fp": (ctypes.c_int, [DG, P, P, P, P, F32, U32, P]),
    "vir_ddcone_bp": (ctypes.c_int, [DG, P, P, P, P, F32, U32, P]),
    "vir_ddsq_fp": (ctypes.c_int, [QG, P, P, F32, U32, P]),
    "vir_ddsq_bp": (ctypes.c_int, [QG, P, P, F32, U32, P]),
    "vir_rd_fp": (ctypes.c_int, [RG, P, P, P, P, F32, U32, P]),
    "vir_rd_bp": (ctypes.c_int, [RG, P, P, P, P, F32, U32, P]),
    "vir_sf_fp": (ctypes.c_int, [SG, P, P, P, P, P, F32, U32, P]),
    "vir_sf_bp": (ctypes.c_int, [SG, P, P, P, P, P, F32, U32, P]),
    "vir_sp_fan_fp": (ctypes.c_int, [SPG, P, P, P, P, F32, U32, P]),
    "vir_sp_fan_bp": (ctypes.c_int, [SPG, P, P, P, P, F32, U32, P]),
    "vir_sep_workspace": (I64, [SEPG]),
    "vir_sep_fp": (ctypes.c_int, [SEPG, P, P, P, P]),
    "vir_sep_bp": (ctypes.c_int, [SEPG, P, P, P, P]),
    "vir_sirt_residual": (ctypes.c_int, [P, P, P, I64, P]),
    "vir_tv": (ctypes.c_int, [P, I64, I32, I32, I32, F32, F32, P, P, P, U32, P]),
    "vir_sqdist": (ctypes.c_int, [P, P, P, I64, P, P, P]),
    "vir_tv_chambolle": (ctypes.c_int, [I32, I32, I32, I32, P, P, P, P, F32, F32, I32, P, P]),
    "vir_pdhg_dual_tv": (ctypes.c_int, [I32, I32, I32, P, P, F32, P]),
    "vir_pdhg_dual_data": (ctypes.c_int, [I64, P, P, P, F32, P, P]),
    "vir_pdhg_primal": (ctypes.c_int, [I32, I32, I32, P, P, P, F32, P, P, P]),
    "vir_geom_circular": (ctypes.c_int, [I32, I64, P, P, I32, P, I32, P, ctypes.c_double,
                                         ctypes.c_double, P, P, P]),
    "vir_geom_square": (ctypes.c_int, [I64, P, P, I64, ctypes.c_double, P, P, P, P, P]),
    "vir_sirt_update": (ctypes.c_int, [P, P, P, F32, I64, P]),
    "vir_phantom": (ctypes.c_int, [I32, I32, I32, I32, P, P, P]),
}

_lib = None


class VirError(RuntimeError):
    pass


def lib():
    """Load libvir_amd.so (once) with every signature declared."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise VirError(f"{LIB_PATH} not built: run `python -m vir_amd.build`")
        handle = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        _lib = handle
    return _lib


def check(rc, what=""):
    if rc != VIR_OK:
        msg = lib().vir_last_error()
        msg = msg.decode() if msg else ""
        if rc == VIR_EINVAL:
            raise ValueError(f"{what}: {msg}")
        raise VirError(f"{what}: {msg} (rc={rc})")


def make_grid(n, d, o, eps):
    g = VirGrid()
    for a in range(3):
        g.n[a] = int(n[a])
        g.d[a] = float(np.float32(d[a]))
        g.o[a] = float(np.float32(o[a]))
    g.eps = float(eps)
    return g


def ptr(t):
    """Device pointer of a torch tensor (or None)."""
    if t is None:
        return None
    return ctypes.c_void_p(t.data_ptr())


def stream_ptr(stream=None):
    import torch
    s = stream if stream is not None else torch.cuda.current_stream()
    return ctypes.c_void_p(s.cuda_stream)
