"""CPU: the C-ABI boundary and host-side logic (no GPU compute).

* libvir_amd.so loads and exports exactly the entry points declared in
  include/vir_amd.h, and the ctypes binding declares each of them;
* vir_amd.proj_geom.geom_circular reproduces the reference's endpoints bit
  for bit (golden vectors from proj_geom.py);
* per-view projection matrices map every ray onto its own index;
* argument validation in the C ABI rejects bad input without launching.
"""
import ctypes
import os
import re

import numpy as np
import pytest

from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "vir_amd.h")


def declared_functions():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(vir_\w+)\s*\(", txt)))


def test_library_exports_header_symbols():
    from vir_amd import _lib
    so = ctypes.CDLL(_lib.LIB_PATH)
    names = declared_functions()
    assert len(names) >= 10
    for n in names:
        assert hasattr(so, n), n
    assert set(names) == set(_lib.SIGNATURES), set(names) ^ set(_lib.SIGNATURES)
    lib = _lib.lib()
    assert lib.vir_version() >= 1


def test_invalid_arguments_rejected():
    from vir_amd import _lib
    lib = _lib.lib()
    g = _lib.make_grid((4, 4, 0), (1, 1, 1), (0, 0, 0), 1e-5)  # nZ = 0 -> invalid
    rc = lib.vir_siddon_count(None, None, 10, g, None, None, None)
    assert rc == -1
    assert b"grid n" in lib.vir_last_error()
    g = _lib.make_grid((4, 4, 1), (1, -1, 1), (0, 0, 0), 1e-5)
    assert lib.vir_siddon_fp(None, None, 1, g, None, None, 0, None, None) == -1
    g = _lib.make_grid((4, 4, 1), (1, 1, 1), (0, 0, 0), 1e-5)
    assert lib.vir_siddon_fp(None, None, 3, g, None, None, 0, None, None) == -1
    assert lib.vir_siddon_bp(None, None, 0, 1, 1, None, g, None, None, 0, None, None) == -1
    assert lib.vir_siddon_count_guarded(None, None, 1, g, None, None) == -1
    assert lib.vir_siddon_count_guarded(None, None, 0, g, None, None) == -1  # count is NULL
    assert lib.vir_csr_transpose(None, None, None, 1, 1 << 31, 8, None, None, None, None) == -1
    assert lib.vir_siddon_bp_workspace(g, 0) == -1
    # per-ray records: NULL endpoints / records, misaligned records, big detector
    assert lib.vir_siddon_ray_records(None, None, 1, g, None, None) == -1
    assert lib.vir_siddon_ray_records(None, None, 0, g, None, None) == 0  # nothing to do
    assert lib.vir_siddon_ray_records(ctypes.c_void_p(64), ctypes.c_void_p(64), 1, g, None, None) == -1
    assert lib.vir_siddon_ray_records(ctypes.c_void_p(64), ctypes.c_void_p(64), 1, g, ctypes.c_void_p(72),
                                      None) == -1
    assert b"aligned" in lib.vir_last_error()
    assert lib.vir_siddon_bp_rec(None, 1, 1, 1, None, g, None, None, 0, None, None) == -1
    assert lib.vir_siddon_fp_rec(None, 1, g, None, None, 0, None, None) == -1
    assert lib.vir_siddon_fp_rec(ctypes.c_void_p(72), 1, g, ctypes.c_void_p(64), ctypes.c_void_p(64), 0, None,
                                 None) == -1
    assert lib.vir_siddon_fp_rec(None, 0, g, None, None, 0, None, None) == 0
    assert lib.vir_siddon_bp_rec(ctypes.c_void_p(72), 1, 1, 1, ctypes.c_void_p(64), g, ctypes.c_void_p(64),
                                 ctypes.c_void_p(64), 0, None, None) == -1
    assert lib.vir_siddon_bp_rec(ctypes.c_void_p(64), 1, 1 << 15, 1 << 15, ctypes.c_void_p(64), g,
                                 ctypes.c_void_p(64), ctypes.c_void_p(64), 0, None, None) == -1
    assert b"nu*nv" in lib.vir_last_error()
    # batches: at least one item, records required, fused forms refuse VIR_ACCUMULATE
    P64 = ctypes.c_void_p(64)
    assert lib.vir_siddon_fp_rec_batch(P64, 1, g, 0, P64, P64, 0, None, None) == -1
    assert b"nbatch" in lib.vir_last_error()
    assert lib.vir_siddon_fp_rec_batch(None, 1, g, 2, P64, P64, 0, None, None) == -1
    assert lib.vir_siddon_fp_rec_batch(None, 0, g, 2, None, None, 0, None, None) == 0  # nothing to do
    assert lib.vir_siddon_bp_batch_workspace(g, 1, 0) == -1
    assert lib.vir_siddon_bp_batch_workspace(g, 1, 3) == 2 * lib.vir_siddon_bp_workspace(g, 1)
    assert lib.vir_siddon_bp_rec_batch(P64, 1, 1, 1, P64, g, 0, P64, P64, 0, None, None) == -1
    assert lib.vir_siddon_bp_rec_batch(None, 1, 1, 1, P64, g, 2, P64, P64, 0, None, None) == -1
    assert lib.vir_siddon_fp_rec_res_batch(P64, 1, g, 2, P64, P64, P64, P64, _lib.VIR_ACCUMULATE, None) == -1
    assert lib.vir_siddon_bp_rec_upd_batch(P64, 1, 1, 1, P64, g, 2, P64, P64, 1.0, P64, _lib.VIR_ACCUMULATE,
                                           None, None) == -1


@pytest.mark.parametrize("name,kw", [
    ("par2d", dict(geom="par")),
    ("fan2d", dict(geom="fan", src_iso=40.0, det_iso=25.0)),
    ("par3d", dict(geom="par", DetsZ="detz")),
    ("cone3d", dict(geom="cone", src_iso=40.0, det_iso=25.0, DetsZ="detz")),
])
def test_geom_circular_bit_exact(golden, name, kw):
    from vir_amd.proj_geom import geom_circular
    kw = dict(kw)
    if kw.get("DetsZ") == "detz":
        kw["DetsZ"] = golden["geom_detz"]
    s, t = geom_circular(golden["geom_dets"], golden["geom_views"], **kw)
    assert np.array_equal(s, golden["geom_" + name + "_src"])
    assert np.array_equal(t, golden["geom_" + name + "_trg"])


def test_geom_errors():
    from vir_amd.proj_geom import geom_circular
    with pytest.raises(ValueError):
        geom_circular(np.arange(3.0), np.arange(2.0), geom="fan")
    with pytest.raises(ValueError):
        geom_circular(np.arange(3.0), np.arange(2.0), geom="cone", src_iso=1, det_iso=1)
    with pytest.raises(ValueError):
        geom_circular(np.arange(3.0), np.arange(2.0), geom="helix")


@pytest.mark.parametrize("geom", ["par", "fan", "cone", "par3"])
def test_view_projection_matrices(geom):
    from vir_amd.proj_geom import geom_circular, view_projections
    views = np.arange(7) * 2 * np.pi / 7 + 0.1
    du = (np.arange(11) - 5.0) * 1.3
    dv = (np.arange(6) - 2.5) * 0.9
    if geom == "par":
        s, t = geom_circular(du, views, "par")
    elif geom == "fan":
        s, t = geom_circular(du, views, "fan", src_iso=50.0, det_iso=30.0)
    elif geom == "cone":
        s, t = geom_circular(du, views, "cone", src_iso=50.0, det_iso=30.0, DetsZ=dv)
    else:
        s, t = geom_circular(du, views, "par", DetsZ=dv)
    s = s.astype(np.float32)
    t = t.astype(np.float32)
    nv = t.shape[2] if t.ndim == 4 else 1
    s4 = s.reshape(len(views), len(du), nv, 3)
    t4 = t.reshape(len(views), len(du), nv, 3)
    P = view_projections(s4, t4).astype(np.float64).reshape(-1, 3, 4)
    rng = np.random.default_rng(0)
    # a point on ray (a, iu, iv) projects to (iu, iv)
    for _ in range(50):
        a, iu, iv = rng.integers(len(views)), rng.integers(len(du)), rng.integers(nv)
        lam = rng.uniform(0.2, 0.8)
        X = s4[a, iu, iv] + lam * (t4[a, iu, iv] - s4[a, iu, iv])
        h = P[a] @ np.append(X, 1.0)
        assert h[2] > 0
        assert abs(h[0] / h[2] - iu) < 1e-3
        if nv > 1:
            assert abs(h[1] / h[2] - iv) < 1e-3


def test_view_projection_rejects_unstructured():
    from vir_amd.proj_geom import view_projections
    rng = np.random.default_rng(1)
    s = rng.normal(size=(2, 5, 1, 3)).astype(np.float32)
    t = rng.normal(size=(2, 5, 1, 3)).astype(np.float32)
    with pytest.raises(ValueError):
        view_projections(s, t)


def test_oracle_projection_matches_reference_c(golden):
    """The FP64 stored-matrix oracle agrees with the reference's compiled
    projection_c.c (oracle/_ref), and the pair is adjoint."""
    from oracle import projection as op
    from oracle import siddon as osd
    if not os.path.exists(op.REF_SO):
        pytest.skip("oracle/_ref not built (reference sources absent)")
    ref = op.RefProjection()
    key = "cone3d_small"
    ind, ln, off = golden[key + "_ind"], golden[key + "_flen"], golden[key + "_foff"]
    nvox = 16 * 16 * 12
    rng = np.random.default_rng(5)
    x = rng.uniform(0, 1, nvox).astype(np.float32)
    y = rng.uniform(0, 1, len(off) - 1).astype(np.float32)
    f64 = op.fp64(x, ind, ln, off)
    fc = ref.fp(x, ind, ln, off)
    assert np.max(np.abs(f64 - fc)) <= 1e-5 * max(1.0, np.abs(f64).max())
    b64 = op.bp64(y, ind, ln, off, nvox)
    bc = ref.bp(y, ind, ln, off, nvox)
    assert np.max(np.abs(b64 - bc)) <= 1e-5 * max(1.0, np.abs(b64).max())
    lhs = float(np.dot(f64, y.astype(np.float64)))
    rhs = float(np.dot(x.astype(np.float64), b64))
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)


def test_noalloc_backprojector_refuses_oversized_staging():
    """vir_siddon_bp_rec_noalloc checks its fixed staging before anything
    reaches the device: VIR_ACCUMULATE on a single-chunk 256^3 plan needs a
    64 MiB volume, more than its 32 MiB (the pointers are never read)."""
    from vir_amd import _lib
    lib = _lib.lib()
    g = _lib.make_grid((256, 256, 256), (1, 1, 1), (0, 0, 0), 1e-5)
    assert lib.vir_siddon_bp_workspace(g, 4) == 0
    p = ctypes.c_void_p(256)
    rc = lib.vir_siddon_bp_rec_noalloc(p, 4, 8, 8, p, g, p, p, _lib.VIR_ACCUMULATE, None)
    assert rc == -1 and b"staging" in lib.vir_last_error()
