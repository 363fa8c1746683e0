"""CPU: pin the Siddon oracle (oracle/siddon_oracle.c) before trusting it.

1. bit-exact against golden vectors captured from the reference's own
   siddons() (tests/golden/make_siddon_golden.py);
2. against FP64 brute force (per-voxel Liang-Barsky clipping) on small grids;
3. invariants: sum of lengths = chord length of ray-box intersection;
4. the numpy reduction order the oracle assumes for np.linalg.norm.
"""
import hashlib

import numpy as np
import pytest

from conftest import grid_of
from oracle import siddon as osd


def sha(*arrays):
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def _check_unravel(g, key, eps=1e-5):
    nP, dP, oP = grid_of(g, key)
    shape, x, y, z, ln, off, fl, n = osd.trace(g[key + "_src"], g[key + "_trg"], nP, dP, oP, eps)
    assert np.array_equal(off, g[key + "_off"]), key
    assert np.array_equal(np.concatenate([x, y, z]), g[key + "_u"]), key
    assert np.array_equal(ln.view(np.int32), g[key + "_len"].view(np.int32)), key
    assert np.array_equal((fl >> 1) & 1, g[key + "_hit"]), key
    return fl


def test_catalog_bit_exact(golden):
    for k in range(int(golden["n_cat"])):
        key = f"cat{k:02d}"
        fl = _check_unravel(golden, key)
        assert int((fl & 1).any()) == int(golden[key + "_raised"]), key


def test_out_of_grid_case(golden):
    fl = _check_unravel(golden, "oob")
    assert int((fl & 1).any()) == int(golden["oob_raised"]) == 1
    with pytest.raises(ValueError):
        nP, dP, oP = grid_of(golden, "oob")
        osd.siddons(golden["oob_src"], golden["oob_trg"], nP, dP, oP, flat=True)


def test_random_edge_rays_bit_exact(golden):
    for ci in range(int(golden["n_rnd"])):
        key = f"rnd{ci}"
        _check_unravel(golden, key, float(golden[key + "_eps"]))


@pytest.mark.parametrize("name", ["par2d_small", "fan2d_small", "cone3d_small", "cone256"])
def test_structured_flat_bit_exact(golden, name):
    nP, dP, oP = grid_of(golden, name)
    ind, ln, off = osd.siddons(golden[name + "_src"], golden[name + "_trg"], nP, dP, oP, flat=True)
    assert np.array_equal(off, golden[name + "_foff"])
    assert np.array_equal(ind, golden[name + "_ind"])
    assert np.array_equal(ln.view(np.int32), golden[name + "_flen"].view(np.int32))


def test_par128_workload_hashes(golden):
    from vir_amd.proj_geom import geom_circular
    s, t = geom_circular(golden["par128_dets"], golden["par128_views"], "par")
    assert sha(s.astype(np.float32)) == str(golden["par128_src_sha"])
    assert sha(t.astype(np.float32)) == str(golden["par128_trg_sha"])
    ind, ln, off = osd.siddons(s, t, (128, 128, 1), 1.0, 0.0, flat=True)
    assert ind.size == int(golden["par128_nnz"])
    assert sha(ind) == str(golden["par128_ind_sha"])
    assert sha(ln) == str(golden["par128_len_sha"])
    assert sha(off) == str(golden["par128_off_sha"])


def test_norm_reduction_order():
    """np.linalg.norm(axis=-1) of float32 (...,3) == sqrtf((x2+y2)+z2)."""
    rng = np.random.default_rng(3)
    x = (rng.standard_normal((200000, 3)) * rng.choice([1e-3, 1, 1e3], (200000, 3))).astype(np.float32)
    s = x * x
    assert np.array_equal(np.linalg.norm(x, axis=-1), np.sqrt((s[:, 0] + s[:, 1]) + s[:, 2]))


def _brute(src, trg, n, d, o):
    """FP64 per-voxel intersection length by slab clipping."""
    src = np.asarray(src, np.float64)
    trg = np.asarray(trg, np.float64)
    D = trg - src
    L = np.linalg.norm(D)
    out = {}
    lo = [o[a] - n[a] * d[a] / 2 for a in range(3)]
    for ix in range(n[0]):
        for iy in range(n[1]):
            for iz in range(n[2]):
                t0, t1 = 0.0, 1.0
                ok = True
                for a, i in zip(range(3), (ix, iy, iz)):
                    b0 = lo[a] + i * d[a]
                    b1 = b0 + d[a]
                    if D[a] == 0:
                        if not (b0 <= src[a] <= b1):
                            ok = False
                        continue
                    u0, u1 = (b0 - src[a]) / D[a], (b1 - src[a]) / D[a]
                    t0 = max(t0, min(u0, u1))
                    t1 = min(t1, max(u0, u1))
                if ok and t1 > t0:
                    out[(ix, iy, iz)] = (t1 - t0) * L
    return out


def test_against_fp64_brute_force():
    rng = np.random.default_rng(11)
    n, d, o = (6, 5, 4), (1.0, 0.75, 1.5), (0.2, -0.1, 0.3)
    half = np.array(n) * np.array(d) / 2
    for _ in range(60):
        s = rng.uniform(-2, 2, 3) * half + o
        t = rng.uniform(-2, 2, 3) * half + o
        shape, x, y, z, ln, off, fl, nn = osd.trace(s, t, n, d, o)
        got = {}
        for xi, yi, zi, li in zip(x, y, z, ln):
            got[(int(xi), int(yi), int(zi))] = got.get((int(xi), int(yi), int(zi)), 0.0) + float(li)
        ref = _brute(np.float32(s), np.float32(t), n, d, o)
        big_ref = {k: v for k, v in ref.items() if v > 1e-3}
        big_got = {k: v for k, v in got.items() if v > 1e-3}
        assert set(big_ref) == set(big_got)
        for k in big_ref:
            assert abs(big_ref[k] - big_got[k]) <= 1e-4 * max(1.0, big_ref[k])
        assert abs(sum(ref.values()) - sum(got.values())) <= 1e-4 * max(1.0, sum(ref.values()))


@pytest.mark.parametrize("plant", [1, 2, 3, 4])
def test_golden_vectors_catch_planted_mistakes(golden, tmp_path, plant):
    """Compile a deliberately wrong oracle variant (no end trim / no index
    snapping / float32 list for axis-aligned rays / rounded midpoint index)
    and require the golden comparison to reject it."""
    import os
    import subprocess
    src = os.path.join(os.path.dirname(osd.__file__), "siddon_oracle.c")
    so = str(tmp_path / f"plant{plant}.so")
    subprocess.run(["gcc", "-O2", "-fPIC", "-shared", "-std=c11", "-ffp-contract=off",
                    f"-DORACLE_PLANT={plant}", "-o", so, src, "-lm"], check=True)
    lib = osd.load_library(so)
    mismatches = 0
    keys = [f"cat{k:02d}" for k in range(int(golden["n_cat"]))] + \
        [f"rnd{c}" for c in range(int(golden["n_rnd"]))]
    for key in keys:
        nP, dP, oP = grid_of(golden, key)
        eps = float(golden[key + "_eps"]) if key + "_eps" in golden else 1e-5
        shape, x, y, z, ln, off, fl, n = osd.trace(golden[key + "_src"], golden[key + "_trg"],
                                                   nP, dP, oP, eps, lib=lib)
        ok = (np.array_equal(off, golden[key + "_off"]) and
              np.array_equal(np.concatenate([x, y, z]), golden[key + "_u"]) and
              np.array_equal(ln.view(np.int32), golden[key + "_len"].view(np.int32)))
        mismatches += not ok
    for key in ["par2d_small", "fan2d_small", "cone3d_small", "cone256"]:
        nP, dP, oP = grid_of(golden, key)
        shape, x, y, z, ln, off, fl, n = osd.trace(golden[key + "_src"], golden[key + "_trg"],
                                                   nP, dP, oP, lib=lib)
        ok = (np.array_equal(off, golden[key + "_foff"]) and
              np.array_equal(ln.view(np.int32), golden[key + "_flen"].view(np.int32)))
        mismatches += not ok
    assert mismatches > 0


def test_threaded_cpu_path_matches_trace_products():
    """The multi-threaded CPU baseline (trace + float64 FP + float32 BP) equals
    the sparse products of the single-threaded trace."""
    from oracle import projection as op
    from vir_amd.proj_geom import geom_circular
    views = np.arange(6) * np.pi / 3 + 0.2
    s, t = geom_circular((np.arange(20) - 9.5) * 1.1, views, "cone", src_iso=40.0, det_iso=30.0,
                         DetsZ=(np.arange(10) - 4.5) * 1.1)
    nP = (12, 11, 9)
    rng = np.random.default_rng(0)
    vol = rng.uniform(0, 1, nP).astype(np.float32)
    ind, ln, off = osd.siddons(s, t, nP, 0.9, 0.0, flat=True)
    ref_fp = op.fp64(vol, ind, ln, off).astype(np.float32)
    ref_bp = op.bp64(ref_fp, ind, ln, off, int(np.prod(nP)))
    for nth in (1, 4):
        sino, back, nnz = osd.fp_bp_threads(s, t, vol, nP, 0.9, nthreads=nth)
        assert nnz == ind.size
        assert np.array_equal(sino, ref_fp)
        np.testing.assert_allclose(back.reshape(-1), ref_bp.reshape(-1), rtol=1e-5, atol=1e-5)


def test_csr_operator_equals_scipy_products():
    """oracle.projection.CsrOperator (float64 C products used for SIRT
    references at workload size) against scipy.sparse and fp64/bp64."""
    import scipy.sparse as sp
    from oracle import projection as op
    from oracle import sirt_ref
    rng = np.random.default_rng(21)
    nrows, ncols = 300, 90
    counts = rng.integers(0, 12, nrows)
    off = np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)
    ind = rng.integers(0, ncols, off[-1]).astype(np.int32)
    ln = rng.uniform(0, 2, off[-1]).astype(np.float32)
    A = op.CsrOperator(ind, ln, off, ncols, nthreads=3)
    S = sp.csr_matrix((ln.astype(np.float64), ind, off), shape=(nrows, ncols))
    x = rng.uniform(-1, 1, ncols)
    y = rng.uniform(-1, 1, nrows)
    assert A.shape == (nrows, ncols) and A.T.shape == (ncols, nrows)
    assert np.allclose(A @ x, S @ x, rtol=0, atol=1e-12)
    assert np.allclose(A.T @ y, S.T @ y, rtol=0, atol=1e-12)
    assert np.allclose(A @ x, op.fp64(x, ind, ln, off), rtol=0, atol=1e-12)
    assert np.allclose(A.T @ y, op.bp64(y, ind, ln, off, ncols), rtol=0, atol=1e-12)
    b = S @ rng.uniform(0, 1, ncols)
    assert np.allclose(sirt_ref.sirt(A, b, 7), sirt_ref.sirt(S, b, 7), rtol=1e-12, atol=1e-14)
