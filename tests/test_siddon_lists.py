"""Siddon list utilities (sys_mat/siddon.py:12-633).

CPU: index/format transforms on object arrays (round trips, involutions,
trimming, counting, flattening) and the float32 reduction loops of the
oracle pinned against float64 sums.  GPU: list2array / rays_ave /
rays_weight_ave reproduce the reference's float32 loops bit for bit.
"""
import numpy as np
import pytest

from oracle import siddon as osd
from oracle import siddon_lists_ref as lref

NP = (6, 5, 4)


def _sdlist(ravel=False, seed=0):
    from vir_amd.proj_geom import geom_circular
    views = np.arange(5) * np.pi / 5 + 0.1
    s, t = geom_circular((np.arange(7) - 3.0) * 0.8, views, "par", DetsZ=np.array([-0.7, 0.6]))
    return osd.siddons(s, t, NP, 1.0, 0.0, ravel=ravel)


def test_ravel_unravel_roundtrip_and_count():
    from vir_amd.sys_mat import siddon as sd
    L = _sdlist()
    R = sd.list_ravel(np.copy(L), NP)
    U = sd.list_unravel(np.copy(R), NP)
    cnt = sd.list_count_elem(L)
    assert cnt.dtype == np.int32 and cnt.shape == L.shape
    for idx, ray in np.ndenumerate(L):
        if ray is None:
            assert U[idx] is None and cnt[idx] == 0
            continue
        assert cnt[idx] == ray[0].size
        assert np.array_equal(R[idx][0], np.ravel_multi_index((ray[0], ray[1], ray[2]), NP))
        for k in range(4):
            assert np.array_equal(U[idx][k], ray[k])


def test_flatten_matches_csr():
    from vir_amd.sys_mat import siddon as sd
    L = _sdlist()
    ind, ln, off = sd.list_flatten(L, nPixels=NP)
    assert ind.dtype == np.int32 and ln.dtype == np.float32 and off.dtype == np.int32
    for k, ray in enumerate(L.flat):
        seg = slice(off[k], off[k + 1])
        if ray is None:
            assert off[k] == off[k + 1]
        else:
            assert np.array_equal(ind[seg], np.ravel_multi_index((ray[0], ray[1], ray[2]), NP))
            assert np.array_equal(ln[seg], ray[3])
    with pytest.raises(ValueError):
        sd.list_flatten(L)


def test_flip_shift_boundary_delete():
    from vir_amd.sys_mat import siddon as sd
    L = _sdlist()
    ray = next(r for r in L.flat if r is not None)
    f = sd.ray_flip(ray, x=3, z=2)
    assert np.array_equal(f[0], 6 - ray[0]) and np.array_equal(f[1], ray[1])
    assert np.array_equal(f[2], 4 - ray[2])
    back = sd.ray_flip(f, x=3, z=2)
    assert all(np.array_equal(a, b) for a, b in zip(back, ray))
    # an axis value of 0 does not flip (truthiness, as in the reference)
    assert np.array_equal(sd.ray_flip(ray, x=0)[0], ray[0])
    sh = sd.ray_index_shift(ray, [2, -1, 0])
    assert np.array_equal(sh[0], ray[0] + 2) and np.array_equal(sh[1], ray[1] - 1)
    b = sd.ray_boundary(sh, NP)
    keep = (sh[0] < NP[0]) & (sh[1] >= 0)
    assert np.array_equal(b[3], ray[3][keep])
    D = sd.list_del_ind(np.copy(L), [1, 2, 3], range(5), [0, 1])
    for idx, r in np.ndenumerate(D):
        if r is not None:
            assert np.isin(r[0], [1, 2, 3]).all() and np.isin(r[2], [0, 1]).all()
            orig = L[idx]
            k = np.isin(orig[0], [1, 2, 3]) & np.isin(orig[2], [0, 1])
            assert np.array_equal(r[3], orig[3][k])
    with pytest.raises(ValueError):
        sd.list_index_shift(np.copy(L), [1, 1, 1], ravel=True)


def test_reduction_loops_pinned():
    import scipy.sparse as sp
    rng = np.random.default_rng(0)
    ind = rng.integers(0, 120, 3000).astype(np.int32)
    ln = rng.uniform(0, 2, 3000).astype(np.float32)
    a = lref.list2array_loop(ind, ln, NP)
    col = np.asarray(sp.csr_matrix((ln.astype(np.float64), ind, [0, ind.size]),
                                   shape=(1, 120)).sum(0)).reshape(NP)
    np.testing.assert_allclose(a, col, rtol=1e-5)
    # merging k copies of one ray and dividing by k returns the ray (duplicates merged)
    u, v = lref.rays_ave_loop(np.tile(ind[:50], 3), np.tile(ln[:50], 3), 3)
    uu, vv = lref.rays_ave_loop(ind[:50], ln[:50], 1)
    assert np.array_equal(u, uu)
    np.testing.assert_allclose(v, vv, rtol=1e-6)


@pytest.mark.gpu
def test_gpu_reductions_bit_exact():
    from vir_amd.sys_mat import siddon as sd
    L = _sdlist()
    ind, ln, off = sd.list_flatten(L, nPixels=NP)
    assert np.array_equal(sd.list2array(L, NP), lref.list2array_loop(ind, ln, NP))
    assert np.array_equal(sd.list2array((ind, ln, off), NP, flat=True),
                          lref.list2array_loop(ind, ln, NP))
    got = sd.rays_ave((ind, ln, off), ravel=True, flat=True, nRays=L.size)
    u, v = lref.rays_ave_loop(ind, ln, L.size)
    assert np.array_equal(got[0], u) and np.array_equal(got[1], v)
    un = sd.rays_ave(L, nPixels=NP)
    assert np.array_equal(np.ravel_multi_index(tuple(un[:3]), NP), u)
    assert np.array_equal(un[3], v)
    # per-view average over the detector axis
    per_view = sd.list_ave(L, nPixels=NP, axis=(1, 2))
    assert per_view.shape == (L.shape[0],)
    w = np.linspace(0.5, 1.5, L.shape[1] * L.shape[2]).reshape(L.shape[1:])
    wa = sd.list_weight_ave(L, w, nPixels=NP, axis=(1, 2))
    r0 = L[0]
    wl = np.empty(r0.shape, dtype=object)
    for idx, r in np.ndenumerate(r0):
        wl[idx] = None if r is None else [np.ravel_multi_index(tuple(r[:3]), NP), w[idx] * r[3]]
    fi, fl, _ = sd.list_flatten(wl, ravel=True)
    u0, v0 = lref.rays_ave_loop(fi, fl, 1)
    assert np.array_equal(np.ravel_multi_index(tuple(wa[0][:3]), NP), u0)
    assert np.array_equal(wa[0][3], v0)
    conv = sd.list_convole(L, np.array([1.0, 2.0, 1.0]), nPixels=NP, axis=(1,))
    assert conv.shape == L.shape
