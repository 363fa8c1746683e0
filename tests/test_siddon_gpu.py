"""GPU parity of the Siddon path (libvir_amd through its C ABI).

* the GPU trace equals the reference's siddons() bit for bit (golden
  vectors) and the oracle on fresh inputs;
* the on-the-fly forward projector equals the FP64 oracle within 1e-6
  relative (compensated f32 ray sums);
* the brick-gather backprojector equals the oracle's transpose A^T within
  f32 summation error, single-ray backprojections are exact, repeated runs
  are bitwise identical, and <Ax, y> = <x, A^T y> to 1e-5 relative;
* the stored-matrix operators in sequential mode equal the reference's own
  compiled projection_c.c bit for bit.
"""
import hashlib
import os

import numpy as np
import pytest

from conftest import grid_of

pytestmark = pytest.mark.gpu


def sha(*arrays):
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def _sd():
    from vir_amd.sys_mat import siddon
    return siddon


def test_trace_catalog_and_random_bit_exact(golden):
    sd = _sd()
    keys = [f"cat{k:02d}" for k in range(int(golden["n_cat"]))] + \
        [f"rnd{c}" for c in range(int(golden["n_rnd"]))]
    for key in keys:
        print("case", key, flush=True)
        nP, dP, oP = grid_of(golden, key)
        eps = float(golden[key + "_eps"]) if key + "_eps" in golden else 1e-5
        rays = sd.siddons(golden[key + "_src"], golden[key + "_trg"], nP, dP, oP, epsilon=eps)
        rays = rays.reshape(-1)
        off = golden[key + "_off"]
        u = golden[key + "_u"]
        nnz = off[-1]
        gx, gy, gz = u[:nnz], u[nnz:2 * nnz], u[2 * nnz:]
        gl = golden[key + "_len"]
        hit = golden[key + "_hit"]
        for r, ray in enumerate(rays):
            if not hit[r]:
                assert ray is None, (key, r)
                continue
            s, e = off[r], off[r + 1]
            assert np.array_equal(ray[0], gx[s:e]), (key, r)
            assert np.array_equal(ray[1], gy[s:e]), (key, r)
            assert np.array_equal(ray[2], gz[s:e]), (key, r)
            assert np.array_equal(ray[3].view(np.int32), gl[s:e].view(np.int32)), (key, r)
        if key.startswith("cat") and int(golden[key + "_raised"]):
            with pytest.raises(ValueError):
                sd.siddons(golden[key + "_src"], golden[key + "_trg"], nP, dP, oP, flat=True)


@pytest.mark.parametrize("name", ["par2d_small", "fan2d_small", "cone3d_small", "cone256"])
def test_trace_flat_bit_exact(golden, name):
    sd = _sd()
    nP, dP, oP = grid_of(golden, name)
    ind, ln, off = sd.siddons(golden[name + "_src"], golden[name + "_trg"], nP, dP, oP, flat=True)
    assert off.dtype == np.int32 and ind.dtype == np.int32 and ln.dtype == np.float32
    assert np.array_equal(off, golden[name + "_foff"])
    assert np.array_equal(ind, golden[name + "_ind"])
    assert np.array_equal(ln.view(np.int32), golden[name + "_flen"].view(np.int32))


def test_trace_par128_workload_hashes(golden):
    sd = _sd()
    from vir_amd.proj_geom import geom_circular
    s, t = geom_circular(golden["par128_dets"], golden["par128_views"], "par")
    ind, ln, off = sd.siddons(s, t, (128, 128, 1), 1.0, 0.0, flat=True)
    assert ind.size == int(golden["par128_nnz"])
    assert sha(ind) == str(golden["par128_ind_sha"])
    assert sha(ln) == str(golden["par128_len_sha"])
    assert sha(off) == str(golden["par128_off_sha"])


def test_trace_ravel_format(golden):
    sd = _sd()
    from oracle import siddon as osd
    key = "rnd2"
    nP, dP, oP = grid_of(golden, key)
    src, trg = golden[key + "_src"], golden[key + "_trg"]
    a = sd.siddons(src, trg, nP, dP, oP, ravel=True)
    b = osd.siddons(src, trg, nP, dP, oP, ravel=True)
    for ra, rb in zip(a.reshape(-1), b.reshape(-1)):
        if rb is None:
            assert ra is None
            continue
        assert np.array_equal(ra[0], rb[0]) and np.array_equal(ra[1], rb[1])


def _geometry(kind, rng):
    from vir_amd.proj_geom import geom_circular
    if kind == "par2d":
        v = np.arange(23) * np.pi / 23
        d = (np.arange(41) - 20.0) * 0.9
        s, t = geom_circular(d, v, "par")
        return s, t, (32, 32, 1), 1.0
    if kind == "fan2d":
        v = np.arange(19) * 2 * np.pi / 19
        d = (np.arange(52) - 25.5) * 1.0
        s, t = geom_circular(d, v, "fan", src_iso=60.0, det_iso=45.0)
        return s, t, (40, 40, 1), 1.0
    if kind == "cone3d":
        v = np.arange(13) * 2 * np.pi / 13
        du = (np.arange(36) - 17.5) * 1.0
        dv = (np.arange(24) - 11.5) * 1.0
        s, t = geom_circular(du, v, "cone", src_iso=80.0, det_iso=80.0, DetsZ=dv)
        return s, t, (40, 36, 20), 0.5
    if kind == "par3d":
        v = np.arange(11) * np.pi / 11
        du = (np.arange(30) - 14.5) * 1.0
        dv = (np.arange(10) - 4.5) * 1.0
        s, t = geom_circular(du, v, "par", DetsZ=dv)
        return s, t, (24, 24, 10), 1.0
    raise ValueError(kind)


@pytest.mark.parametrize("kind", ["par2d", "fan2d", "cone3d", "par3d"])
def test_fp_bp_against_oracle(kind):
    import torch
    sd = _sd()
    from oracle import projection as op
    from oracle import siddon as osd
    rng = np.random.default_rng(7)
    s, t, nP, dP = _geometry(kind, rng)
    print("case", kind, flush=True)
    P = sd.SiddonProjector(s, t, nP, dP)
    assert P.proj is not None, "structured layout expected"
    ind, ln, off = osd.siddons(s, t, nP, dP, 0.0, flat=True)
    nvox = int(np.prod(nP))
    x = rng.uniform(0, 1, nvox).astype(np.float32)
    y = rng.uniform(0, 1, P.nrays).astype(np.float32)
    # forward
    fa = P.fp(torch.from_numpy(x).cuda().reshape(nP)).cpu().numpy().reshape(-1)
    f64 = op.fp64(x, ind, ln, off)
    assert np.max(np.abs(fa - f64)) <= 1e-6 * max(1.0, np.abs(f64).max())
    # adjoint
    ba = P.bp(torch.from_numpy(y).cuda().reshape(P.ray_shape)).cpu().numpy().reshape(-1)
    b64 = op.bp64(y, ind, ln, off, nvox)
    err = np.abs(ba - b64)
    assert np.max(err) <= 2e-6 * max(1.0, np.abs(b64).max()), np.max(err)
    # determinism
    bb = P.bp(torch.from_numpy(y).cuda().reshape(P.ray_shape)).cpu().numpy().reshape(-1)
    assert np.array_equal(ba.view(np.int32), bb.view(np.int32))
    # dot-product adjoint test
    lhs = float(np.dot(fa.astype(np.float64), y))
    rhs = float(np.dot(x.astype(np.float64), ba))
    assert abs(lhs - rhs) <= 1e-5 * abs(lhs)


@pytest.mark.parametrize("kind", ["par2d", "fan2d", "cone3d", "par3d"])
def test_unguarded_kernels_same_bits(kind):
    """vir_siddon_count_guarded finds no ray needing the range-checked
    division on the cone-beam set (source outside every plane, no zero
    direction component), every ray of the 2-D and parallel 3-D sets needs
    the float64 list (zero z component); where the count is 0, FP and BP
    with VIR_UNGUARDED equal the dynamically dispatched kernels bit for bit."""
    import torch
    from vir_amd import _lib
    from vir_amd.device import stream
    sd = _sd()
    rng = np.random.default_rng(11)
    s, t, nP, dP = _geometry(kind, rng)
    print("case", kind, flush=True)
    P = sd.SiddonProjector(s, t, nP, dP)
    if kind == "cone3d":
        assert P.guarded_rays == 0 and P._ung == _lib.VIR_UNGUARDED
    else:
        assert P.guarded_rays > 0 and P._ung == 0
        return
    L = _lib.lib()
    x = torch.from_numpy(rng.uniform(0, 1, P.nvox).astype(np.float32)).cuda()
    y = torch.from_numpy(rng.uniform(0, 1, P.nrays).astype(np.float32)).cuda()
    outs = []
    for fl in (0, _lib.VIR_UNGUARDED):
        f = torch.empty(P.nrays, dtype=torch.float32, device="cuda")
        b = torch.empty(P.nvox, dtype=torch.float32, device="cuda")
        _lib.check(L.vir_siddon_fp(_lib.ptr(P.src), _lib.ptr(P.trg), P.nrays, P.grid, _lib.ptr(x),
                                   _lib.ptr(f), fl, None, stream()), "fp")
        nviews, nu, nv = P.layout
        _lib.check(L.vir_siddon_bp(_lib.ptr(P.src), _lib.ptr(P.trg), nviews, nu, nv, _lib.ptr(P.proj),
                                   P.grid, _lib.ptr(y), _lib.ptr(b), fl, _lib.ptr(P._ws), stream()), "bp")
        outs.append((f.cpu().numpy(), b.cpu().numpy()))
    assert np.array_equal(outs[0][0].view(np.int32), outs[1][0].view(np.int32))
    assert np.array_equal(outs[0][1].view(np.int32), outs[1][1].view(np.int32))


@pytest.mark.parametrize("kind", ["par2d", "fan2d", "cone3d", "par3d", "cone256"])
def test_bp_from_records_same_bits(kind):
    """vir_siddon_bp_rec / vir_siddon_fp_rec, reading the per-ray records that
    vir_siddon_ray_records builds once, equal vir_siddon_bp / vir_siddon_fp,
    which recompute each (candidate) ray's setup, bit for bit: overwrite and
    VIR_ACCUMULATE, with and without
    VIR_UNGUARDED where its precondition holds.  cone256: 8 views of the
    256^3 cone-beam workload."""
    import torch
    from vir_amd import _lib
    from vir_amd.device import stream
    sd = _sd()
    rng = np.random.default_rng(12)
    if kind == "cone256":
        from vir_amd.workloads import cone256_geometry
        s, t, grid = cone256_geometry(np.arange(0, 360, 45))
        nP, dP = grid["nPixels"], grid["dPixels"]
    else:
        s, t, nP, dP = _geometry(kind, rng)
    print("case", kind, flush=True)
    P = sd.SiddonProjector(s, t, nP, dP)
    assert P.rec is not None and P.rec.numel() == P.nrays * 20
    assert sd.SiddonProjector(s, t, nP, dP, records=False).rec is None
    L = _lib.lib()
    y = torch.from_numpy(rng.uniform(-1, 1, P.nrays).astype(np.float32)).cuda()
    z = torch.from_numpy(rng.uniform(0, 1, P.nvox).astype(np.float32)).cuda()
    nviews, nu, nv = P.layout
    for fl in sorted({0, P._ung}):
        for acc in (0, _lib.VIR_ACCUMULATE):
            a, b = z.clone(), z.clone()
            _lib.check(L.vir_siddon_bp(_lib.ptr(P.src), _lib.ptr(P.trg), nviews, nu, nv, _lib.ptr(P.proj),
                                       P.grid, _lib.ptr(y), _lib.ptr(a), fl | acc, _lib.ptr(P._ws), stream()),
                       "bp")
            _lib.check(L.vir_siddon_bp_rec(_lib.ptr(P.rec), nviews, nu, nv, _lib.ptr(P.proj), P.grid,
                                           _lib.ptr(y), _lib.ptr(b), fl | acc, _lib.ptr(P._ws), stream()),
                       "bp_rec")
            a, b = a.cpu().numpy(), b.cpu().numpy()
            assert np.abs(a).max() > 0
            assert np.array_equal(a.view(np.int32), b.view(np.int32)), (kind, fl, acc)
    # the forward projector from records: same bits and out-of-grid count
    x = torch.from_numpy(rng.uniform(-1, 1, P.nvox).astype(np.float32)).cuda()
    for fl in sorted({0, P._ung}):
        for acc in (0, _lib.VIR_ACCUMULATE):
            a, b = y.clone(), y.clone()
            ea, eb = (torch.zeros(1, dtype=torch.int32, device="cuda") for _ in range(2))
            _lib.check(L.vir_siddon_fp(_lib.ptr(P.src), _lib.ptr(P.trg), P.nrays, P.grid, _lib.ptr(x), _lib.ptr(a),
                                       fl | acc, _lib.ptr(ea), stream()), "fp")
            _lib.check(L.vir_siddon_fp_rec(_lib.ptr(P.rec), P.nrays, P.grid, _lib.ptr(x), _lib.ptr(b), fl | acc,
                                           _lib.ptr(eb), stream()), "fp_rec")
            a, b = a.cpu().numpy(), b.cpu().numpy()
            assert np.array_equal(a.view(np.int32), b.view(np.int32)), (kind, fl, acc)
            assert int(ea.item()) == int(eb.item()) == 0
    # the projector's own bp() takes the record path
    c = P.bp(y).cpu().numpy().reshape(-1)
    a = torch.empty(P.nvox, dtype=torch.float32, device="cuda")
    _lib.check(L.vir_siddon_bp(_lib.ptr(P.src), _lib.ptr(P.trg), nviews, nu, nv, _lib.ptr(P.proj), P.grid,
                               _lib.ptr(y), _lib.ptr(a), P._ung, _lib.ptr(P._ws), stream()), "bp")
    assert np.array_equal(c.view(np.int32), a.cpu().numpy().view(np.int32))


@pytest.mark.parametrize("kind", ["par2d", "cone3d"])
def test_bp_single_rays_exact(kind):
    """Backprojecting a one-hot sinogram reproduces that ray's own entries
    exactly (each voxel's lengths summed in entry order)."""
    import torch
    sd = _sd()
    from oracle import siddon as osd
    rng = np.random.default_rng(3)
    s, t, nP, dP = _geometry(kind, rng)
    P = sd.SiddonProjector(s, t, nP, dP)
    shape, x, y, z, ln, off, fl, n = osd.trace(s, t, nP, dP, 0.0)
    nvox = int(np.prod(nP))
    picks = rng.choice(P.nrays, 12, replace=False)
    for r in picks:
        e0, e1 = off[r], off[r + 1]
        ref = np.zeros(nvox, np.float32)
        for xi, yi, zi, li in zip(x[e0:e1], y[e0:e1], z[e0:e1], ln[e0:e1]):
            v = (int(xi) * nP[1] + int(yi)) * nP[2] + int(zi)
            ref[v] = np.float32(ref[v] + li)
        onehot = torch.zeros(P.nrays, dtype=torch.float32, device="cuda")
        onehot[int(r)] = 1.0
        got = P.bp(onehot.reshape(P.ray_shape)).cpu().numpy().reshape(-1)
        assert np.array_equal(got.view(np.int32), ref.view(np.int32)), (kind, r)


def test_unstructured_rays_use_stored_transpose():
    import torch
    sd = _sd()
    from oracle import projection as op
    from oracle import siddon as osd
    rng = np.random.default_rng(9)
    nP, dP = (12, 10, 6), (1.0, 0.8, 1.2)
    half = np.array(nP) * np.array(dP) / 2
    s = rng.uniform(-2, 2, (500, 3)) * half
    t = rng.uniform(-2, 2, (500, 3)) * half
    P = sd.SiddonProjector(s, t, nP, dP)
    assert P.proj is None
    ind, ln, off = osd.siddons(s, t, nP, dP, 0.0, flat=True)
    y = rng.uniform(0, 1, 500).astype(np.float32)
    b = P.bp(torch.from_numpy(y).cuda()).cpu().numpy().reshape(-1)
    b64 = op.bp64(y, ind, ln, off, int(np.prod(nP)))
    assert np.max(np.abs(b - b64)) <= 2e-6 * max(1.0, np.abs(b64).max())


def test_stored_matrix_bit_exact_vs_reference_c(golden):
    """sd_f_proj / sd_b_proj (sequential) == the reference's projection_c.c."""
    from oracle import projection as op
    from vir_amd.projection import proj
    if not os.path.exists(op.REF_SO):
        pytest.skip("oracle/_ref not built")
    ref = op.RefProjection()
    key = "cone3d_small"
    ind, ln, off = golden[key + "_ind"], golden[key + "_flen"], golden[key + "_foff"]
    nP = (16, 16, 12)
    rng = np.random.default_rng(2)
    x = rng.uniform(0, 1, nP).astype(np.float32)
    y = rng.uniform(0, 1, len(off) - 1).astype(np.float32)
    f = proj.sd_f_proj(x, (ind, ln, off), sequential=True)
    assert np.array_equal(f.view(np.int32), ref.fp(x, ind, ln, off).view(np.int32))
    b = proj.sd_b_proj(y, (ind, ln, off), nP, sequential=True).reshape(-1)
    assert np.array_equal(b.view(np.int32), ref.bp(y, ind, ln, off, int(np.prod(nP))).view(np.int32))
    # default (compensated) mode against FP64
    f2 = proj.sd_f_proj(x, (ind, ln, off))
    f64 = op.fp64(x, ind, ln, off)
    assert np.max(np.abs(f2 - f64)) <= 1e-6 * f64.max()
    # object-array sdlists (ravel and unravel) give the same results
    from vir_amd.sys_mat import siddon as sd
    src, trg = golden[key + "_src"], golden[key + "_trg"]
    rav = sd.siddons(src, trg, nP, 0.5, 0.0, ravel=True)
    unr = sd.siddons(src, trg, nP, 0.5, 0.0)
    fr = proj.sd_f_proj(x, rav, ravel=True, flat=False)
    fu = proj.sd_f_proj(x, unr, ravel=False, flat=False)
    assert np.array_equal(fr.reshape(-1), f2) and np.array_equal(fu.reshape(-1), f2)
    br = proj.sd_b_proj(y.reshape(rav.shape), rav, nP, ravel=True)
    bu = proj.sd_b_proj(y.reshape(unr.shape), unr, nP, ravel=False)
    assert np.array_equal(br, bu)


def test_out_of_grid_index_raises_like_reference(golden):
    """A ray whose midpoint index leaves the grid: the reference raises
    ValueError in flat mode (golden 'oob'); ravel=False output keeps the raw
    indices; the on-the-fly projector reports it when asked."""
    import torch
    sd = _sd()
    nP, dP, oP = grid_of(golden, "oob")
    src, trg = golden["oob_src"], golden["oob_trg"]
    assert int(golden["oob_raised"]) == 1
    with pytest.raises(ValueError):
        sd.siddons(src, trg, nP, dP, oP, flat=True)
    with pytest.raises(ValueError):
        sd.siddons(src, trg, nP, dP, oP, ravel=True)
    ray = sd.siddons(src, trg, nP, dP, oP).reshape(-1)[0]
    assert np.array_equal(np.concatenate(ray[:3]), golden["oob_u"])
    assert np.array_equal(ray[3].view(np.int32), golden["oob_len"].view(np.int32))
    P = sd.SiddonProjector(src, trg, nP, dP, oP, structured=False)
    with pytest.raises(ValueError):
        P.fp(torch.ones(nP, device="cuda"), check=True)
