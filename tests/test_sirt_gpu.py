"""GPU SIRT against the float64 recurrence (oracle/sirt_ref.py).

* stored-matrix SIRT (the sirt_test.py drop-in) on a 2-D parallel Siddon
  system: every iterate within 1e-5 relative L2 of FP64;
* on-the-fly Siddon (cone) and distance-driven (fan) SIRT: final iterate
  within 1e-5 relative of FP64 with the same system matrix;
* the HIP-graph-captured iteration equals the eager one bit for bit over
  several replays with eager work in between.
"""
import numpy as np
import pytest

from oracle import dd_ref, sirt_ref
from oracle import siddon as osd

pytestmark = pytest.mark.gpu
TOL = 1e-5


def rel(a, b):
    a = np.asarray(a, np.float64).reshape(-1)
    b = np.asarray(b, np.float64).reshape(-1)
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def _par_problem():
    from vir_amd.proj_geom import geom_circular
    v = np.arange(40) * np.pi / 40
    d = (np.arange(49) - 24.0) * 1.0
    s, t = geom_circular(d, v, "par")
    nP = (32, 32, 1)
    rng = np.random.default_rng(0)
    x_true = np.zeros(nP, np.float32)
    x_true[8:24, 10:22, 0] = rng.uniform(0.5, 1.0, (16, 12))
    return s, t, nP, x_true


def test_sirt_dropin_stored_matrix():
    from vir_amd import sirt as vsirt
    from vir_amd.sys_mat.siddon import siddons
    s, t, nP, x_true = _par_problem()
    sd = siddons(s, t, nP, 1.0, ravel=True)
    ind, ln, off = osd.siddons(s, t, nP, 1.0, 0.0, flat=True)
    A = sirt_ref.csr_matrix(ind, ln, off, int(np.prod(nP)))
    b = (A @ x_true.reshape(-1).astype(np.float64)).astype(np.float32).reshape(s.shape[:-1])
    W_sino = (A @ np.ones(A.shape[1])).astype(np.float32).reshape(b.shape)
    W_img = (A.T @ np.ones(A.shape[0])).astype(np.float32).reshape(nP)
    its = vsirt.sirt(b, sd, W_sino, W_img, nP, gamma=1.0, iters=12)
    ref = sirt_ref.sirt(A, b, 12, W_sino=W_sino, W_img=W_img, return_all=True)
    for i in range(12):
        assert rel(its[..., i], ref[i]) <= TOL, i
    # the iterates approach the consistent solution
    assert rel(its[..., -1], x_true) < rel(its[..., 0], x_true)


def test_sirt_on_the_fly_siddon_cone():
    import torch
    from vir_amd.proj_geom import geom_circular
    from vir_amd.sirt import SIRT
    from vir_amd.sys_mat.siddon import SiddonProjector
    v = np.arange(16) * 2 * np.pi / 16
    du = (np.arange(30) - 14.5) * 1.0
    dv = (np.arange(14) - 6.5) * 1.0
    s, t = geom_circular(du, v, "cone", src_iso=60.0, det_iso=60.0, DetsZ=dv)
    nP, dP = (20, 20, 10), 0.5
    rng = np.random.default_rng(1)
    x_true = rng.uniform(0, 1, nP).astype(np.float32)
    ind, ln, off = osd.siddons(s, t, nP, dP, 0.0, flat=True)
    A = sirt_ref.csr_matrix(ind, ln, off, int(np.prod(nP)))
    b = (A @ x_true.reshape(-1).astype(np.float64)).astype(np.float32)
    P = SiddonProjector(s, t, nP, dP)
    S = SIRT(P, torch.from_numpy(b).cuda())
    x = S.run(15).cpu().numpy()
    ref = sirt_ref.sirt(A, b, 15)
    assert rel(x, ref) <= TOL


def test_sirt_dd_fan_and_graph_replay():
    import torch
    from vir_amd.sirt import SIRT
    from vir_amd.sys_mat.dd import DDProjector
    nx = ny = 24
    nu, na = 40, 30
    ang = np.linspace(0, 2 * np.pi, na, endpoint=False).astype(np.float32)
    DSO, DSD, du = 60.0, 90.0, 1.2
    A = dd_ref.overlap_brute_2d(nx, ny, ang, nu, du, 0.0, 1.0, geom="fan",
                                DSO=DSO, DSD=DSD)
    rng = np.random.default_rng(2)
    x_true = np.zeros((nx, ny), np.float32)
    x_true[6:18, 5:19] = rng.uniform(0.2, 1.0, (12, 14))
    b = (A @ x_true.reshape(-1).astype(np.float64)).astype(np.float32).reshape(na, nu)
    P = DDProjector((nx, ny), ang, nu, "fan", DSO=DSO, DSD=DSD, du=du)
    S = SIRT(P, torch.from_numpy(b).cuda(), x0=np.full((nx, ny), 1e-5, np.float32))
    x = S.run(20).cpu().numpy()
    ref = sirt_ref.sirt(A, b, 20, x0=np.full(nx * ny, np.float32(1e-5)))
    assert rel(x, ref) <= TOL
    # graph replay == eager, several replays, eager work in between
    Sg = SIRT(P, torch.from_numpy(b).cuda(), graph=True)
    Se = SIRT(P, torch.from_numpy(b).cuda(), graph=False)
    for k in range(5):
        xg = Sg.step().clone()
        xe = Se.step().clone()
        torch.cuda.synchronize()
        junk = torch.randn(1000, device="cuda").sum()  # unrelated eager work
        assert torch.equal(xg, xe), k
        del junk


def test_tv_denoise_reduces_total_variation():
    import torch
    from vir_amd.tv import denoise_tv_chambolle
    rng = np.random.default_rng(4)
    img = np.zeros((32, 32), np.float32)
    img[8:24, 8:24] = 1.0
    noisy = img + rng.normal(0, 0.2, img.shape).astype(np.float32)
    den = denoise_tv_chambolle(torch.from_numpy(noisy).cuda(), weight=0.2).cpu().numpy()

    def tv(a):
        return np.abs(np.diff(a, axis=0)).sum() + np.abs(np.diff(a, axis=1)).sum()
    assert tv(den) < 0.5 * tv(noisy)
    assert np.linalg.norm(den - img) < np.linalg.norm(noisy - img)


def _fused_case(kind):
    import torch
    from vir_amd.geometry import circular_projections, circular_rays
    from vir_amd.serve import cone_geometry, fan_geometry
    from vir_amd.sys_mat.dd import DDProjector
    from vir_amd.sys_mat.siddon import SiddonProjector
    from vir_amd.workloads import phantom_device
    if kind == "fan":
        g = fan_geometry(96, 60)
        P = DDProjector(g["img_shape"], g["ang"], g["nu"], "fan", DSO=g["DSO"], DSD=g["DSD"],
                        du=g["du"], detector="arc")
    else:
        # 24^3: the brick gather splits the views into chunks (partial
        # volumes summed by the reduction kernel); 192^3: one chunk, the
        # update is written by the gather itself
        n, views = (24, 20) if kind == "cone_chunked" else (192, 4)
        g = cone_geometry(n, views)
        src, trg = circular_rays(g["dets_y"], g["views"], "cone", g["src_iso"], g["det_iso"],
                                 g["dets_z"])
        mats = circular_projections(g["dets_y"], g["views"], "cone", g["src_iso"], g["det_iso"],
                                    g["dets_z"])
        P = SiddonProjector(src, trg, g["img_shape"], 1.0, proj=mats)
    b = P.fp(phantom_device(g["img_shape"], 8, 21))
    torch.cuda.synchronize()
    return P, b


@pytest.mark.parametrize("kind", ["cone_chunked", "cone_direct", "fan"])
def test_fused_sirt_steps_equal_separate_steps(kind):
    """The residual written by the forward projection and the update
    written by the backprojection give the iterates of the separate
    elementwise steps bit for bit; graph replays give the eager iterates."""
    import torch
    from vir_amd import _lib
    from vir_amd.sirt import SIRT
    P, b = _fused_case(kind)
    if kind == "cone_chunked":
        assert int(_lib.lib().vir_siddon_bp_workspace(P.grid, P.layout[0])) > 0
    elif kind == "cone_direct":
        assert int(_lib.lib().vir_siddon_bp_workspace(P.grid, P.layout[0])) == 0
    gamma = 0.9 if kind == "fan" else 1.0
    ref = SIRT(P, b, gamma=gamma, fused=False)
    B = SIRT(P, b, gamma=gamma, fused=True)
    G = SIRT(P, b, gamma=gamma, fused=True, graph=True)
    assert B.fused and not ref.fused and G.fused
    for it in range(3):
        ref.step()
        B.step()
        G.step()
        torch.cuda.synchronize()
        xr, xb, xg = (S.x.cpu().numpy() for S in (ref, B, G))
        assert np.abs(xr).max() > 0
        assert np.array_equal(xr.view(np.int32), xb.view(np.int32)), (kind, it)
        assert np.array_equal(xr.view(np.int32), xg.view(np.int32)), (kind, it)
    # the fused entry points refuse VIR_ACCUMULATE
    import ctypes
    from vir_amd.device import stream
    if kind == "fan":
        rc = _lib.lib().vir_dd2d_bp_upd(P.g, _lib.ptr(P.cos), _lib.ptr(P.sin), _lib.ptr(B.r),
                                        _lib.ptr(B.W_img), ctypes.c_float(1.0), _lib.ptr(B.x),
                                        P.fp_scale, _lib.VIR_ACCUMULATE, stream())
    else:
        rc = _lib.lib().vir_siddon_fp_rec_res(_lib.ptr(P.rec), P.nrays, P.grid, _lib.ptr(B.x),
                                              _lib.ptr(b), _lib.ptr(B.W_sino), _lib.ptr(B.r),
                                              _lib.VIR_ACCUMULATE, stream())
    assert rc == _lib.VIR_EINVAL


@pytest.mark.parametrize("kind", ["cone_chunked", "cone_direct"])
def test_batched_sirt_equals_one_sirt_per_item(kind):
    """SIRT(nbatch=3) on one projector -- every projection tracing the rays
    once for the batch -- gives each item's iterates of its own SIRT bit for
    bit, eagerly and as graph replays; projectors without batched steps are
    refused."""
    import torch
    from vir_amd.sirt import SIRT
    P, b = _fused_case(kind)
    bs = torch.stack([b * (1.0 + 0.5 * j) for j in range(3)])
    own = [SIRT(P, bs[j], gamma=0.8) for j in range(3)]
    Bt = SIRT(P, bs, gamma=0.8, nbatch=3)
    Gt = SIRT(P, bs, gamma=0.8, nbatch=3, graph=True)
    for it in range(3):
        for S in own + [Bt, Gt]:
            S.step()
        torch.cuda.synchronize()
        for j in range(3):
            xo = own[j].x.cpu().numpy().view(np.int32)
            assert np.array_equal(xo, Bt.x[j].cpu().numpy().view(np.int32)), (kind, it, j)
            assert np.array_equal(xo, Gt.x[j].cpu().numpy().view(np.int32)), (kind, it, j)
    F, bf = _fused_case("fan")
    with pytest.raises(ValueError):
        SIRT(F, torch.stack([bf, bf]), nbatch=2)
