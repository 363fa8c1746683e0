"""GPU parity at the benchmark workloads' own sizes (a subset of views where
the float64 oracle would otherwise take too long), through the C ABI.

  cone 256^3, 360 views, 512 x 384 detector (Siddon FP + brick-gather BP):
      four full views: the GPU trace equals the C oracle's bit for bit; FP
      and BP within 1e-5 of float64 products over that trace; adjoint;
  parallel 128^2, 180 views x 185 bins: sd_b_proj(compat=True,
      sequential=True) equals the reference's numpy backprojector bit for bit;
  fan 512^2, 720 views, 888 equiangular bins (distance-driven): eight views
      against the float64 overlap matrix, and two SIRT iterations against the
      float64 recurrence with that matrix;
  parallel 3-D 8 x 256 x 256 x 64, 180 views, 384 x 64 detector (separable
      footprint, the autograd workload): four views, every batch item,
      against the float64 separable operator;
  serving trace (Poisson 200/s, seed 1005): the first twelve requests --
      cone 256^3 and 64^3, fan 1024^2 down to 128^2 -- give the same bits
      alone and in the mix; the smallest fan and cone requests are within
      1e-5 of the float64 SIRT recurrence.
"""
import numpy as np
import pytest

from oracle import dd_ref, pd_ref, sirt_ref
from oracle import projection as op
from oracle import siddon as osd

pytestmark = pytest.mark.gpu
TOL = 1e-5


def rel(a, b):
    a = np.asarray(a, np.float64).reshape(-1)
    b = np.asarray(b, np.float64).reshape(-1)
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def _say(msg):
    print(msg, flush=True)


def test_cone256_four_views_trace_fp_bp_adjoint():
    import torch
    from vir_amd.sys_mat.siddon import SiddonProjector, siddons
    from vir_amd.workloads import cone256_geometry, phantom
    views = [0, 45, 101, 270]
    s, t, g = cone256_geometry(views)
    nP, dP = g["nPixels"], g["dPixels"]
    _say("cone256: oracle trace of 4 views")
    ind, ln, off = osd.siddons(s, t, nP, dP, 0.0, flat=True)
    _say(f"cone256: {ind.size} entries; GPU trace")
    gi, gl, go = siddons(s, t, nP, dP, 0.0, flat=True)
    assert np.array_equal(go, off) and np.array_equal(gi, ind)
    assert np.array_equal(gl.view(np.int32), ln.view(np.int32))
    del gi, gl, go
    P = SiddonProjector(s, t, nP, dP, structured=True)
    assert P.proj is not None and P.guarded_rays == 0
    vol = phantom(nP, 20, seed=1002)
    rng = np.random.default_rng(7)
    y = rng.uniform(0, 1, P.ray_shape).astype(np.float32)
    _say("cone256: GPU FP/BP")
    fx = P.fp(vol).cpu().numpy().reshape(-1)
    by = P.bp(torch.from_numpy(y).cuda()).cpu().numpy().reshape(-1)
    _say("cone256: float64 products")
    A = op.CsrOperator(ind, ln, off, int(np.prod(nP)))
    x64 = vol.cpu().numpy().reshape(-1).astype(np.float64)
    f64 = A @ x64
    b64 = A.T @ y.reshape(-1)
    assert rel(fx, f64) <= 1e-6
    assert rel(by, b64) <= 1e-6
    assert np.max(np.abs(by - b64)) <= 1e-5 * np.abs(b64).max()
    lhs = float(np.dot(fx.astype(np.float64), y.reshape(-1)))
    rhs = float(np.dot(x64, by.astype(np.float64)))
    assert abs(lhs - rhs) <= 1e-5 * abs(lhs)


def test_par128_compat_backprojection_is_the_reference_numpy_one():
    from vir_amd.projection.proj import sd_b_proj
    from vir_amd.sys_mat.siddon import siddons
    from vir_amd.workloads import par128_geometry
    s, t, g = par128_geometry()
    nP = g["nPixels"]
    sd = siddons(s, t, nP, g["dPixels"], ravel=True)
    ind, ln, off = siddons(s, t, nP, g["dPixels"], flat=True)
    rng = np.random.default_rng(3)
    y = rng.uniform(0, 1, sd.shape).astype(np.float32)
    ref = op.sd_b_proj_numpy(y, ind, ln, off, nP)
    got = sd_b_proj(y, sd, nP, compat=True, sequential=True)
    assert got.dtype == np.float32 and got.shape == ref.shape
    assert np.array_equal(got.view(np.int32), ref.view(np.int32))
    # the default (exact transpose) differs exactly where a ray repeats a voxel
    full = sd_b_proj(y, sd, nP)
    nvox = int(np.prod(nP))
    rows = np.repeat(np.arange(off.size - 1), np.diff(off))
    assert np.unique(rows.astype(np.int64) * nvox + ind).size < ind.size   # the workload has repeats
    assert rel(full, op.bp64(y, ind, ln, off, nvox)) <= 1e-6
    assert not np.array_equal(full, got)


def _fan512(views):
    from vir_amd.serve import fan_geometry
    g = fan_geometry(512, 720)
    ang = g["ang"][views]
    A = dd_ref.overlap_sparse_2d(512, 512, ang, g["nu"], float(np.float32(g["du"])), 0.0, 1.0,
                                 geom="arc", DSO=float(np.float32(g["DSO"])),
                                 DSD=float(np.float32(g["DSD"])))
    return g, ang, A


def test_fan512_eight_views_and_two_sirt_iterations():
    import torch
    from vir_amd.sirt import SIRT
    from vir_amd.sys_mat.dd import DDProjector
    from vir_amd.workloads import phantom
    views = np.array([0, 45, 90, 135, 180, 333, 500, 719])
    g, ang, A = _fan512(views)
    P = DDProjector(g["img_shape"], ang, g["nu"], "fan", DSO=g["DSO"], DSD=g["DSD"], du=g["du"],
                    detector="arc")
    x = phantom((512, 512), 10, seed=1003)
    fx = P.fp(x).cpu().numpy()
    assert fx.shape == (8, 888)
    assert rel(fx, A @ x.cpu().numpy().reshape(-1).astype(np.float64)) <= TOL
    rng = np.random.default_rng(4)
    y = rng.uniform(0, 1, (8, 888)).astype(np.float32)
    by = P.bp(torch.from_numpy(y).cuda()).cpu().numpy()
    assert rel(by, A.T @ y.reshape(-1).astype(np.float64)) <= TOL
    b = fx
    S = SIRT(P, b)
    S.run(2)
    ref = sirt_ref.sirt(A, b, 2)
    assert rel(S.x.cpu().numpy(), ref) <= TOL


def test_par3d_sf_workload_four_views_all_batch_items():
    import torch
    from vir_amd.sys_mat.pd import SFProjector
    from vir_amd.workloads import phantom_device
    shape, nb = (256, 256, 64), 8
    ang = (np.arange(180) * np.pi / 180)[[0, 45, 97, 150]]
    P = SFProjector(shape, ang, 384, "par3d", nv=64, du=1.0, dv=1.0, batch=nb)
    vols = torch.stack([phantom_device(shape, 20, 1004 + k) for k in range(nb)])
    fp_, wt = pd_ref.par3d_ops(shape, ang, 384, 64, du=1.0, dv=1.0, d_pix=1.0)
    fx = P.fp(vols).cpu().numpy()
    rng = np.random.default_rng(5)
    y = rng.uniform(0, 1, P.sino_shape).astype(np.float32)
    by = P.bp(torch.from_numpy(y).cuda()).cpu().numpy()
    vh = vols.cpu().numpy()
    for k in range(nb):
        _say(f"par3d: batch item {k}")
        assert rel(fx[k], fp_(vh[k])) <= TOL
        assert rel(by[k], wt(y[k])) <= TOL       # fp_scale = d_pix = 1


def test_serving_trace_prefix_alone_equals_mix_and_fp64():
    import torch
    from vir_amd.serve import (Server, SirtBackend, cone_geometry, fan_geometry, poisson_trace,
                               run_alone)
    from vir_amd.sys_mat.dd import DDProjector
    from vir_amd.sys_mat.siddon import SiddonProjector
    from vir_amd.proj_geom import geom_circular
    from vir_amd.workloads import phantom_device
    reqs = poisson_trace(10000, 200.0, seed=1005)[:12]
    be = SirtBackend(n_streams=8)
    srv = Server(be, 200 << 30, max_active=16)
    _say("serving: the mix")
    res = srv.run(reqs)
    assert all(r.ok for r in res), [r.error for r in res if not r.ok]
    assert srv.max_concurrent > 1
    for r, q in zip(res, reqs):
        _say(f"serving: request {q.rid} ({q.kind} {q.n} x {q.views}) alone")
        assert np.array_equal(r.x, run_alone(q)), q
    # float64 recurrence, smallest fan request (128^2 x 402 views)
    q = min((q for q in reqs if q.kind == "fan"), key=lambda q: q.n * q.n * q.views)
    _say(f"serving: fan request {q.rid} against float64")
    g = fan_geometry(q.n, q.views)
    A = dd_ref.overlap_sparse_2d(q.n, q.n, g["ang"], g["nu"], float(np.float32(g["du"])), 0.0, 1.0,
                                 geom="arc", DSO=float(np.float32(g["DSO"])),
                                 DSD=float(np.float32(g["DSD"])))
    P = DDProjector(g["img_shape"], g["ang"], g["nu"], "fan", DSO=g["DSO"], DSD=g["DSD"],
                    du=g["du"], detector="arc")
    b = P.fp(phantom_device(g["img_shape"], 10, q.seed)).cpu().numpy()
    assert rel(res[q.rid].x, sirt_ref.sirt(A, b, q.iters)) <= TOL
    # smallest cone request (64^3 x 288 views, 96 x 96 detector)
    q = min((q for q in reqs if q.kind == "cone"), key=lambda q: q.n ** 3 * q.views)
    _say(f"serving: cone request {q.rid} against float64")
    g = cone_geometry(q.n, q.views)
    s_h, t_h = geom_circular(g["dets_y"], g["views"], "cone", g["src_iso"], g["det_iso"],
                             g["dets_z"])
    ind, ln, off = osd.siddons(s_h, t_h, g["img_shape"], 1.0, 0.0, flat=True)
    A = op.CsrOperator(ind, ln, off, q.n ** 3)
    Ph = SiddonProjector(s_h, t_h, g["img_shape"], 1.0)
    b = Ph.fp(phantom_device(g["img_shape"], 20, q.seed)).cpu().numpy()
    assert rel(res[q.rid].x, sirt_ref.sirt(A, b, q.iters)) <= TOL
    torch.cuda.synchronize()
