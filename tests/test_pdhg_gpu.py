"""GPU Chambolle-Pock TV reconstruction (vir_amd.tv.pdhg_tv) against the
float64 recurrence (oracle/pdhg_ref.py) with the same system matrix: the
iterate after 40 iterations within 1e-5 relative L2, on a stored-matrix-free
Siddon projector (2-D parallel) and a distance-driven fan projector."""
import numpy as np
import pytest

from oracle import dd_ref, pdhg_ref, sirt_ref
from oracle import siddon as osd

pytestmark = pytest.mark.gpu
TOL = 1e-5


def rel(a, b):
    a = np.asarray(a, np.float64).reshape(-1)
    b = np.asarray(b, np.float64).reshape(-1)
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def _truth(n, seed):
    rng = np.random.default_rng(seed)
    x = np.zeros((n, n), np.float32)
    x[n // 4:3 * n // 4, n // 3:2 * n // 3] = 1.0
    x += 0.05 * rng.normal(size=x.shape).astype(np.float32)
    return x


def test_pdhg_siddon_parallel_matches_fp64():
    import torch
    from vir_amd.proj_geom import geom_circular
    from vir_amd.sys_mat.siddon import SiddonProjector
    from vir_amd.tv import pdhg_tv
    n = 24
    s, t = geom_circular((np.arange(35) - 17.0) * 1.0, np.arange(30) * np.pi / 30 + 0.01, "par")
    ind, ln, off = osd.siddons(s, t, (n, n, 1), 1.0, 0.0, flat=True)
    A = sirt_ref.csr_matrix(ind, ln, off, n * n).toarray()
    x = _truth(n, 0)
    g = (A @ x.reshape(-1).astype(np.float64)).astype(np.float32)
    P = SiddonProjector(s, t, (n, n, 1), 1.0)
    u = pdhg_tv(P, torch.from_numpy(g).cuda(), iters=40, lam=0.5).cpu().numpy()
    ref = pdhg_ref.pdhg(A, g, (n, n), 40, lam=0.5)
    assert rel(u.reshape(n, n), ref) <= TOL


def test_pdhg_dd_fan_matches_fp64():
    import torch
    from vir_amd.sys_mat.dd import DDProjector
    from vir_amd.tv import pdhg_tv
    # detector inside the inscribed circle: no grazing rays, whose tiny A 1
    # would make sigma_q = 1/(A 1) amplify float32 rounding
    n, nu, na = 20, 24, 24
    ang = np.linspace(0, 2 * np.pi, na, endpoint=False).astype(np.float32)
    A = dd_ref.overlap_brute_2d(n, n, ang, nu, 1.2, 0.0, 1.0, geom="fan",
                                DSO=50.0, DSD=80.0)
    x = _truth(n, 1)
    g = (A @ x.reshape(-1).astype(np.float64)).astype(np.float32)
    P = DDProjector((n, n), ang, nu, "fan", DSO=50.0, DSD=80.0, du=1.2)
    assert (A @ np.ones(n * n)).min() > 5.0
    u = pdhg_tv(P, torch.from_numpy(g.reshape(na, nu)).cuda(), iters=40, lam=1.0).cpu().numpy()
    ref = pdhg_ref.pdhg(A, g, (n, n), 40, lam=1.0)
    assert rel(u, ref) <= TOL
