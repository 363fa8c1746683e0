"""GPU separable-footprint prototypes (sys_mat/sp.py, sys_mat/fp_2d_par.py)
against float64 oracles:

  sf_fp_fan_2d / sf_bp_fan_2d: oracle/sp_ref.fan_matrix (pinned by the
      triangle-mass identity and the parallel limit), adjoint to 1e-5;
  forward_project / back_project (the separable toy): oracle/sp_ref.sep_*;
  separable_footprint_fp_2d: pd_ref's parallel footprint matrix without its
      amplitude (pd_ref pinned in test_pd_oracle.py);
  siddons_fp_2d: path lengths from the C Siddon oracle along the same rays.
"""
import numpy as np
import pytest

from oracle import pd_ref, sp_ref
from oracle import projection as op
from oracle import siddon as osd

pytestmark = pytest.mark.gpu
TOL = 1e-5


def rel(a, b):
    a = np.asarray(a, np.float64).reshape(-1)
    b = np.asarray(b, np.float64).reshape(-1)
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.mark.parametrize("nx,ny,na,nu,DSO,DSD,du,d", [
    (40, 33, 24, 70, 80.0, 130.0, 0.9, 1.1),
    (256, 256, 12, 444, 500.0, 750.0, 1.0, 1.0),
])
def test_triangle_fan_pair(nx, ny, na, nu, DSO, DSD, du, d):
    import torch
    from vir_amd.sys_mat.sp import SPFanProjector, sf_bp_fan_2d, sf_fp_fan_2d
    rng = np.random.default_rng(nx)
    angs = np.sort(rng.uniform(0, 2 * np.pi, na))
    W = sp_ref.fan_matrix(nx, ny, angs, nu, DSO, DSD, du=du, d_pix=d)
    img = rng.uniform(0, 1, (nx, ny)).astype(np.float32)
    img[: nx // 5] = 0.0                      # zero pixels (skipped by the reference)
    sino = sf_fp_fan_2d(img, angs, nu, DSO, DSD, du=du, d_pix=d)
    assert sino.shape == (na, nu) and sino.dtype == np.float32
    assert rel(sino, d * (W @ img.reshape(-1).astype(np.float64))) <= TOL
    y = rng.uniform(0, 1, (na, nu)).astype(np.float32)
    bp = sf_bp_fan_2d(y, angs, (nx, ny), DSO, DSD, du=du, d_pix=d)
    assert rel(bp, d * (W.T @ y.reshape(-1).astype(np.float64))) <= TOL
    P = SPFanProjector((nx, ny), angs, nu, DSO, DSD, du=du, d_pix=d)
    x = torch.from_numpy(rng.normal(size=(nx, ny)).astype(np.float32)).cuda()
    yy = torch.from_numpy(rng.normal(size=(na, nu)).astype(np.float32)).cuda()
    lhs = float((P.fp(x).double() * yy.double()).sum())
    rhs = float((x.double() * P.bp(yy).double()).sum())
    assert abs(lhs - rhs) <= 1e-5 * abs(lhs)


def test_separable_toy_pair():
    from vir_amd.sys_mat.sp import CTGeometry, back_project, forward_project
    geom = CTGeometry((20, 18, 12), (1.0, 1.1, 0.9), (40, 36), (0.8, 0.7), 30.0, 20.0)
    g = (20, 18, 12, 40, 36, 1.0, 1.1, 0.9, 0.8, 0.7, 30.0, 50.0)
    rng = np.random.default_rng(3)
    vol = rng.uniform(0, 1, (20, 18, 12)).astype(np.float32)
    proj = forward_project(vol, geom)
    assert proj.shape == (36, 40) and proj.dtype == np.float32
    assert rel(proj, sp_ref.sep_fp(vol, *g)) <= 1e-6
    y = rng.uniform(0, 1, (36, 40)).astype(np.float32)
    back = back_project(y, geom)
    assert back.shape == (20, 18, 12)
    assert rel(back, sp_ref.sep_bp(y, *g)) <= 1e-6


def test_rectangle_footprint_parallel():
    from vir_amd.sys_mat.fp_2d_par import separable_footprint_fp_2d
    rng = np.random.default_rng(4)
    nx, ny, nu, d_det, d = 30, 26, 60, 0.9, 1.1
    angs = np.linspace(0, np.pi, 25, endpoint=False)
    W = pd_ref.matrix("par", (nx, ny), angs, nu, du=d_det, d_pix=d)
    amp = 1.0 / (np.abs(np.sin(angs)) + np.abs(np.cos(angs)))
    W = W / np.repeat(amp, nu)[:, None]
    img = rng.uniform(0, 1, (nx, ny)).astype(np.float32)
    got = separable_footprint_fp_2d(img, angs, nu, d_det=d_det, d_pix=d)
    assert got.shape == (25, nu)
    assert rel(got, W @ img.reshape(-1).astype(np.float64)) <= TOL


def test_siddon_parallel_prototype():
    from vir_amd.sys_mat.fp_2d_par import siddons_fp_2d
    rng = np.random.default_rng(5)
    nx, ny, nd, d_det, d = 32, 28, 50, 0.9, 1.05
    angs = np.linspace(0.05, np.pi, 23, endpoint=False)
    img = rng.uniform(0, 1, (nx, ny)).astype(np.float32)
    got = siddons_fp_2d(img, angs, nd, d_det=d_det, d_pix=d)
    det = d_det * (np.arange(nd) - nd / 2 + 0.5)
    c, s = np.cos(angs)[:, None], np.sin(angs)[:, None]
    L = 100.0
    src = np.stack([-det * s - L * c, det * c - L * s, np.zeros((23, nd))], -1)
    trg = np.stack([-det * s + L * c, det * c + L * s, np.zeros((23, nd))], -1)
    ind, ln, off = osd.siddons(src, trg, (nx, ny, 1), d, 0.0, flat=True)
    ref = op.fp64(img, ind, ln, off).reshape(23, nd)
    assert rel(got, ref) <= TOL
