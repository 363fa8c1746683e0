"""Differentiable projection and loss terms on the GPU.

* TV value and gradient against a float64 torch restatement of the same
  formula (autograd on the CPU);
* least-squares value/gradient against float64 numpy;
* project(): gradient of <y, A x> equals A^T y, dot-product test 1e-5;
* the full loss 0.5*||A x - b||^2 + 1e-3*TV(x) on a batched 3-D parallel
  separable-footprint system: autograd gradient against central finite
  differences on 64 random voxels (seeded).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _tv64(x, eps):
    """sum sqrt(|D x|^2 + eps^2) - eps, forward differences along the last
    three axes, zero at the last index (float64, CPU)."""
    import torch
    d = []
    for ax in (-3, -2, -1):
        g = torch.zeros_like(x)
        n = x.shape[ax]
        hi = [slice(None)] * x.dim()
        lo = [slice(None)] * x.dim()
        hi[ax] = slice(1, n)
        lo[ax] = slice(0, n - 1)
        g[tuple(lo)] = x[tuple(hi)] - x[tuple(lo)]
        d.append(g)
    return (torch.sqrt(d[0] ** 2 + d[1] ** 2 + d[2] ** 2 + eps * eps) - eps).sum()


def test_total_variation_value_and_gradient():
    import torch
    from vir_amd.autograd import total_variation
    rng = np.random.default_rng(0)
    x = rng.normal(size=(3, 17, 12, 9)).astype(np.float32)
    x[0, 4:10, 3:8, 2:6] = 1.0       # a flat block: |D x| = 0 inside
    eps = 1e-2
    xt = torch.from_numpy(x).cuda().requires_grad_(True)
    v = total_variation(xt, eps=eps)
    v.backward()
    x64 = torch.from_numpy(x).double().requires_grad_(True)
    ref = _tv64(x64, eps)
    ref.backward()
    assert abs(float(v.detach()) - float(ref.detach())) <= 1e-6 * abs(float(ref.detach()))
    g, gr = xt.grad.cpu().double(), x64.grad
    assert float((g - gr).norm() / gr.norm()) <= 1e-5
    # 2-D images: TV over the last two axes only
    y = torch.from_numpy(rng.normal(size=(20, 14)).astype(np.float32)).cuda()
    v2 = float(total_variation(y, eps=eps, ndim=2))
    r2 = float(_tv64(torch.from_numpy(y.cpu().numpy()).double()[None], eps))
    assert abs(v2 - r2) <= 1e-6 * abs(r2)


def test_least_squares():
    import torch
    from vir_amd.autograd import least_squares
    rng = np.random.default_rng(1)
    a = rng.normal(size=100003).astype(np.float32)
    b = rng.normal(size=100003).astype(np.float32)
    at = torch.from_numpy(a).cuda().requires_grad_(True)
    v = least_squares(at, torch.from_numpy(b).cuda())
    (3.0 * v).backward()
    d = a.astype(np.float64) - b
    assert abs(float(v.detach()) - 0.5 * (d * d).sum()) <= 1e-9 * (d * d).sum()
    np.testing.assert_allclose(at.grad.cpu().numpy(), 3.0 * (a - b), rtol=1e-6)


def test_project_gradient_is_backprojection():
    import torch
    from vir_amd.autograd import project
    from vir_amd.sys_mat.pd import SFProjector
    ang = np.linspace(0, np.pi, 24, endpoint=False)
    P = SFProjector((20, 20, 8), ang, 32, "par3d", nv=8, batch=2)
    rng = np.random.default_rng(2)
    x = torch.from_numpy(rng.normal(size=(2, 20, 20, 8)).astype(np.float32)).cuda()
    y = torch.from_numpy(rng.normal(size=(2, 24, 32, 8)).astype(np.float32)).cuda()
    xr = x.clone().requires_grad_(True)
    (project(xr, P) * y).sum().backward()
    assert torch.equal(xr.grad, P.bp(y))
    lhs = float((P.fp(x).double() * y.double()).sum())
    rhs = float((x.double() * P.bp(y).double()).sum())
    assert abs(lhs - rhs) <= 1e-5 * abs(lhs)


def test_loss_gradient_matches_central_differences():
    import torch
    from vir_amd.autograd import least_squares, project, total_variation
    from vir_amd.sys_mat.pd import SFProjector
    from vir_amd.workloads import phantom
    nb, shape = 2, (32, 32, 16)
    ang = np.arange(30) * np.pi / 30
    P = SFProjector(shape, ang, 48, "par3d", nv=16, batch=nb)
    truth = torch.stack([phantom(shape, 6, seed=s, device="cuda") for s in range(nb)])
    b = P.fp(truth)
    rng = np.random.default_rng(3)
    x0 = (truth + torch.from_numpy(rng.uniform(-0.2, 0.2, truth.shape).astype(np.float32)).cuda())
    lam, eps = 1e-3, 1e-3

    def loss(x):
        return least_squares(project(x, P), b) + lam * total_variation(x, eps=eps)

    x = x0.clone().requires_grad_(True)
    loss(x).backward()
    g = x.grad.reshape(-1)
    idx = rng.choice(x0.numel(), 64, replace=False)
    h = 5e-2   # the data term is quadratic: central differences are exact but for rounding
    flat = x0.reshape(-1)
    fd = []
    for i in idx:
        xp = flat.clone()
        xp[i] += h
        xm = flat.clone()
        xm[i] -= h
        fp_ = float(loss(xp.reshape(x0.shape)))
        fm_ = float(loss(xm.reshape(x0.shape)))
        fd.append((fp_ - fm_) / (2 * h))
    fd = np.array(fd)
    ga = g[torch.from_numpy(idx).cuda()].double().cpu().numpy()
    scale = np.abs(ga).max()
    assert np.abs(fd - ga).max() <= 1e-3 * scale, (fd, ga)
