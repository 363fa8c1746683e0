"""Differentiable projection and reconstruction losses.

`project(x, P)` applies any projector with an exact adjoint pair
(SFProjector, DDProjector, SiddonProjector): forward P.fp, backward P.bp,
both hand-written HIP kernels.  `least_squares` and `total_variation` are
loss terms whose values are reduced in double on the device and whose
gradients come from their own kernels (libvir_amd.so vir_sqdist, vir_tv).

    x.requires_grad_(True)
    loss = least_squares(project(x, P), b) + 1e-3 * total_variation(x)
    loss.backward()

The reference reconstructs with SIRT and a TV prox (tv.py:200-294,
tests/Test_SIRT/sirt_test.py:37-38); this module is the gradient form of
the same data term and regulariser.
"""
import torch

from . import _lib
from .device import require_cuda_f32, stream

TV_WORK = 2048  # doubles, VIR_TV_WORK


def _work(device):
    return torch.empty(TV_WORK, dtype=torch.float64, device=device)


class Project(torch.autograd.Function):
    """y = A x with dL/dx = A^T dL/dy."""

    @staticmethod
    def forward(ctx, x, P):
        ctx.P = P
        return P.fp(x.detach())

    @staticmethod
    def backward(ctx, g):
        return ctx.P.bp(g.contiguous()), None


def project(x, P):
    return Project.apply(x, P)


class LeastSquares(torch.autograd.Function):
    """0.5 * ||a - b||^2 as a float64 scalar; dL/da = a - b, dL/db = b - a."""

    @staticmethod
    def forward(ctx, a, b):
        require_cuda_f32(a, b, what="least_squares inputs")
        if a.device != b.device:
            raise ValueError("least_squares inputs are on different devices")
        a = a.detach().contiguous()
        b = b.detach().contiguous()
        if a.numel() != b.numel():
            raise ValueError("a and b differ in size")
        r = torch.empty_like(a)
        val = torch.empty((), dtype=torch.float64, device=a.device)
        _lib.check(_lib.lib().vir_sqdist(_lib.ptr(a), _lib.ptr(b), _lib.ptr(r), a.numel(),
                                         _lib.ptr(val), _lib.ptr(_work(a.device)), stream()),
                   "vir_sqdist")
        ctx.save_for_backward(r)
        return val

    @staticmethod
    def backward(ctx, gv):
        (r,) = ctx.saved_tensors
        ga = r * gv.to(r.dtype)
        gb = -ga if ctx.needs_input_grad[1] else None
        return (ga if ctx.needs_input_grad[0] else None), gb


def least_squares(a, b):
    return LeastSquares.apply(a, b)


def _tv_shape(x, ndim):
    s = tuple(x.shape)
    if ndim is None:
        ndim = 3 if len(s) >= 3 else 2
    if ndim not in (2, 3) or len(s) < ndim:
        raise ValueError("total_variation needs 2 or 3 trailing spatial dimensions")
    sp = s[len(s) - ndim:]
    nb = 1
    for v in s[:len(s) - ndim]:
        nb *= v
    nx, ny = sp[0], sp[1]
    nz = sp[2] if ndim == 3 else 1
    return nb, nx, ny, nz


class TotalVariation(torch.autograd.Function):
    """sum_i sqrt(|D x|_i^2 + eps^2) - eps over each volume of a batch
    (forward differences along the last `ndim` axes), a float64 scalar."""

    @staticmethod
    def forward(ctx, x, eps, ndim):
        require_cuda_f32(x, what="total_variation input")
        x = x.detach().contiguous()
        nb, nx, ny, nz = _tv_shape(x, ndim)
        grad = torch.empty_like(x)
        val = torch.empty((), dtype=torch.float64, device=x.device)
        _lib.check(_lib.lib().vir_tv(_lib.ptr(x), nb, nx, ny, nz, float(eps), 1.0, _lib.ptr(grad),
                                     _lib.ptr(val), _lib.ptr(_work(x.device)), 0, stream()),
                   "vir_tv")
        ctx.save_for_backward(grad)
        return val

    @staticmethod
    def backward(ctx, gv):
        (grad,) = ctx.saved_tensors
        return grad * gv.to(grad.dtype), None, None


def total_variation(x, eps=1e-3, ndim=None):
    return TotalVariation.apply(x, eps, ndim)
