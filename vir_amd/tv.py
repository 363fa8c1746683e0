"""Total-variation denoising and TV-regularised reconstruction on the GPU.

denoise_tv_chambolle: the denoiser the reference applies to SIRT iterates
(scikit-image's denoise_tv_chambolle, tests/Test_SIRT/sirt_test.py:13,38;
not installed here) restated from the published algorithm -- A. Chambolle,
"An algorithm for total variation minimization and applications", J. Math.
Imaging Vis. 20 (2004) 89-97, eq. (9) -- and run by libvir_amd's
vir_tv_chambolle kernels.

pdhg_tv: the Chambolle-Pock reconstruction of tv.py:229-294 over any
projector.
"""
import numpy as np
import torch

TV_WORK = 2048  # floats, VIR_TV_WORK


def denoise_tv_chambolle(image, weight=0.1, eps=2.0e-4, max_num_iter=200, tau=None, check_every=10):
    """u = argmin ||u - image||^2 / 2 + weight * TV(u) over an image of 1 to 3
    axes (isotropic TV, forward differences), by Chambolle's dual
    projection iteration with step tau (default 1/(4 ndim), the step the
    paper proves convergent).

    Stopping rule: every `check_every` iterations the largest change of the
    dual variable q = weight p is read back; the loop ends when it is at
    most eps * weight (|p| <= 1), or after max_num_iter iterations.  eps=0
    runs exactly max_num_iter.  (scikit-image's eps bounds a relative
    energy change instead; the two rules stop at different points.)
    torch in -> torch (float32, same device) out; numpy in -> numpy out."""
    from . import _lib
    from .device import as_device_f32, stream
    torch_in = isinstance(image, torch.Tensor)
    g = as_device_f32(image)
    shape = tuple(g.shape)
    nd = len(shape)
    if nd < 1 or nd > 3:
        raise ValueError("denoise_tv_chambolle handles images of 1 to 3 axes")
    if not weight > 0:
        raise ValueError("weight must be > 0")
    dims = shape + (1,) * (3 - nd)
    tau = 1.0 / (4.0 * nd) if tau is None else float(tau)
    q = torch.zeros((nd,) + shape, dtype=torch.float32, device=g.device)
    v = torch.empty_like(g)
    u = torch.empty_like(g)
    dmax = torch.empty(TV_WORK, dtype=torch.float32, device=g.device)
    L = _lib.lib()
    done = 0
    while done < max_num_iter:
        k = min(int(check_every), int(max_num_iter) - done)
        _lib.check(L.vir_tv_chambolle(nd, dims[0], dims[1], dims[2], _lib.ptr(g), _lib.ptr(q), _lib.ptr(v),
                                      None, float(weight), tau, k, _lib.ptr(dmax) if eps > 0 else None,
                                      stream()), "vir_tv_chambolle")
        done += k
        if eps > 0 and float(dmax[0]) <= eps * weight:
            break
    _lib.check(L.vir_tv_chambolle(nd, dims[0], dims[1], dims[2], _lib.ptr(g), _lib.ptr(q), _lib.ptr(v),
                                  _lib.ptr(u), float(weight), tau, 0, None, stream()), "vir_tv_chambolle")
    return u if torch_in else u.cpu().numpy()


def _shape3(shape):
    s = tuple(int(v) for v in shape)
    if len(s) == 2:
        return s + (1,)
    if len(s) == 3:
        return s
    raise ValueError("pdhg_tv needs a 2-D or 3-D image")


def pdhg_tv(P, g, iters=100, lam=1.0, theta=1.0, x0=None):
    """TV-regularised reconstruction by the Chambolle-Pock primal-dual
    iteration of tv.py:229-294 with diagonal step sizes: sigma_z = 1/2
    (1/grad_abs(1)), sigma_q = 1/(A 1) (0 where A 1 = 0), tau = 1/(A^T 1 +
    div_abs(1)) = 1/(A^T 1 + 2*ndim).  P: any projector with fp/bp on device
    tensors; g: the sinogram.  Returns u (device)."""
    from . import _lib
    from .device import as_device_f32, stream
    img_shape = tuple(getattr(P, "img_shape", None) or getattr(P, "nPixels"))
    sino_shape = tuple(getattr(P, "sino_shape", None) or getattr(P, "ray_shape"))
    nx, ny, nz = _shape3(img_shape if len(img_shape) != 3 or img_shape[2] != 1 else img_shape[:2])
    ndim = 3 if nz > 1 else 2
    g = as_device_f32(g).reshape(sino_shape)
    dev = g.device
    ones_i = torch.ones(img_shape, dtype=torch.float32, device=dev)
    ones_s = torch.ones(sino_shape, dtype=torch.float32, device=dev)
    A1 = P.fp(ones_i)
    sigma_q = torch.where(A1 != 0, 1.0 / A1, torch.zeros_like(A1))
    tau = 1.0 / (P.bp(ones_s) + 2.0 * ndim)
    u = (torch.zeros(img_shape, dtype=torch.float32, device=dev) if x0 is None
         else as_device_f32(x0).reshape(img_shape).clone())
    ubar = u.clone()
    z = torch.zeros((ndim,) + img_shape, dtype=torch.float32, device=dev)
    q = torch.zeros(sino_shape, dtype=torch.float32, device=dev)
    Au = torch.empty(sino_shape, dtype=torch.float32, device=dev)
    ATq = torch.empty(img_shape, dtype=torch.float32, device=dev)
    L = _lib.lib()
    for _ in range(int(iters)):
        st = stream()
        _lib.check(L.vir_pdhg_dual_tv(nx, ny, nz, _lib.ptr(ubar), _lib.ptr(z), 0.5, st),
                   "vir_pdhg_dual_tv")
        P.fp(ubar, out=Au)
        _lib.check(L.vir_pdhg_dual_data(q.numel(), _lib.ptr(Au), _lib.ptr(g), _lib.ptr(sigma_q),
                                        float(lam), _lib.ptr(q), st), "vir_pdhg_dual_data")
        P.bp(q, out=ATq)
        _lib.check(L.vir_pdhg_primal(nx, ny, nz, _lib.ptr(z), _lib.ptr(ATq), _lib.ptr(tau),
                                     float(theta), _lib.ptr(u), _lib.ptr(ubar), st),
                   "vir_pdhg_primal")
    return u
