"""Total-variation denoising (Chambolle's projection algorithm).

The reference regularises SIRT iterates with
skimage.restoration.denoise_tv_chambolle (tests/Test_SIRT/sirt_test.py:13,38),
which is not available here.  This restates the published algorithm
(A. Chambolle, "An algorithm for total variation minimization and
applications", JMIV 20, 2004): dual ascent on p with step tau = 1/(2*ndim),
u = f - div(p), stopping when the relative energy change drops below eps.
Runs on the device with elementwise tensor operations.
"""
import torch


def _div(p):
    # negative adjoint of the forward difference (zero flux at the ends)
    nd = p.shape[0]
    d = -p.sum(0)
    for ax in range(nd):
        sl_d = [slice(None)] * nd
        sl_p = [slice(None)] * nd
        sl_d[ax] = slice(1, None)
        sl_p[ax] = slice(0, -1)
        d[tuple(sl_d)] += p[ax][tuple(sl_p)]
    return d


def denoise_tv_chambolle(image, weight=0.1, eps=2.0e-4, max_num_iter=200):
    """TV denoising of `image` (any dimension): argmin_u ||u - f||^2/2 +
    weight * TV(u), iterated until |dE| < eps * E_0."""
    f = image if isinstance(image, torch.Tensor) else torch.as_tensor(image)
    f = f.to(torch.float32)
    nd = f.dim()
    p = torch.zeros((nd,) + tuple(f.shape), dtype=f.dtype, device=f.device)
    g = torch.zeros_like(p)
    tau = 1.0 / (2.0 * nd)
    e_init = e_prev = None
    out = f
    for i in range(max_num_iter):
        if i > 0:
            d = _div(p)
            out = f + d
            energy = (d * d).sum()
        else:
            out = f
            energy = torch.zeros((), dtype=f.dtype, device=f.device)
        g.zero_()
        for ax in range(nd):
            sl = [slice(None)] * nd
            sl[ax] = slice(0, -1)
            g[ax][tuple(sl)] = torch.diff(out, dim=ax)
        norm = torch.sqrt((g * g).sum(0))
        energy = energy + weight * norm.sum()
        norm = 1.0 + norm * (tau / weight)
        p = (p - tau * g) / norm
        e = float(energy) / f.numel()
        if i == 0:
            e_init = e_prev = e
        else:
            if abs(e_prev - e) < eps * e_init:
                break
            e_prev = e
    return out


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
