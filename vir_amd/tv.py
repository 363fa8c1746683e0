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
