"""SIRT driver of the translational square geometry (sys_mat/sirt_dd.py).

`sirt(sino, W_sino, W_img, gem_prms, img_prms, src_prms, det_prms,
srt_prms)` keeps the reference's call surface (sirt_dd.py:37-76): x0 =
1e-5 everywhere, iterations 1 .. iters-1, each

    r = (b - A x) / W_sino   where W_sino != 0      (fast_divide1, :23-26)
    u = A^T r / W_img        where W_img  != 0      (fast_divide2, :30-33)
    x = x + gamma * u

with A = dd_fp_square and A^T = dd_bp_square, run on the device with the
residual and update fused into one kernel each and one iteration captured
into a HIP graph.  numpy in -> numpy out.
"""
import numpy as np
import torch

from ..sirt import SIRT
from .dd import SquareDDProjector


def sirt(sino, W_sino, W_img, gem_prms, img_prms, src_prms, det_prms, srt_prms, graph=True):
    DSO, DSD = gem_prms
    nx, ny, nz, d_pix = img_prms
    nsrc_p, nsrc_z, dsrc_p, dsrc_z, ssrc_p, ssrc_z = src_prms
    nu, nv, du, dv, su, sv = det_prms
    gamma, iters = srt_prms
    torch_in = isinstance(sino, torch.Tensor)
    P = SquareDDProjector((nx, ny, nz), nu, nv, nsrc_p, nsrc_z, DSO, DSD, du=du, dv=dv,
                          dsrc_p=dsrc_p, dsrc_z=dsrc_z, su=su, sv=sv, ssrc_p=ssrc_p,
                          ssrc_z=ssrc_z, d_pix=d_pix)
    x0 = np.full((nx, ny, nz), np.float32(1e-5), np.float32)
    S = SIRT(P, sino, x0=x0, W_sino=W_sino, W_img=W_img, gamma=gamma, graph=graph)
    S.run(max(int(iters) - 1, 0))
    return S.x if torch_in else S.x.cpu().numpy()
