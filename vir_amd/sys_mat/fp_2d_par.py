"""The 2-D parallel-beam prototypes of sys_mat/fp_2d_par.py on the GPU.

  siddons_fp_2d(img, Angs, n_dets, d_det=1.0, d_pix=1.0)      fp_2d_par.py:11-64
      exact path lengths of rays with direction (cos, sin) through the bin
      centres d_det (k - n/2 + 1/2) along (-sin, cos) -- the Amanatides-Woo
      parallel projector's geometry and output (vir_rd_fp, sys_mat/rd.py)
  separable_footprint_fp_2d(img, Angs, n_dets, d_det=1.0, d_pix=1.0)
                                                              fp_2d_par.py:67-125
      each pixel's corners projected on -sin x + cos y, the rectangle's
      overlap with every bin divided by d_det -- pd.py's parallel footprint
      with amplitude 1 and no d_pix factor (vir_sf_fp with
      VIR_SF_UNIT_AMPLITUDE, output scale 1)

Both return float32 (na, n_dets); numpy in -> numpy out, CUDA tensors in ->
CUDA tensors out.  The reference loops in float64 and accumulates in
float32; results agree to float32 rounding.
"""
import numpy as np
import torch

from .pd import SFProjector
from .rd import RDProjector


def _wrap(x, torch_in):
    return x if torch_in else x.cpu().numpy()


def siddons_fp_2d(img, Angs, n_dets, d_det=1.0, d_pix=1.0):
    torch_in = isinstance(img, torch.Tensor)
    P = RDProjector(tuple(img.shape), np.asarray(Angs).reshape(-1), n_dets, "par", du=d_det, d_pix=d_pix)
    return _wrap(P.fp(img), torch_in)


def separable_footprint_fp_2d(img, Angs, n_dets, d_det=1.0, d_pix=1.0):
    torch_in = isinstance(img, torch.Tensor)
    P = SFProjector(tuple(img.shape), np.asarray(Angs).reshape(-1), n_dets, "par", du=d_det, d_pix=d_pix,
                    unit_amplitude=True)
    from ..device import as_device_f32, out_tensor
    x = as_device_f32(img)
    out = out_tensor(None, P.sino_shape, x.device)
    P._call("vir_sf_fp", x, out, 1.0, False)
    return _wrap(out, torch_in)
