"""Every projector family against an analytic sinogram, as the reference's
own checks compare dd / rd / pd against analytic circle sinograms
(sys_mat/test_fp_bp.py, sys_mat/test_scaling.py).

A disc (radius 30, centre (10, -5)) on a 128^2 grid of unit pixels,
area-weighted by 8x8 supersampling; 90 parallel views over [0, pi), 185
bins of width 1 centred on the axis.  Its sinogram is 2 sqrt(r^2 - s^2),
s the signed distance of the ray from the centre.  Siddon, distance-driven,
separable-footprint, Amanatides-Woo and Joseph forward projections must all
be within 2 % relative L2 of it (discretisation error of the pixel model)
and within 2 % of each other -- one scaling convention across families.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

N, R, CX, CY = 128, 30.0, 10.0, -5.0
ANG = np.arange(90) * np.pi / 90
U = np.arange(185) - 92.0


def _disc():
    ss = 8
    c = (np.arange(N * ss) + 0.5) / ss - N / 2
    X, Y = np.meshgrid(c, c, indexing="ij")
    inside = ((X - CX) ** 2 + (Y - CY) ** 2 <= R * R).astype(np.float64)
    return inside.reshape(N, ss, N, ss).mean(axis=(1, 3)).astype(np.float32)


def _analytic():
    s = U[None, :] - (-np.sin(ANG)[:, None] * CX + np.cos(ANG)[:, None] * CY)
    return 2 * np.sqrt(np.clip(R * R - s * s, 0, None))


def rel(a, b):
    a = np.asarray(a, np.float64).reshape(-1)
    b = np.asarray(b, np.float64).reshape(-1)
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def test_all_families_match_the_analytic_sinogram():
    from vir_amd.proj_geom import geom_circular
    from vir_amd.sys_mat import dd, pd, rd
    from vir_amd.sys_mat.siddon import SiddonProjector
    img = _disc()
    ref = _analytic()
    s, t = geom_circular(U, ANG, "par")
    sino = {
        "siddon": SiddonProjector(s, t, (N, N, 1), 1.0).fp(img.reshape(N, N, 1)).cpu().numpy(),
        "dd": dd.dd_fp_par_2d(img, ANG, U.size),
        "sf": pd.pd_fp_par_2d(img, ANG, U.size),
        "aw": rd.aw_fp_par_2d(img, ANG, U.size),
        "joseph": rd.aw_fp_par_2d(img, ANG, U.size, joseph=True),
    }
    for k, v in sino.items():
        assert v.shape == ref.shape, k
        assert rel(v, ref) < 0.02, (k, rel(v, ref))
    names = list(sino)
    for i in range(len(names)):
        for j in range(i + 1, len(names)):
            assert rel(sino[names[i]], sino[names[j]]) < 0.02, (names[i], names[j])
