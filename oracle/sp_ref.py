"""ORACLE: the separable-footprint prototypes of sys_mat/sp.py in float64 --
test infrastructure only.

* fan_matrix: the system matrix of sf_fp_fan_2d (sp.py:12-87) without its
  d_pix factor, rows (view, bin), columns (x, y) C-order: pixel centre
  (x, y) = d (i - n/2 + 1/2), xp = c x + s y, yp = -s x + c y,
  u0 = DSD xp/(DSO - yp), h = DSD d/(DSO - yp), weight at bin centre u_b =
  du (b - nu/2 + 1/2): max(0, 1 - |u_b - u0|/h) (scipy.sparse CSR);
* sep_fp / sep_bp: the toy model of sp.py:135-211, proj[v, u] =
  sum vol[x,y,z] f(u/Dd - x/Ds) f(v/Dd - y/Ds) f(z/Ds), f(r) = max(0, 1-|r|),
  as dense weight matrices contracted with einsum, and its transpose.

Pinned in tests/test_sp_oracle.py: a triangle of integer half-width m bins
sums to m over the bins wherever its support lies on the detector (sum_k
max(0, 1 - |t - k|/m) = m), so every pixel on the central line at a
magnification that makes h an integer number of bins has column sum h/du;
in the parallel limit (DSO = DSD -> inf) the fan matrix becomes the parallel
triangle footprint max(0, 1 - |u_b - (c x + s y)|/d); a single voxel of the
toy model projects to the outer product of its two detector footprints times
f(z/Ds); the toy's forward and backward are transposes.
"""
import numpy as np


def _centres(n, d):
    return d * (np.arange(n, dtype=np.float64) - n / 2 + 0.5)


def fan_matrix(nx, ny, angs, nu, DSO, DSD, du=1.0, d_pix=1.0):
    import scipy.sparse as sp
    angs = np.asarray(angs, np.float64).reshape(-1)
    x = _centres(nx, d_pix)[:, None]
    y = _centres(ny, d_pix)[None, :]
    ub = _centres(nu, du)
    rows, cols, vals = [], [], []
    pix = np.arange(nx * ny)
    for a, th in enumerate(angs):
        c, s = np.cos(th), np.sin(th)
        xp = (c * x + s * y).reshape(-1)
        yp = (-s * x + c * y).reshape(-1)
        u0 = DSD * xp / (DSO - yp)
        h = DSD * d_pix / (DSO - yp)
        b0 = np.floor((u0 - h) / du + nu / 2).astype(np.int64)
        span = int(np.max(np.ceil(2 * h / du))) + 3
        for k in range(span):
            b = b0 + k
            ok = (b >= 0) & (b < nu)
            bc = np.clip(b, 0, nu - 1)
            w = np.maximum(0.0, 1.0 - np.abs(ub[bc] - u0) / h)
            ok &= w > 0
            rows.append(a * nu + bc[ok])
            cols.append(pix[ok])
            vals.append(w[ok])
    return sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                         shape=(angs.size * nu, nx * ny))


def _fp1(r):
    return np.maximum(1.0 - np.abs(r), 0.0)


def sep_weights(nx, ny, nz, nu, nv, dx, dy, dz, su, sv, Ds, Dd):
    """(Wx (nu, nx), Wy (nv, ny), wz (nz,))."""
    Wx = _fp1(_centres(nu, su)[:, None] / Dd - _centres(nx, dx)[None, :] / Ds)
    Wy = _fp1(_centres(nv, sv)[:, None] / Dd - _centres(ny, dy)[None, :] / Ds)
    wz = _fp1(_centres(nz, dz) / Ds)
    return Wx, Wy, wz


def sep_fp(vol, *geom):
    Wx, Wy, wz = sep_weights(*geom)
    return np.einsum("ux,vy,z,xyz->vu", Wx, Wy, wz, np.asarray(vol, np.float64))


def sep_bp(proj, *geom):
    Wx, Wy, wz = sep_weights(*geom)
    return np.einsum("ux,vy,z,vu->xyz", Wx, Wy, wz, np.asarray(proj, np.float64))
