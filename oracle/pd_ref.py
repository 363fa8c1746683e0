"""ORACLE: pixel-driven separable-footprint projectors in float64 -- test
infrastructure only (tests, smoke and the bench's CPU leg).

Restates sys_mat/pd.py:
  * footprint of pixel (ix, iy) in view a: the four corners projected onto
    the detector, rectangle [min, max]                   (pd.py:196-211, 296-306, 420-423)
      parallel  p = -sin*x + cos*y                          (pd.py:188-189)
      fan       u = DSD*(p)/(DSO - o), o = cos*x + sin*y    (pd.py:40-42, 282-285)
      cone      u = fm * p_corner, fm = DSD/(DSO - o_centre) (pd.py:415-423)
  * amplitude  parallel/cone 1/(|sin|+|cos|)               (pd.py:181, 392)
               fan 1/(|cos(a+g)|+|sin(a+g)|), g = atan(u_c/DSD) (pd.py:291-298)
  * cone v footprint  (fm/cos g) * [z_l, z_r]              (pd.py:416, 436-437)
  * weight per bin  overlap/du (* overlap_v/dv)            (pd.py:55-65, 84-102)
  * forward output * d_pix; backprojection / na / d_pix * du (* dv)
                                                           (pd.py:231-234, 347-350, 482-485)

`matrix(...)` builds the dense system matrix W (weights without d_pix) so
that FP = d_pix * W img and BP = du(*dv)/(na*d_pix) * W^T sino.
`fan_fp_loop` follows the reference's fan loop literally, including its
handling of zero pixels: the skip path (pd.py:321-325) leaves oy_bnd_l
unchanged, so the next nonzero pixel's centre, and with it the amplitude,
is taken from a stale boundary.  `stale_oy=False` gives the intended value.

Pinned in tests/test_pd_oracle.py by closed forms: the parallel views at
0 and 90 degrees are row/column sums, mass conservation sum_u sino*du =
d_pix^2 * sum(img), the fan footprint of the centre pixel, the fan and cone
models' parallel limit (DSO, DSD -> inf), and the cone v-magnification of an
off-axis voxel.
"""
import numpy as np


def _bnd(n, d):
    return d * (np.arange(n + 1, dtype=np.float64) - n / 2.0)


def det_bnd(n, d, s):
    return d * (np.arange(n + 1, dtype=np.float64) - n / 2.0 + s)


def footprints(mode, nx, ny, d_pix, ang, DSO=None, DSD=None):
    """(u0, u1, amp, cm) arrays of shape (nx, ny) for one view angle.
    mode: "par", "fan" or "cone"; cm is the cone v-magnification (1 else)."""
    ang = float(ang)
    c, s = np.cos(ang), np.sin(ang)
    xb, yb = _bnd(nx, d_pix), _bnd(ny, d_pix)
    xl, xr = xb[:-1, None], xb[1:, None]
    yl, yr = yb[None, :-1], yb[None, 1:]
    pxl, pxr, pyl, pyr = -s * xl, -s * xr, c * yl, c * yr
    if mode == "par":
        P = [pyl + pxl, pyl + pxr, pyr + pxl, pyr + pxr]
        u0 = np.minimum.reduce(np.broadcast_arrays(*P))
        u1 = np.maximum.reduce(np.broadcast_arrays(*P))
        amp = np.full((nx, ny), 1.0 / (abs(s) + abs(c)))
        return u0, u1, amp, np.ones((nx, ny))
    oxl, oxr, oyl, oyr = c * xl, c * xr, s * yl, s * yr
    pxc, oxc = (pxl + pxr) / 2, (oxl + oxr) / 2
    pyc, oyc = (pyl + pyr) / 2, (oyl + oyr) / 2
    uc = DSD * (pxc + pyc) / (DSO - (oxc + oyc))
    gam = np.arctan(uc / DSD)
    if mode == "fan":
        P = [DSD * (px + py) / (DSO - (ox + oy))
             for (px, ox) in ((pxl, oxl), (pxr, oxr)) for (py, oy) in ((pyl, oyl), (pyr, oyr))]
        u0 = np.minimum.reduce(np.broadcast_arrays(*P))
        u1 = np.maximum.reduce(np.broadcast_arrays(*P))
        amp = 1.0 / (np.abs(np.cos(ang + gam)) + np.abs(np.sin(ang + gam)))
        return u0, u1, amp, np.ones((nx, ny))
    if mode == "cone":
        fm = DSD / (DSO - (oxc + oyc))
        P = [fm * (pyl + pxl), fm * (pyl + pxr), fm * (pyr + pxl), fm * (pyr + pxr)]
        u0 = np.minimum.reduce(np.broadcast_arrays(*P))
        u1 = np.maximum.reduce(np.broadcast_arrays(*P))
        amp = np.full((nx, ny), 1.0 / (abs(s) + abs(c)))
        return u0, u1, amp, fm / np.cos(gam)
    raise ValueError(mode)


def _overlap(a0, a1, bnd):
    """(..., nb) positive overlaps of [a0, a1] with the bins of `bnd`."""
    o = np.minimum(a1[..., None], bnd[1:]) - np.maximum(a0[..., None], bnd[:-1])
    return np.where(o > 0, o, 0.0)


def matrix(mode, img_shape, ang_arr, nu, du=1.0, su=0.0, d_pix=1.0, DSO=None, DSD=None,
           nv=1, dv=1.0, sv=0.0):
    """Dense W: rows (view, u[, v]), columns (x, y[, z]); modes "par", "fan",
    "cone" and "par3d" (parallel per z-slice, rows of height dv along z)."""
    nx, ny = img_shape[:2]
    nz = img_shape[2] if mode in ("cone", "par3d") else 1
    ub = det_bnd(nu, du, su)
    blocks = []
    for a in np.asarray(ang_arr, np.float64).reshape(-1):
        fmode = "par" if mode == "par3d" else mode
        u0, u1, amp, cm = footprints(fmode, nx, ny, d_pix, a, DSO, DSD)
        wu = _overlap(u0, u1, ub) / du * amp[..., None]          # (nx, ny, nu)
        if mode in ("par", "fan"):
            blocks.append(wu.reshape(nx * ny, nu).T)
            continue
        vb = det_bnd(nv, dv, sv)
        zb = _bnd(nz, d_pix)
        v0 = cm[..., None] * zb[:-1]                             # (nx, ny, nz)
        v1 = cm[..., None] * zb[1:]
        wv = _overlap(v0, v1, vb) / dv                           # (nx, ny, nz, nv)
        blk = wu[:, :, None, :, None] * wv[:, :, :, None, :]     # (nx, ny, nz, nu, nv)
        blocks.append(blk.reshape(nx * ny * nz, nu * nv).T)
    return np.concatenate(blocks, 0)


def fp(W, img, d_pix):
    return d_pix * (W @ np.asarray(img, np.float64).reshape(-1))


def bp_reference(W, sino, na, d_pix, du, dv=None):
    out = W.T @ np.asarray(sino, np.float64).reshape(-1) / na / d_pix * du
    return out * dv if dv is not None else out


def fan_fp_loop(img, ang_arr, nu, DSO, DSD, du=1.0, su=0.0, d_pix=1.0, stale_oy=True):
    """The fan forward loop of pd.py:255-350 in float64, visiting pixels in
    the reference's order.  With stale_oy=True the skip path of a zero
    pixel keeps the previous oy boundary (pd.py:321-325)."""
    img = np.asarray(img, np.float64)
    nx, ny = img.shape
    xb, yb = _bnd(nx, d_pix), _bnd(ny, d_pix)
    ub = det_bnd(nu, du, su)
    sino = np.zeros((len(ang_arr), nu))
    for ia, a in enumerate(np.asarray(ang_arr, np.float64)):
        c, s = np.cos(a), np.sin(a)
        px, py, ox, oy = -s * xb, c * yb, c * xb, s * yb
        for ix in range(nx):
            pxc, oxc = (px[ix] + px[ix + 1]) / 2, (ox[ix] + ox[ix + 1]) / 2
            oy_l = oy[0]
            for iy in range(ny):
                if img[ix, iy] == 0:
                    if not stale_oy:
                        oy_l = oy[iy + 1]
                    continue
                pyc, oyc = (py[iy] + py[iy + 1]) / 2, (oy_l + oy[iy + 1]) / 2
                uc = DSD * (pxc + pyc) / (DSO - (oxc + oyc))
                g = np.arctan(uc / DSD)
                amp = 1.0 / (abs(np.cos(a + g)) + abs(np.sin(a + g)))
                P = [DSD * (px[i] + py[j]) / (DSO - (ox[i] + oy[j]))
                     for i in (ix, ix + 1) for j in (iy, iy + 1)]
                lo, hi = min(P), max(P)
                o = np.minimum(hi, ub[1:]) - np.maximum(lo, ub[:-1])
                sino[ia] += img[ix, iy] * amp * np.where(o > 0, o, 0.0) / du
                oy_l = oy[iy + 1]
    return sino * d_pix


def par3d_ops(img_shape, ang_arr, nu, nv, du=1.0, dv=1.0, su=0.0, sv=0.0, d_pix=1.0):
    """Forward and adjoint of matrix("par3d", ...) without the dense matrix,
    for workload-sized volumes: per view, the 2-D parallel footprint
    weights Wu (nu, nx*ny) as a sparse matrix and the slab/row overlaps
    Wv (nz, nv), so that  W img = Wu @ (img(nx*ny, nz) @ Wv)  per view.
    Returns (fp(img) -> (na, nu, nv) = d_pix W img, wt(sino) -> W^T sino).
    Pinned against matrix("par3d") in tests/test_pd_oracle.py."""
    import scipy.sparse as sp
    nx, ny, nz = img_shape
    ub = det_bnd(nu, du, su)
    vb = det_bnd(nv, dv, sv)
    zb = _bnd(nz, d_pix)
    Wv = _overlap(zb[:-1], zb[1:], vb) / dv                     # (nz, nv)
    mats = []
    for a in np.asarray(ang_arr, np.float64).reshape(-1):
        u0, u1, amp, _ = footprints("par", nx, ny, d_pix, a)
        u0, u1, amp = u0.reshape(-1), u1.reshape(-1), amp.reshape(-1)
        first = np.clip(np.searchsorted(ub, u0, side="right") - 1, 0, nu - 1)
        span = int(np.max(np.ceil((u1 - u0) / du))) + 2
        rows, cols, vals = [], [], []
        for k in range(span):
            iu = np.minimum(first + k, nu - 1)
            ov = np.minimum(u1, ub[iu + 1]) - np.maximum(u0, ub[iu])
            ok = (first + k < nu) & (ov > 0)
            rows.append(iu[ok])
            cols.append(np.nonzero(ok)[0])
            vals.append((ov / du * amp)[ok])
        mats.append(sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                                  shape=(nu, nx * ny)))

    def fp_(img):
        x = np.asarray(img, np.float64).reshape(nx * ny, nz) @ Wv
        return d_pix * np.stack([M @ x for M in mats])

    def wt(sino):
        y = np.asarray(sino, np.float64).reshape(len(mats), nu, nv)
        out = sum(M.T @ (y[i] @ Wv.T) for i, M in enumerate(mats))
        return np.asarray(out).reshape(nx, ny, nz)
    return fp_, wt
