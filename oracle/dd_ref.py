"""ORACLE: distance-driven projectors, float64 restatement -- test
infrastructure only.

Restates the reference's distance-driven merge sweeps (sys_mat/dd.py) in
float64 with pure-Python loops (small cases only):

* geometry per view: _dd_par_geom dd.py:471-523, _dd_fan_geom :526-585,
  _dd_fp_cone_geom :588-645 (driving axis |sin| >= |cos| -> x-driven,
  projected boundaries, monotone reordering, ray scales);
* sweeps: _dd_fp_par_sweep :81-126 / _dd_bp_par_sweep :129-161,
  _dd_fp_fan_sweep :164-221 / _dd_bp_fan_sweep :224-277,
  _dd_fp_cone_sweep :284-379 / _dd_bp_cone_sweep :383-468;
* drivers and output scaling: dd_p_par_2d :675-725, dd_p_fan_2d :748-803,
  dd_p_cone_3d :838-888 (FP * d_pix/du [* d_pix/dv]; BP / na / d_pix).

`cone_fp(..., stale_buffer=True)` reproduces the reference's cone forward
sweep literally, including its partial re-zeroing of the u buffer
(dd.py:317, `tmp_u[iu_min:iu_max] = 0.0`, while :374 adds the whole buffer),
which carries values from previous orthogonal lines into later ones; the
default (False) zeroes the whole buffer -- the projector the sweep intends.

Pins (tests/test_dd_oracle.py): mass conservation, single-pixel impulse
responses against closed-form rectangle overlaps, a brute-force per-(pixel,
bin) overlap integral, and the FP/BP scaling relation.
"""
import numpy as np


def _trig(ang):
    """cos/sin per view as the reference evaluates them: np.cos / np.sin of
    the float32 angle array (as_float32 dd.py:29-35, dd.py:702-703), widened
    to float64.  At a diagonal view |sin| == |cos| holds only in float32,
    and that tie fixes the driving axis."""
    a32 = np.asarray(ang, np.float32).reshape(-1)
    return np.cos(a32).astype(np.float64), np.sin(a32).astype(np.float64)


def _bnd(n, d):
    return d * (np.arange(n + 1, dtype=np.float64) - n / 2.0)


def _cen(b):
    return (b[:-1] + b[1:]) / 2.0


def _ubnd(nu, du, su):
    return du * (np.arange(nu + 1, dtype=np.float64) - nu / 2.0 + su)


def par_geom(nx, ny, d_pix, c, s):
    """_dd_par_geom: (x_driven, p_drv_bnd, p_orth, ray_scale, reversed)."""
    xb, yb = _bnd(nx, d_pix), _bnd(ny, d_pix)
    xc, yc = _cen(xb), _cen(yb)
    if abs(s) >= abs(c):
        xd, scale = True, 1.0 / abs(s)
        pdrv, porth = -s * xb, c * yc
    else:
        xd, scale = False, 1.0 / abs(c)
        pdrv, porth = c * yb, -s * xc
    rev = pdrv[1] < pdrv[0]
    if rev:
        pdrv = pdrv[::-1]
    return xd, pdrv, porth, scale, rev


def fan_geom(nx, ny, d_pix, c, s, DSO, DSD, cone=False):
    """_dd_fan_geom / _dd_fp_cone_geom."""
    xb, yb = _bnd(nx, d_pix), _bnd(ny, d_pix)
    xc, yc = _cen(xb), _cen(yb)
    if abs(s) >= abs(c):
        xd, rs = True, abs(s)
        pdrv, porth, odrv, oorth = -s * xb, c * yc, c * xb, s * yc
    else:
        xd, rs = False, abs(c)
        pdrv, porth, odrv, oorth = c * yb, -s * xc, s * yb, c * xc
    rev = pdrv[1] < pdrv[0]
    if rev:
        pdrv, odrv = pdrv[::-1], odrv[::-1]
    r = np.sqrt((DSO - oorth) ** 2 + porth ** 2)
    scl = r / ((DSO - oorth) * rs)
    return xd, pdrv, porth, odrv, oorth, scl, rev


def _img_trm(img, xd, rev):
    t = img if xd else img.T
    return t[::-1] if rev else t


def _sweep_par(vec_img, pb, ub, scl, out, bp):
    """Merge sweep of one orthogonal line (_dd_fp_par_sweep / _dd_bp_par_sweep)."""
    npix, nu = len(pb) - 1, len(ub) - 1
    ip = iu = 0
    while ip < npix and iu < nu:
        pl, pr = pb[ip], pb[ip + 1]
        ul, ur = ub[iu], ub[iu + 1]
        if not bp and pr < pl:
            pl, pr = pr, pl
        ol, orr = max(pl, ul), min(pr, ur)
        if orr > ol:
            if bp:
                out[ip] += vec_img[iu] * (orr - ol) * scl
            else:
                out[iu] += vec_img[ip] * (orr - ol) * scl
        if pr < ur:
            ip += 1
        else:
            iu += 1


def par_2d(img_or_sino, ang, nu_or_shape, du=1.0, su=0.0, d_pix=1.0, bp=False):
    """dd_fp_par_2d / dd_bp_par_2d in float64."""
    ang = np.asarray(ang, np.float64)
    ub = _ubnd(nu_or_shape if not bp else img_or_sino.shape[1], du, su)
    if bp:
        sino = np.asarray(img_or_sino, np.float64)
        nx, ny = nu_or_shape
        img_x = np.zeros((nx, ny))
        img_y = np.zeros((ny, nx))
    else:
        img = np.asarray(img_or_sino, np.float64)
        nx, ny = img.shape
        sino = np.zeros((ang.size, nu_or_shape))
    trig_c, trig_s = _trig(ang)
    for ia in range(ang.size):
        c, s = trig_c[ia], trig_s[ia]
        xd, pdrv, porth, scl, rev = par_geom(nx, ny, d_pix, c, s)
        if bp:
            tgt = img_x if xd else img_y
            view = tgt[::-1] if rev else tgt
            for io in range(view.shape[1]):
                col = np.zeros(view.shape[0])
                _sweep_par(sino[ia], pdrv + porth[io], ub, scl, col, True)
                view[:, io] += col
        else:
            trm = _img_trm(img, xd, rev)
            for io in range(trm.shape[1]):
                _sweep_par(trm[:, io], pdrv + porth[io], ub, scl, sino[ia], False)
    if bp:
        return (img_x + img_y.T) / ang.size / d_pix
    return sino * d_pix / du


def _fan_bounds(pdrv, porth, odrv, oorth, DSO, DSD):
    return DSD * (pdrv + porth) / (DSO - (odrv + oorth))


def fan_2d(img_or_sino, ang, nu_or_shape, DSO, DSD, du=1.0, su=0.0, d_pix=1.0, bp=False):
    """dd_fp_fan_2d / dd_bp_fan_2d (flat detector at DSD) in float64.
    The forward sweep swaps reversed intervals (dd.py:199-200), the back
    sweep does not (dd.py:256-263); for geometries where every projected
    pixel row is monotone the two are exact transposes up to scaling."""
    ang = np.asarray(ang, np.float64)
    if bp:
        sino = np.asarray(img_or_sino, np.float64)
        nx, ny = nu_or_shape
        ub = _ubnd(sino.shape[1], du, su)
        img_x, img_y = np.zeros((nx, ny)), np.zeros((ny, nx))
    else:
        img = np.asarray(img_or_sino, np.float64)
        nx, ny = img.shape
        ub = _ubnd(nu_or_shape, du, su)
        sino = np.zeros((ang.size, nu_or_shape))
    trig_c, trig_s = _trig(ang)
    for ia in range(ang.size):
        c, s = trig_c[ia], trig_s[ia]
        xd, pdrv, porth, odrv, oorth, scl, rev = fan_geom(nx, ny, d_pix, c, s, DSO, DSD)
        if bp:
            tgt = img_x if xd else img_y
            view = tgt[::-1] if rev else tgt
            for io in range(view.shape[1]):
                col = np.zeros(view.shape[0])
                pb = _fan_bounds(pdrv, porth[io], odrv, oorth[io], DSO, DSD)
                _sweep_par(sino[ia], pb, ub, scl[io], col, True)
                view[:, io] += col
        else:
            trm = _img_trm(img, xd, rev)
            for io in range(trm.shape[1]):
                pb = _fan_bounds(pdrv, porth[io], odrv, oorth[io], DSO, DSD)
                _sweep_par(trm[:, io], pb, ub, scl[io], sino[ia], False)
    if bp:
        return (img_x + img_y.T) / ang.size / d_pix
    return sino * d_pix / du


def cone_fp(img, ang, nu, nv, DSO, DSD, du=1.0, dv=1.0, su=0.0, sv=1.0, d_pix=1.0,
            stale_buffer=False):
    """dd_fp_cone_3d in float64 (dd.py:815-825, 838-888, sweep :284-379)."""
    img = np.asarray(img, np.float64)
    ang = np.asarray(ang, np.float64)
    nx, ny, nz = img.shape
    zb = _bnd(nz, d_pix)
    ub, vb = _ubnd(nu, du, su), _ubnd(nv, dv, sv)
    u0, v0 = ub[0], vb[0]
    sino = np.zeros((ang.size, nu, nv))
    trig_c, trig_s = _trig(ang)
    for ia in range(ang.size):
        c, s = trig_c[ia], trig_s[ia]
        xd, pdrv, porth, odrv, oorth, scl, rev = fan_geom(nx, ny, d_pix, c, s, DSO, DSD, True)
        vol = img if xd else img.transpose(1, 0, 2)
        if rev:
            vol = vol[::-1]
        npix, no = vol.shape[0], vol.shape[1]
        tmp = np.zeros(nu)
        for io in range(no):
            zmag = DSD / (DSO - oorth[io])
            pmag = DSD / (DSO - (odrv + oorth[io]))
            pb = pmag * (pdrv + porth[io])
            iu_min = max(int((pb[0] - u0) / du), 0)
            iu_max = min(int((pb[-1] - u0) / du) + 1, nu)
            for iz in range(nz):
                vl, vr = zmag * zb[iz], zmag * zb[iz + 1]
                iv_min = max(int((vl - v0) / dv), 0)
                iv_max = min(int((vr - v0) / dv) + 1, nv)
                if iv_max <= 0 or iv_min >= nv:
                    continue
                if stale_buffer:
                    tmp[iu_min:iu_max] = 0.0
                else:
                    tmp[:] = 0.0
                ip, iu = 0, iu_min
                if iu >= nu:
                    continue
                ul, ur = ub[iu], ub[iu + 1]
                pl, pr = pb[0], pb[1]
                val = vol[ip, io, iz] * scl[io]
                while True:
                    ol, orr = max(pl, ul), min(pr, ur)
                    if orr > ol:
                        tmp[iu] += val * (orr - ol)
                    if pr < ur:
                        ip += 1
                        if ip == npix:
                            break
                        pl, pr = pr, pb[ip + 1]
                        val = vol[ip, io, iz] * scl[io]
                    else:
                        iu += 1
                        if iu == nu:
                            break
                        ul, ur = ur, ub[iu + 1]
                for iv in range(iv_min, iv_max):
                    ovl, ovr = max(vl, vb[iv]), min(vr, vb[iv + 1])
                    if ovr <= ovl:
                        continue
                    sino[ia, :, iv] += tmp * (ovr - ovl)
    return sino * d_pix ** 2 / (du * dv)


def cone_bp(sino, ang, img_shape, DSO, DSD, du=1.0, dv=1.0, su=0.0, sv=0.0, d_pix=1.0):
    """dd_bp_cone_3d in float64 (dd.py:829-835, sweep :383-468)."""
    sino = np.asarray(sino, np.float64)
    ang = np.asarray(ang, np.float64)
    nx, ny, nz = img_shape
    na, nu, nv = sino.shape
    zb = _bnd(nz, d_pix)
    ub, vb = _ubnd(nu, du, su), _ubnd(nv, dv, sv)
    v0 = vb[0]
    img_x = np.zeros((nx, ny, nz))
    img_y = np.zeros((ny, nx, nz))
    trig_c, trig_s = _trig(ang)
    for ia in range(ang.size):
        c, s = trig_c[ia], trig_s[ia]
        xd, pdrv, porth, odrv, oorth, scl, rev = fan_geom(nx, ny, d_pix, c, s, DSO, DSD, True)
        tgt = img_x if xd else img_y
        vol = tgt[::-1] if rev else tgt
        npix, no = vol.shape[0], vol.shape[1]
        for io in range(no):
            zp = zb * DSD / (DSO - oorth[io])
            pb = DSD * (pdrv + porth[io]) / (DSO - (odrv + oorth[io]))
            for iz in range(nz):
                zl, zr = zp[iz], zp[iz + 1]
                iv0 = max(int(np.floor((zl - v0) / dv)), 0)
                iv1 = min(int(np.ceil((zr - v0) / dv)), nv)
                for iv in range(iv0, iv1):
                    ovl, ovr = max(zl, vb[iv]), min(zr, vb[iv + 1])
                    if ovr <= ovl:
                        continue
                    ov = ovr - ovl
                    ip = iu = 0
                    pl, pr = pb[0], pb[1]
                    ul, ur = ub[0], ub[1]
                    while True:
                        ol, orr = max(pl, ul), min(pr, ur)
                        if orr > ol:
                            vol[ip, io, iz] += sino[ia, iu, iv] * (orr - ol) * ov * scl[io]
                        if pr < ur:
                            ip += 1
                            if ip == npix:
                                break
                            pl, pr = pr, pb[ip + 1]
                        else:
                            iu += 1
                            if iu == nu:
                                break
                            ul, ur = ur, ub[iu + 1]
    return (img_x + img_y.transpose(1, 0, 2)) / ang.size / d_pix


def overlap_brute_2d(nx, ny, ang, nu, du, su, d_pix, geom="par", DSO=None, DSD=None, sub=64):
    """Independent check: system matrix of the DD model by direct evaluation
    of each pixel's projected driving-axis segment against each bin (no
    sweep), float64.  Returns A of shape (na*nu, nx*ny) with the FP scaling;
    cos/sin per view from _trig, as the sweeps."""
    ang = np.asarray(ang, np.float64)
    trig_c, trig_s = _trig(ang)
    A = np.zeros((ang.size, nu, nx, ny))
    ub = _ubnd(nu, du, su)
    xb, yb = _bnd(nx, d_pix), _bnd(ny, d_pix)
    xc, yc = _cen(xb), _cen(yb)
    for ia in range(ang.size):
        c, s = trig_c[ia], trig_s[ia]
        xdrv = abs(s) >= abs(c)
        for ix in range(nx):
            for iy in range(ny):
                if xdrv:
                    pts = [(xb[ix], yc[iy]), (xb[ix + 1], yc[iy])]
                else:
                    pts = [(xc[ix], yb[iy]), (xc[ix], yb[iy + 1])]
                if geom == "par":
                    us = [-s * x + c * y for x, y in pts]
                    scl = 1.0 / (abs(s) if xdrv else abs(c))
                else:
                    if geom == "arc":
                        # equiangular detector: arc length DSD * gamma
                        us = [DSD * np.arctan2(-s * x + c * y, DSO - (c * x + s * y))
                              for x, y in pts]
                    else:
                        us = [DSD * (-s * x + c * y) / (DSO - (c * x + s * y)) for x, y in pts]
                    po = c * yc[iy] if xdrv else -s * xc[ix]
                    oo = s * yc[iy] if xdrv else c * xc[ix]
                    rs = abs(s) if xdrv else abs(c)
                    scl = np.sqrt((DSO - oo) ** 2 + po ** 2) / ((DSO - oo) * rs)
                lo, hi = min(us), max(us)
                for iu in range(nu):
                    ov = min(hi, ub[iu + 1]) - max(lo, ub[iu])
                    if ov > 0:
                        A[ia, iu, ix, iy] = ov * scl * d_pix / du
    return A.reshape(ang.size * nu, nx * ny)


def overlap_sparse_2d(nx, ny, ang, nu, du, su, d_pix, geom="par", DSO=None, DSD=None):
    """overlap_brute_2d's system matrix (same formula, float64) as a
    scipy.sparse CSR matrix, vectorised over pixels: each pixel's projected
    driving-axis segment [lo, hi] against the few bins it meets.  For
    workload-sized grids (512^2 x 888 bins) where the dense loop is out of
    reach; pinned against overlap_brute_2d in tests/test_dd_oracle.py."""
    import scipy.sparse as sp
    ang = np.asarray(ang, np.float64)
    trig_c, trig_s = _trig(ang)
    ub = _ubnd(nu, du, su)
    xb, yb = _bnd(nx, d_pix), _bnd(ny, d_pix)
    xc, yc = _cen(xb), _cen(yb)
    rows, cols, vals = [], [], []
    pix = np.arange(nx * ny).reshape(nx, ny)
    for ia in range(ang.size):
        c, s = trig_c[ia], trig_s[ia]
        xdrv = abs(s) >= abs(c)
        if xdrv:
            pts = [(xb[:-1, None], yc[None, :]), (xb[1:, None], yc[None, :])]
        else:
            pts = [(xc[:, None], yb[None, :-1]), (xc[:, None], yb[None, 1:])]
        if geom == "par":
            us = [-s * x + c * y for x, y in pts]
            scl = np.full((nx, ny), 1.0 / (abs(s) if xdrv else abs(c)))
        else:
            if geom == "arc":
                us = [DSD * np.arctan2(-s * x + c * y, DSO - (c * x + s * y)) for x, y in pts]
            else:
                us = [DSD * (-s * x + c * y) / (DSO - (c * x + s * y)) for x, y in pts]
            if xdrv:
                po, oo = np.broadcast_to(c * yc[None, :], (nx, ny)), np.broadcast_to(s * yc[None, :], (nx, ny))
            else:
                po, oo = np.broadcast_to(-s * xc[:, None], (nx, ny)), np.broadcast_to(c * xc[:, None], (nx, ny))
            rs = abs(s) if xdrv else abs(c)
            scl = np.sqrt((DSO - oo) ** 2 + po ** 2) / ((DSO - oo) * rs)
        u0, u1 = np.broadcast_arrays(*us)
        lo, hi = np.minimum(u0, u1), np.maximum(u0, u1)
        first = np.clip(np.searchsorted(ub, lo, side="right") - 1, 0, nu - 1)
        span = int(np.max(np.ceil((hi - lo) / du))) + 2
        for k in range(span):
            iu = first + k
            ok = iu < nu
            iuc = np.minimum(iu, nu - 1)
            ov = np.minimum(hi, ub[iuc + 1]) - np.maximum(lo, ub[iuc])
            ok &= ov > 0
            rows.append((ia * nu + iuc)[ok])
            cols.append(pix[ok])
            vals.append((ov * scl * d_pix / du)[ok])
    return sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                         shape=(ang.size * nu, nx * ny))
