"""ORACLE: distance-driven projector of the translational square geometry in
float64 -- test infrastructure only.

Restates dd.py:903-1313 (dd_fp_square / dd_fp_translational /
dd_bp_square / dd_bp_translational) as an explicit matrix W with
sinogram rows (src_p, src_z, v, u, side) and volume columns (x, y, z):

  * boundaries: pf.boundspace(n, d, s) = k*d + (s - d*n/2), sources
    pf.censpace(n, d, s) = k*d + (s - d*(n-1)/2) (pf.py:42-55);
  * slice io (orthogonal axis) magnified by M = DSD / (DSO - o_bnd[io]),
    the slice's lower boundary (dd.py:981), slices with M <= 0 skipped;
  * voxel p-interval M*p_bnd - (M*src_p - ssrc_p) against
    boundspace(nu, du, su + ssrc_p); z likewise against
    boundspace(nv, dv, sv + ssrc_z); weight overlap_u * overlap_v;
  * sides (dd.py:1003-1006): 0 img[io, ip, iz], 1 img[n-1-ip, io, iz],
    2 img[n-1-io, n-1-ip, iz], 3 img[ip, n-1-io, iz].

FP = d_pix^2/(du*dv) * W img (dd.py:930); BP = W^T sino (dd.py:1135).
Pinned in tests/test_ddsq_oracle.py by the closed-form footprint of a single
voxel (total weight (M d)^2 when the footprint is on the detector), the
90-degree rotation relation between sides, and the adjoint identity of the
reference's BP loop restated separately.
"""
import numpy as np


def boundspace(n, d, s=0.0):
    return np.arange(n + 1) * d + (s - d * n * 0.5)


def censpace(n, d, s=0.0):
    return np.arange(n) * d + (s - d * (n - 1) * 0.5)


def _ovl(lo, hi, bnd):
    o = np.minimum(hi[..., None], bnd[1:]) - np.maximum(lo[..., None], bnd[:-1])
    return np.where(o > 0, o, 0.0)


def _side_index(side, n, io, ip):
    if side == 0:
        return io, ip
    if side == 1:
        return n - 1 - ip, io
    if side == 2:
        return n - 1 - io, n - 1 - ip
    return ip, n - 1 - io


def matrix(img_shape, nu, nv, nsrc_p, nsrc_z, DSO, DSD, du=1.0, dv=1.0, dsrc_p=1.0, dsrc_z=1.0,
           su=0.0, sv=0.0, ssrc_p=0.0, ssrc_z=0.0, d_pix=1.0):
    n, n2, nz = img_shape
    assert n == n2
    pb, zb = boundspace(n, d_pix), boundspace(nz, d_pix)
    ub, vb = boundspace(nu, du, su + ssrc_p), boundspace(nv, dv, sv + ssrc_z)
    sp, sz = censpace(nsrc_p, dsrc_p, ssrc_p), censpace(nsrc_z, dsrc_z, ssrc_z)
    W = np.zeros((nsrc_p, nsrc_z, nv, nu, 4, n, n, nz))
    for io in range(n):
        M = DSD / (DSO - pb[io])
        if M <= 0:
            continue
        for isp in range(nsrc_p):
            ps = M * sp[isp] - ssrc_p
            wu = _ovl(M * pb[:-1] - ps, M * pb[1:] - ps, ub)          # (n_p, nu)
            for isz in range(nsrc_z):
                zs = M * sz[isz] - ssrc_z
                wv = _ovl(M * zb[:-1] - zs, M * zb[1:] - zs, vb)      # (nz, nv)
                blk = np.einsum("pu,zv->vupz", wu, wv)               # (nv, nu, n_p, nz)
                for side in range(4):
                    for ip in range(n):
                        ix, iy = _side_index(side, n, io, ip)
                        W[isp, isz, :, :, side, ix, iy, :] += blk[:, :, ip, :]
    return W.reshape(nsrc_p * nsrc_z * nv * nu * 4, n * n * nz)


def fp(W, img, d_pix, du, dv, sino_shape):
    return (d_pix ** 2 / (du * dv) * (W @ np.asarray(img, np.float64).reshape(-1))).reshape(sino_shape)


def bp(W, sino, img_shape):
    return (W.T @ np.asarray(sino, np.float64).reshape(-1)).reshape(img_shape)


def fp_loop(img, nu, nv, nsrc_p, nsrc_z, DSO, DSD, du=1.0, dv=1.0, dsrc_p=1.0, dsrc_z=1.0,
            su=0.0, sv=0.0, ssrc_p=0.0, ssrc_z=0.0, d_pix=1.0):
    """dd_fp_square followed literally: the transposed copies imgX, imgY
    (dd.py:920-923) and the per-slice loops (dd.py:958-1098), float64."""
    img = np.asarray(img, np.float64)
    imgX = np.ascontiguousarray(img.transpose(1, 2, 0))
    imgY = np.ascontiguousarray(img.transpose(0, 2, 1))
    no, nz, nP = imgX.shape
    pb, ob, zb = boundspace(nP, d_pix), boundspace(no, d_pix), boundspace(nz, d_pix)
    ub, vb = boundspace(nu, du, su + ssrc_p), boundspace(nv, dv, sv + ssrc_z)
    sp, sz = censpace(nsrc_p, dsrc_p, ssrc_p), censpace(nsrc_z, dsrc_z, ssrc_z)
    sino = np.zeros((nsrc_p, nsrc_z, nv, nu, 4))
    for isp in range(nsrc_p):
        for io in range(no):
            M = DSD / (DSO - ob[io])
            if M <= 0:
                continue
            ps = M * sp[isp] - ssrc_p
            for iz in range(nz):
                cols = (imgY[io, iz], imgX[io, iz], imgY[no - 1 - io, iz], imgX[no - 1 - io, iz])
                tmp = np.zeros((nu, 4))
                for ip in range(nP):
                    pl, pr = M * pb[ip] - ps, M * pb[ip + 1] - ps
                    vals = (cols[0][ip], cols[1][nP - 1 - ip], cols[2][nP - 1 - ip], cols[3][ip])
                    for iu in range(nu):
                        o = min(pr, ub[iu + 1]) - max(pl, ub[iu])
                        if o > 0:
                            for f in range(4):
                                tmp[iu, f] += vals[f] * o
                for isz in range(nsrc_z):
                    zs = M * sz[isz] - ssrc_z
                    vl, vr = M * zb[iz] - zs, M * zb[iz + 1] - zs
                    for iv in range(nv):
                        o = min(vr, vb[iv + 1]) - max(vl, vb[iv])
                        if o > 0:
                            sino[isp, isz, iv] += tmp * o
    return sino * d_pix ** 2 / (du * dv)


def bp_loop(sino, img_shape, DSO, DSD, du=1.0, dv=1.0, dsrc_p=1.0, dsrc_z=1.0, su=0.0, sv=0.0,
            ssrc_p=0.0, ssrc_z=0.0, d_pix=1.0):
    """dd_bp_square followed literally: accumulation into imgX, imgY and the
    final transposes (dd.py:1104-1313), float64."""
    sino = np.asarray(sino, np.float64)
    nx, ny, nz = img_shape
    imgX = np.zeros((ny, nz, nx))
    imgY = np.zeros((nx, nz, ny))
    no, _, nP = imgX.shape
    nsrc_p, nsrc_z, nv, nu, _ = sino.shape
    pb, ob, zb = boundspace(nP, d_pix), boundspace(no, d_pix), boundspace(nz, d_pix)
    ub, vb = boundspace(nu, du, su + ssrc_p), boundspace(nv, dv, sv + ssrc_z)
    sp, sz = censpace(nsrc_p, dsrc_p, ssrc_p), censpace(nsrc_z, dsrc_z, ssrc_z)
    for io in range(no):
        M = DSD / (DSO - ob[io])
        if M <= 0:
            continue
        for isp in range(nsrc_p):
            ps = M * sp[isp] - ssrc_p
            for iz in range(nz):
                tmp = np.zeros((nu, 4))
                for isz in range(nsrc_z):
                    zs = M * sz[isz] - ssrc_z
                    vl, vr = M * zb[iz] - zs, M * zb[iz + 1] - zs
                    for iv in range(nv):
                        o = min(vr, vb[iv + 1]) - max(vl, vb[iv])
                        if o > 0:
                            tmp += sino[isp, isz, iv] * o
                for ip in range(nP):
                    pl, pr = M * pb[ip] - ps, M * pb[ip + 1] - ps
                    acc = np.zeros(4)
                    for iu in range(nu):
                        o = min(pr, ub[iu + 1]) - max(pl, ub[iu])
                        if o > 0:
                            acc += tmp[iu] * o
                    imgX[io, iz, ip] += acc[0]
                    imgY[io, iz, nP - 1 - ip] += acc[1]
                    imgX[no - 1 - io, iz, nP - 1 - ip] += acc[2]
                    imgY[no - 1 - io, iz, ip] += acc[3]
    return imgX.transpose(0, 2, 1) + imgY.transpose(2, 0, 1)
