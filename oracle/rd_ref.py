"""ORACLE: ray-driven projectors of sys_mat/rd.py in float64 -- test
infrastructure only (pure-Python loops, small cases).

Restates, per ray, the reference's geometry and loops:
  * rays: bin centres du*(iu - (nu/2 - 0.5) + su) (rd.py:483), cone rows
    dv*(iv - (nu/2 - 0.5) + sv) (rd.py:691, centred with nu as the
    reference does); parallel origin u*(-sin, cos), direction (cos, sin)
    (rd.py:489-499); fan/cone from the source DSO*(cos, sin) to the flat
    detector -(DSD-DSO)*(cos, sin) + u*(-sin, cos) [+ v z] (rd.py:580-605,
    697-733), unit direction;
  * box intersection _intersect_bounding (rd.py:51-64), entry voxel
    _calc_ir0 (rd.py:66-73), stepping _fp_step_init (rd.py:76-108);
  * AW traversal: FP value * length (rd.py:111-148, 178-213), BP splat of
    the same lengths (rd.py:151-175, 215-247); cone FP scaled by
    DSD/sqrt(DSD^2+u^2+v^2) (rd.py:784-785);
  * Joseph: samples every 0.5 from t_entry while t <= t_exit; 2-D FP clamps
    and interpolates bilinearly (rd.py:250-289), 3-D FP truncates and keeps
    inside cells (rd.py:292-353); BP splats with floor indices and
    per-corner bounds (rd.py:356-445);
  * drivers: FP unscaled; BP / na / d_pix^2 * du (2-D), / na * du * dv (cone).

`clamp_parallel=True` follows the reference for rays parallel to an axis
lying outside the image (their entry index is clamped onto the border row);
False drops them (the GPU projectors' behaviour).

Pinned in tests/test_rd_oracle.py: AW forward = the Siddon oracle's path
lengths on the same rays (a different algorithm); Joseph forward of constant
and linear images in closed form; BP = the transpose of FP where the
reference pairs are adjoint.
"""
import math

import numpy as np

EPS = 1e-6
STEP = 0.5


def _bounds(n, d):
    return -d * n / 2.0, d * n / 2.0, -d * (n / 2.0 - 0.5)


def rays(mode, ang, nu, du=1.0, su=0.0, DSO=None, DSD=None, nv=1, dv=1.0, sv=0.0):
    """List of (ia, iu, iv, origin(3), unit direction(3), u, v)."""
    out = []
    ucen = nu / 2.0 - 0.5
    for ia, a in enumerate(np.asarray(ang, np.float64)):
        c, s = math.cos(a), math.sin(a)
        for iu in range(nu):
            u = du * (iu - ucen + su)
            for iv in range(nv if mode == "cone" else 1):
                v = dv * (iv - ucen + sv) if mode == "cone" else 0.0
                if mode == "par":
                    o, h = (u * -s, u * c, 0.0), (c, s, 0.0)
                else:
                    o = (DSO * c, DSO * s, 0.0)
                    det = (-(DSD - DSO) * c - u * s, -(DSD - DSO) * s + u * c, v)
                    vec = [det[k] - o[k] for k in range(3)]
                    m = math.sqrt(sum(x * x for x in vec))
                    h = tuple(x / m for x in vec)
                out.append((ia, iu, iv, o, h, u, v))
    return out


def _entry(o, h, shape, d, clamp_parallel):
    na = len(shape)
    t0, t1 = -math.inf, math.inf
    for a in range(na):
        lo, hi, _ = _bounds(shape[a], d)
        if abs(h[a]) > EPS:
            ta, tb = (lo - o[a]) / h[a], (hi - o[a]) / h[a]
            t0, t1 = max(t0, min(ta, tb)), min(t1, max(ta, tb))
        elif not clamp_parallel and not (lo <= o[a] < hi):
            return None
    if t1 <= t0:
        return None
    return t0, t1


def _aw_segments(o, h, shape, d, clamp_parallel):
    """[(voxel index tuple, length)] of the AW traversal."""
    e = _entry(o, h, shape, d, clamp_parallel)
    if e is None:
        return []
    t0, t1 = e
    na = len(shape)
    idx, dirs, step, tn = [], [], [], []
    for a in range(na):
        lo, hi, _ = _bounds(shape[a], d)
        r = o[a] + t0 * h[a]
        r = min(max(r, lo + EPS), hi - EPS)
        i = int(math.floor((r - lo) / d))
        idx.append(i)
        if h[a] > 0:
            dirs.append(1)
            r_next = lo + (i + 1) * d
        else:
            dirs.append(-1)
            r_next = lo + i * d
        if abs(h[a]) > EPS:
            step.append(d / abs(h[a]))
            tn.append(max((r_next - o[a]) / h[a], t0))
        else:
            step.append(math.inf)
            tn.append(math.inf)
    segs = []
    t = t0
    while t < t1:
        if any(idx[a] < 0 or idx[a] >= shape[a] for a in range(na)):
            break
        a = min(range(na), key=lambda k: (tn[k], k))
        te = min(tn[a], t1)
        segs.append((tuple(idx), te - t))
        t = te
        tn[a] += step[a]
        idx[a] += dirs[a]
    return segs


def _joseph_samples(o, h, shape, d, clamp_parallel):
    e = _entry(o, h, shape, d, clamp_parallel)
    if e is None:
        return []
    t0, t1 = e
    na = len(shape)
    c0 = [_bounds(shape[a], d)[2] for a in range(na)]
    out = []
    k = 0
    while True:
        t = t0 + k * STEP
        if not t <= t1:
            break
        out.append(tuple((o[a] + t * h[a] - c0[a]) / d for a in range(na)))
        k += 1
    return out


def fp(img, mode, ang, nu, du=1.0, su=0.0, d_pix=1.0, DSO=None, DSD=None, nv=1, dv=1.0,
       sv=0.0, joseph=False, clamp_parallel=False):
    img = np.asarray(img, np.float64)
    shape = img.shape
    na = len(shape)
    R = rays(mode, ang, nu, du, su, DSO, DSD, nv, dv, sv)
    out = np.zeros((len(ang), nu, nv) if mode == "cone" else (len(ang), nu))
    for ia, iu, iv, o, h, u, v in R:
        acc = 0.0
        if not joseph:
            for ix, ln in _aw_segments(o, h, shape, d_pix, clamp_parallel):
                acc += img[ix] * ln
            if mode == "cone":
                acc *= DSD / math.sqrt(DSD ** 2 + u ** 2 + v ** 2)
        else:
            for f in _joseph_samples(o, h, shape, d_pix, clamp_parallel):
                if na == 2:
                    fx = min(max(f[0], 0.0), shape[0] - 2.0)
                    fy = min(max(f[1], 0.0), shape[1] - 2.0)
                    ix, iy = int(fx), int(fy)
                    dx, dy = fx - ix, fy - iy
                    val = (img[ix, iy] * (1 - dx) * (1 - dy) + img[ix + 1, iy] * dx * (1 - dy) +
                           img[ix, iy + 1] * (1 - dx) * dy + img[ix + 1, iy + 1] * dx * dy)
                    acc += val * STEP
                else:
                    ii = [int(x) for x in f]      # truncation toward zero
                    if all(0 <= ii[a] < shape[a] - 1 for a in range(3)):
                        dx, dy, dz = (f[a] - ii[a] for a in range(3))
                        ix, iy, iz = ii
                        c000, c100 = img[ix, iy, iz], img[ix + 1, iy, iz]
                        c010, c110 = img[ix, iy + 1, iz], img[ix + 1, iy + 1, iz]
                        c001, c101 = img[ix, iy, iz + 1], img[ix + 1, iy, iz + 1]
                        c011, c111 = img[ix, iy + 1, iz + 1], img[ix + 1, iy + 1, iz + 1]
                        c00 = c000 + dx * (c100 - c000)
                        c10 = c010 + dx * (c110 - c010)
                        c01 = c001 + dx * (c101 - c001)
                        c11 = c011 + dx * (c111 - c011)
                        c0 = c00 + dy * (c10 - c00)
                        c1 = c01 + dy * (c11 - c01)
                        acc += (c0 + dz * (c1 - c0)) * STEP
        if mode == "cone":
            out[ia, iu, iv] = acc
        else:
            out[ia, iu] = acc
    return out


def bp(sino, mode, ang, img_shape, du=1.0, su=0.0, d_pix=1.0, DSO=None, DSD=None, dv=1.0,
       sv=0.0, joseph=False, clamp_parallel=False, reference_scale=True):
    """The reference's backprojection (splat) in float64, with its driver
    scaling when reference_scale."""
    sino = np.asarray(sino, np.float64)
    shape = tuple(img_shape)
    na_ax = len(shape)
    nu = sino.shape[1]
    nv = sino.shape[2] if mode == "cone" else 1
    img = np.zeros(shape)
    for ia, iu, iv, o, h, u, v in rays(mode, ang, nu, du, su, DSO, DSD, nv, dv, sv):
        y = sino[ia, iu, iv] if mode == "cone" else sino[ia, iu]
        if y == 0:
            continue
        if not joseph:
            for ix, ln in _aw_segments(o, h, shape, d_pix, clamp_parallel):
                img[ix] += y * ln
        else:
            for f in _joseph_samples(o, h, shape, d_pix, clamp_parallel):
                ii = [int(math.floor(x)) for x in f]
                dd = [f[a] - ii[a] for a in range(na_ax)]
                for corner in range(1 << na_ax):
                    idx, w = [], 1.0
                    for a in range(na_ax):
                        if corner >> a & 1:
                            idx.append(ii[a] + 1)
                            w *= dd[a]
                        else:
                            idx.append(ii[a])
                            w *= 1 - dd[a]
                    if all(0 <= idx[a] < shape[a] for a in range(na_ax)):
                        img[tuple(idx)] += y * STEP * w
    if not reference_scale:
        return img
    if mode == "cone":
        return img / len(ang) * du * dv
    return img / len(ang) / d_pix / d_pix * du


def ray_endpoints(mode, ang, nu, du=1.0, su=0.0, DSO=None, DSD=None, nv=1, dv=1.0, sv=0.0,
                  reach=1e3):
    """(src, trg) float64 (nrays, 3) of the rays, for cross-checks with the
    Siddon path-length oracle."""
    R = rays(mode, ang, nu, du, su, DSO, DSD, nv, dv, sv)
    src = np.array([[o[k] - (reach if mode == "par" else 0.0) * h[k] for k in range(3)]
                    for _, _, _, o, h, _, _ in R])
    trg = np.array([[o[k] + reach * h[k] for k in range(3)] for _, _, _, o, h, _, _ in R])
    return src, trg
