"""Colour phases of the brick-gather Siddon backprojector never race.

The backprojector (vir_amd/csrc/siddon.hip, k_siddon_bp_wave) accumulates a
16x8x8 sub-brick in on-chip memory without atomics.  Per (sub-brick, view)
it takes the candidate rays (iu, iv) of a detector rectangle and walks them
in colour phases -- rays with iu = cu (mod su), iv = cv (mod sv) -- whose
lanes run together.  That is race-free only if no two distinct rays of one
phase have an entry in the same voxel.

This file restates the kernel's rectangle and strides (view_foot) in numpy
float32, in the kernel's operation order, and checks on the C oracle's trace
(the entries the backprojector visits) that
  * every ray with an entry in a sub-brick lies in its candidate rectangle;
  * the distinct rays of one view that meet one voxel have distinct colours;
on two views of the 256^3 cone-beam workload (512 x 384 detector) and on a
cone beam whose source sits at the projector's in-front limit.  A stride
taken from a fraction of the bound (a plausible mistake) must collide.
"""
import numpy as np
import pytest

from oracle import siddon as osd
from vir_amd.proj_geom import check_in_front, geom_circular, view_projections
from vir_amd.workloads import cone256_geometry

F = np.float32
SB = (16, 8, 8)  # BpShape3 sub-brick


def _project(P, x, y, z):
    u = ((P[0] * x + P[1] * y) + P[2] * z) + P[3]
    v = ((P[4] * x + P[5] * y) + P[6] * z) + P[7]
    w = ((P[8] * x + P[9] * y) + P[10] * z) + P[11]
    return u, v, w


def view_foot(P, planes, d, b0, b1, nu, nv, stride_scale=F(1.05)):
    """Candidate rectangle and colour strides of sub-bricks [b0, b1) (arrays
    (B, 3)) for one view's float32 matrix P (12,)."""
    P = np.asarray(P, F)
    kG = F(0.05)
    lo = [planes[a][b0[:, a]] - kG * d[a] for a in range(3)]
    hi = [planes[a][b1[:, a]] + kG * d[a] for a in range(3)]
    gu, gv, gw = [], [], []
    for c in range(8):
        u, v, w = _project(P, hi[0] if c & 4 else lo[0], hi[1] if c & 2 else lo[1],
                           hi[2] if c & 1 else lo[2])
        gu.append(u / w)
        gv.append(v / w)
        gw.append(w)
    gu, gv, gw = np.stack(gu), np.stack(gv), np.stack(gw)
    assert (gw > 0).all()
    umin, umax, vmin, vmax, wmin = gu.min(0), gu.max(0), gv.min(0), gv.max(0), gw.min(0)
    ku = np.zeros_like(umin)
    kv = np.zeros_like(umin)
    for i in range(3):
        c = P[8 + i]
        ku = ku + d[i] * np.maximum(np.abs(P[i] - umin * c), np.abs(P[i] - umax * c))
        kv = kv + d[i] * np.maximum(np.abs(P[4 + i] - vmin * c), np.abs(P[4 + i] - vmax * c))
    fu, fv = ku / wmin, kv / wmin
    rm, sm = F(0.075), F(0.15)
    iu0 = np.maximum(np.ceil(umin - rm), 0).astype(np.int64)
    iu1 = np.minimum(np.floor(umax + rm), nu - 1).astype(np.int64)
    iv0 = np.maximum(np.ceil(vmin - rm), 0).astype(np.int64)
    iv1 = np.minimum(np.floor(vmax + rm), nv - 1).astype(np.int64)
    su = np.floor(fu * stride_scale + sm).astype(np.int64) + 1
    sv = np.ones_like(su) if nv == 1 else np.floor(fv * stride_scale + sm).astype(np.int64) + 1
    su = np.minimum(su, np.maximum(iu1 - iu0 + 1, 1))
    sv = np.minimum(sv, np.maximum(iv1 - iv0 + 1, 1))
    return iu0, iu1, iv0, iv1, su, sv


def colour_check(s, t, nP, dP, stride_scale=F(1.05)):
    """(entries outside their sub-brick's rectangle, same-colour distinct ray
    pairs sharing a voxel) summed over the views of (s, t)."""
    nviews, nu, nv = t.shape[:3]
    mats = view_projections(s, t)
    n, d, o = osd.grid_params(nP, dP, 0.0)
    assert check_in_front(mats, n, d, o)   # a geometry the projector accepts
    planes = [osd.planes(int(n[a]), float(d[a]), float(o[a])) for a in range(3)]
    nb = [(int(n[a]) + SB[a] - 1) // SB[a] for a in range(3)]
    bidx = np.stack(np.meshgrid(*[np.arange(k) for k in nb], indexing="ij"), -1).reshape(-1, 3)
    b0 = bidx * np.array(SB)
    b1 = np.minimum(b0 + np.array(SB), n)
    outside = collisions = 0
    for a in range(nviews):
        _, x, y, z, _, off, _, _ = osd.trace(s[a], t[a], nP, dP)
        ray = np.repeat(np.arange(nu * nv), np.diff(off))
        vox = (x.astype(np.int64) * n[1] + y) * n[2] + z
        # one ray can repeat a voxel (near-tie segments); one lane walks it
        pair = np.unique(vox * (nu * nv) + ray)
        vox, ray = pair // (nu * nv), pair % (nu * nv)
        iu, iv = ray // nv, ray % nv
        xx, yy, zz = vox // (n[1] * n[2]), (vox // n[2]) % n[1], vox % n[2]
        bi = ((xx // SB[0]) * nb[1] + yy // SB[1]) * nb[2] + zz // SB[2]
        iu0, iu1, iv0, iv1, su, sv = view_foot(mats[a], planes, d.astype(F), b0, b1, nu, nv,
                                               stride_scale)
        outside += int((~((iu >= iu0[bi]) & (iu <= iu1[bi]) & (iv >= iv0[bi]) &
                          (iv <= iv1[bi]))).sum())
        key = vox * int((su * sv).max()) + (iu % su[bi]) * sv[bi] + iv % sv[bi]
        key.sort()
        collisions += int((key[1:] == key[:-1]).sum())
    return outside, collisions


def test_cone256_workload_views():
    s, t, g = cone256_geometry([0, 45, 101])
    assert colour_check(s, t, g["nPixels"], g["dPixels"]) == (0, 0)


def _near_source():
    v = np.arange(8) * 2 * np.pi / 8 + 0.3
    du = (np.arange(96) - 47.5) * 1.0
    dv = (np.arange(48) - 23.5) * 1.0
    # 32 x 32 x 16 unit voxels; the volume grown by 2 voxels reaches 25.5
    # from the axis at 45 degrees: the source is just outside it
    s, t = geom_circular(du, v, "cone", src_iso=26.0, det_iso=30.0, DetsZ=dv)
    return s.astype(np.float32), t.astype(np.float32), (32, 32, 16)


def test_source_at_the_in_front_limit():
    s, t, nP = _near_source()
    assert colour_check(s, t, nP, 1.0) == (0, 0)


@pytest.mark.parametrize("scale", [0.5, 0.7])
def test_a_too_small_stride_collides(scale):
    s, t, g = cone256_geometry([45])
    _, coll = colour_check(s, t, g["nPixels"], g["dPixels"], stride_scale=F(scale))
    assert coll > 0
