"""Ray geometry: source/target endpoints of circular and square trajectories.

Same call surface and the same float64 arithmetic as the reference's
proj_geom.py (geom_square :12-57, geom_circular :61-148), so the float32
endpoints that siddons() consumes are bit-identical.  Also derives, per
view, the 3x4 projection matrix the brick-gather backprojector uses to find
a voxel brick's detector footprint.
"""
import numpy as np


def square_steps(distY, Angs):
    """Per-angle target offsets (D, E) = (cos(a) distY 5, sin(a) distY 5) of
    geom_square, evaluated with numpy in the input dtypes as the reference
    does (proj_geom.py:51-54): float32 angles give float32 offsets."""
    return np.cos(Angs) * distY * 5, np.sin(Angs) * distY * 5


def geom_square(BinsX, distY, Angs):
    """Source/target coordinates of a four-sided translational fan geometry
    (proj_geom.py:12-57), float64 arrays (nBins, nAngs, 4, 3), z = 0.

    Side k of the square holds the detectors at BinsX b along one edge, at
    distance distY from the centre; each looks across the square at angle a:

        side   source        target
        0      ( b, -d)      ( b + D, -d + E)
        1      (-d,  b)      (-d + E,  b + D)
        2      (-b,  d)      (-b - D,  d - E)
        3      ( d, -b)      ( d - E, -b - D)

    with d = distY and (D, E) = square_steps(distY, a); every coordinate is
    the source value plus or minus one offset in float64, so the endpoints
    are the reference's bits (tests/golden/geom_square_golden.npz)."""
    BinsX = np.asarray(BinsX)
    Angs = np.asarray(Angs)
    nBins, nAngs = BinsX.size, Angs.size
    D, E = square_steps(distY, Angs)
    D = np.asarray(D, np.float64)[None, :]
    E = np.asarray(E, np.float64)[None, :]
    b = BinsX.reshape(-1).astype(np.float64)[:, None]
    nb = (-BinsX).reshape(-1).astype(np.float64)[:, None]
    d = np.float64(distY)
    md = np.float64(-distY)
    src = np.zeros((nBins, nAngs, 4, 3))
    trg = np.zeros((nBins, nAngs, 4, 3))
    sides = ((b, md, +1, +1), (md, b, +1, +1), (nb, d, -1, -1), (d, nb, -1, -1))
    for k, (sx, sy, sgn_d, sgn_e) in enumerate(sides):
        src[:, :, k, 0] = sx
        src[:, :, k, 1] = sy
        # side 0/2 move x by D and y by E; side 1/3 move x by E and y by D
        dx, dy = (D, E) if k % 2 == 0 else (E, D)
        trg[:, :, k, 0] = sx + dx if sgn_d > 0 else sx - dx
        trg[:, :, k, 1] = sy + dy if sgn_e > 0 else sy - dy
    return src, trg


def geom_circular(DetsY, Views, geom="par", src_iso=None, det_iso=None, DetsZ=None):
    """Source/target coordinates of a circular orbit (proj_geom.py:61-148).

    geom: "par" (rays from -src_iso to det_iso along x, detectors at DetsY
    [and DetsZ]), "fan" (point source at (-src_iso, 0)), "cone" (point source,
    2-D detector DetsY x DetsZ).  Each view rotates the in-plane coordinates
    by the row-vector product [x y] @ [[cos, sin], [-sin, cos]].
    Returns float64 (src, trg) of shape Views.shape + DetsY.shape
    [+ DetsZ.shape] + (3,)."""
    DetsY = np.asarray(DetsY)
    Views = np.asarray(Views)
    if DetsZ is not None:
        DetsZ = np.asarray(DetsZ)
    if geom == "fan":
        if src_iso is None:
            raise ValueError("src_iso must be provided with geom=fan")
        if det_iso is None:
            raise ValueError("det_iso must be provided with geom=fan")
        if DetsZ is None:
            z_src, z_trg, y_trg = 0, 0, DetsY
        else:
            z_src, z_trg, y_trg = DetsZ, DetsZ, DetsY[:, np.newaxis]
        y_src = 0.0
    elif geom == "cone":
        if src_iso is None:
            raise ValueError("src_iso must be provided with geom=cone")
        if det_iso is None:
            raise ValueError("det_iso must be provided with geom=cone")
        if DetsZ is None:
            raise ValueError("DetsZ must be provided with geom=cone")
        z_src, z_trg = 0, DetsZ
        y_src, y_trg = 0.0, DetsY[:, np.newaxis]
    elif geom == "par":
        if DetsZ is None:
            z_src, z_trg, y_src, y_trg = 0, 0, DetsY, DetsY
        else:
            z_src, z_trg = DetsZ, DetsZ
            y_src, y_trg = DetsY[:, np.newaxis], DetsY[:, np.newaxis]
        if det_iso is None:
            det_iso = np.max(DetsY) * 100
        if src_iso is None:
            src_iso = np.max(DetsY) * 100
    else:
        raise ValueError("Cannot process geom: ", geom)

    if DetsZ is None:
        Bins = Views.shape + DetsY.shape
    else:
        Bins = Views.shape + DetsY.shape + DetsZ.shape

    src = np.stack([np.broadcast_to(-src_iso, Bins), np.broadcast_to(y_src, Bins),
                    np.broadcast_to(z_src, Bins)], axis=-1)
    trg = np.stack([np.broadcast_to(det_iso, Bins), np.broadcast_to(y_trg, Bins),
                    np.broadcast_to(z_trg, Bins)], axis=-1)
    for i, view in np.ndenumerate(Views):
        r_mat = np.array([[np.cos(view), np.sin(view)], [-np.sin(view), np.cos(view)]])
        src[i][..., :2] = np.matmul(np.stack([src[i][..., 0], src[i][..., 1]], axis=-1), r_mat)
        trg[i][..., :2] = np.matmul(np.stack([trg[i][..., 0], trg[i][..., 1]], axis=-1), r_mat)
    return (src, trg)


def check_in_front(mats, nPixels, dPixels, origin, margin=2.0):
    """True when every corner of the volume grown by `margin` voxels lies
    strictly in front of every view's source (P2.X~ > 0), i.e. no source
    sits inside or next to the volume."""
    n = np.asarray(nPixels, np.float64)
    d = np.broadcast_to(np.asarray(dPixels, np.float64), (3,))
    o = np.broadcast_to(np.asarray(origin, np.float64), (3,))
    half = n * d / 2 + margin * d
    corners = np.array([[o[0] + sx * half[0], o[1] + sy * half[1], o[2] + sz * half[2], 1.0]
                        for sx in (-1, 1) for sy in (-1, 1) for sz in (-1, 1)])
    w = np.asarray(mats, np.float64).reshape(-1, 3, 4)[:, 2, :] @ corners.T
    return bool(np.all(w > 1e-6 * np.abs(w).max()))


def check_projections(src, trg, mats, tol=0.05):
    """Raise ValueError unless each view's 3x4 matrix maps its own rays onto
    their ray indices: the four corner rays and the centre ray of every view
    (src, trg: (nviews, nu, nv, 3) numpy arrays or device tensors; only
    those 5 rays per view are read) must project to (iu, iv) within `tol`
    index units.  The midpoint of a ray is on the ray for a point source and
    for a parallel bundle alike, so it is the point projected."""
    nviews, nu, nv = (int(s) for s in trg.shape[:3])
    pick = sorted({(0, 0), (0, nv - 1), (nu - 1, 0), (nu - 1, nv - 1), (nu // 2, nv // 2)})
    iu_idx = np.array([p[0] for p in pick])
    iv_idx = np.array([p[1] for p in pick])
    if hasattr(trg, "is_cuda"):
        import torch
        ti = torch.as_tensor(iu_idx, device=trg.device)
        tv = torch.as_tensor(iv_idx, device=trg.device)
        s = src[:, ti, tv].double().cpu().numpy()
        t = trg[:, ti, tv].double().cpu().numpy()
    else:
        s = np.asarray(src, np.float64)[:, iu_idx, iv_idx]
        t = np.asarray(trg, np.float64)[:, iu_idx, iv_idx]
    M = np.asarray(mats, np.float64).reshape(nviews, 3, 4)
    X = np.concatenate([0.5 * (s + t), np.ones(s.shape[:2] + (1,))], axis=2)
    h = np.einsum("aij,akj->aki", M, X)
    with np.errstate(divide="ignore", invalid="ignore"):
        pu = h[..., 0] / h[..., 2]
        pv = h[..., 1] / h[..., 2]
    err = np.abs(pu - iu_idx[None, :])
    if nv > 1:
        err = np.maximum(err, np.abs(pv - iv_idx[None, :]))
    bad = ~(np.isfinite(err) & (err <= tol) & (h[..., 2] > 0))
    if bad.any():
        a = int(np.nonzero(bad.any(axis=1))[0][0])
        raise ValueError(f"view {a}: the projection matrix does not map the view's rays onto their "
                         f"detector indices (error {np.nanmax(err[a]):.3g} > {tol})")


def view_projections(src, trg, tol=0.05):
    """Per-view 3x4 projection matrices for a structured ray set.

    src, trg: float32 arrays (nviews, nu, nv, 3).  Each view must be either a
    point source (every ray of the view starts at the same point) or a
    parallel bundle (one direction), with targets on an affine detector
    grid trg(iu, iv) = T0 + iu*Eu + iv*Ev.  Returns float32 (nviews, 12) such
    that for a world point X: iu = (P0.X~)/(P2.X~), iv = (P1.X~)/(P2.X~), the
    fractional ray index of the ray through X, with P2.X~ > 0 in front of the
    source.  Raises ValueError when a view does not fit the model within
    `tol` ray-index units (every ray is checked)."""
    src = np.asarray(src, np.float64)
    trg = np.asarray(trg, np.float64)
    nviews, nu, nv = trg.shape[:3]
    mats = np.zeros((nviews, 12), np.float64)
    iu = np.arange(nu, dtype=np.float64)[:, None]
    iv = np.arange(nv, dtype=np.float64)[None, :]
    for a in range(nviews):
        T = trg[a]
        S = src[a]
        T0 = T[0, 0]
        Eu = (T[nu - 1, 0] - T0) / (nu - 1) if nu > 1 else None
        Ev = (T[0, nv - 1] - T0) / (nv - 1) if nv > 1 else None
        D0 = T0 - S[0, 0]
        point = np.allclose(S, S[0, 0], rtol=0, atol=1e-6 * (1 + np.abs(S[0, 0]).max()))
        if Eu is None:
            # single detector column: any direction orthogonal to the ray
            Eu = np.cross(D0, [0.0, 0.0, 1.0])
            if np.linalg.norm(Eu) < 1e-12:
                Eu = np.cross(D0, [0.0, 1.0, 0.0])
            Eu = Eu / np.linalg.norm(Eu)
        if Ev is None:
            Ev = np.cross(D0, Eu)
            Ev = Ev / np.linalg.norm(Ev)
            flat_v = True
        else:
            flat_v = False
        N = np.cross(Eu, Ev)
        A = np.cross(Ev, N)   # u = (Q - T0).A / (Eu.A)
        B = np.cross(N, Eu)   # v = (Q - T0).B / (Ev.B)
        if point:
            Sp = S[0, 0]
            nd = N @ (T0 - Sp)
            # u*w with w = N.(X - S): Q - T0 = (S - T0) + t (X - S), t = nd / N.(X-S)
            cu = ((Sp - T0) @ A) * N + nd * A
            cv = ((Sp - T0) @ B) * N + nd * B
            sgn = 1.0 if nd > 0 else -1.0
            ru = sgn * np.concatenate([cu, [-cu @ Sp]]) / (Eu @ A)
            rv = sgn * np.concatenate([cv, [-cv @ Sp]]) / (Ev @ B)
            rw = sgn * np.concatenate([N, [-N @ Sp]])
        else:
            Dir = D0
            dirs = T - S
            if not np.allclose(dirs, Dir, rtol=0, atol=1e-5 * (1 + np.abs(Dir).max())):
                raise ValueError(f"view {a}: rays neither share a source nor a direction")
            nd = N @ Dir
            if abs(nd) < 1e-12:
                raise ValueError(f"view {a}: rays parallel to the detector plane")
            # Q = X + t Dir, t = N.(T0 - X)/nd; u = (Q - T0).A/(Eu.A)
            # (Q - T0).A = (X - T0).A + (N.(T0 - X)/nd) Dir.A
            gu = A - (Dir @ A) / nd * N
            gv = B - (Dir @ B) / nd * N
            ru = np.concatenate([gu, [-gu @ T0]]) / (Eu @ A)
            rv = np.concatenate([gv, [-gv @ T0]]) / (Ev @ B)
            rw = np.array([0.0, 0.0, 0.0, 1.0])
        if flat_v:
            rv = np.zeros(4)
        M = np.stack([ru, rv, rw])
        # every ray must project onto its own index
        X = np.concatenate([T.reshape(-1, 3), np.ones((nu * nv, 1))], axis=1)
        if not point:
            X = np.concatenate([(0.5 * (S + T)).reshape(-1, 3), np.ones((nu * nv, 1))], axis=1)
        h = X @ M.T
        pu = (h[:, 0] / h[:, 2]).reshape(nu, nv)
        pv = (h[:, 1] / h[:, 2]).reshape(nu, nv)
        err = np.max(np.abs(pu - iu))
        if not flat_v:
            err = max(err, np.max(np.abs(pv - iv)))
        if not np.isfinite(err) or err > tol:
            raise ValueError(f"view {a}: detector is not an affine ray grid (err {err:.3g})")
        mats[a] = M.reshape(-1)
    return mats.astype(np.float32)
