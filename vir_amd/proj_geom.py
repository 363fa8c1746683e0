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
    `tol` ray-index units (every ray is checked).  Views are fitted in
    groups of about 2^18 rays (arrays over the view axis)."""
    nviews, nu, nv = trg.shape[:3]
    step = max(1, (1 << 18) // max(1, nu * nv))
    return np.concatenate([_fit_views(src[a:a + step], trg[a:a + step], tol, a)
                           for a in range(0, nviews, step)], axis=0)


def _fit_views(src, trg, tol, first):
    """view_projections of a group of views; `first` is the index of the
    group's first view (for the error messages)."""
    src = np.asarray(src, np.float64)
    trg = np.asarray(trg, np.float64)
    nviews, nu, nv = trg.shape[:3]

    def dot(x, y):   # per-view dot products of (nviews, 3) rows
        return (x[:, 0] * y[:, 0] + x[:, 1] * y[:, 1]) + x[:, 2] * y[:, 2]

    def unit(x):
        return x / np.linalg.norm(x, axis=1, keepdims=True)

    T0 = trg[:, 0, 0]
    S0 = src[:, 0, 0]
    D0 = T0 - S0
    atol = 1e-6 * (1 + np.abs(S0).max(axis=1))
    point = np.all(np.abs(src - S0[:, None, None, :]) <= atol[:, None, None, None], axis=(1, 2, 3))
    if nu > 1:
        Eu = (trg[:, nu - 1, 0] - T0) / (nu - 1)
    else:
        # single detector column: any direction orthogonal to the ray
        Eu = np.cross(D0, [0.0, 0.0, 1.0])
        alt = np.cross(D0, [0.0, 1.0, 0.0])
        Eu = unit(np.where((np.linalg.norm(Eu, axis=1) < 1e-12)[:, None], alt, Eu))
    flat_v = nv == 1
    Ev = unit(np.cross(D0, Eu)) if flat_v else (trg[:, 0, nv - 1] - T0) / (nv - 1)
    N = np.cross(Eu, Ev)
    A = np.cross(Ev, N)   # u = (Q - T0).A / (Eu.A)
    B = np.cross(N, Eu)   # v = (Q - T0).B / (Ev.B)
    euA, evB = dot(Eu, A), dot(Ev, B)
    M = np.zeros((nviews, 3, 4), np.float64)
    # point sources: u*w with w = N.(X - S): Q - T0 = (S - T0) + t (X - S), t = nd / N.(X-S)
    P_ = np.nonzero(point)[0]
    if P_.size:
        Sp, T0p, Np, Ap, Bp = S0[P_], T0[P_], N[P_], A[P_], B[P_]
        nd = dot(Np, T0p - Sp)
        cu = dot(Sp - T0p, Ap)[:, None] * Np + nd[:, None] * Ap
        cv = dot(Sp - T0p, Bp)[:, None] * Np + nd[:, None] * Bp
        sgn = np.where(nd > 0, 1.0, -1.0)[:, None]
        M[P_, 0] = sgn * np.concatenate([cu, -dot(cu, Sp)[:, None]], axis=1) / euA[P_, None]
        M[P_, 1] = sgn * np.concatenate([cv, -dot(cv, Sp)[:, None]], axis=1) / evB[P_, None]
        M[P_, 2] = sgn * np.concatenate([Np, -dot(Np, Sp)[:, None]], axis=1)
    # parallel bundles: Q = X + t Dir, t = N.(T0 - X)/nd; u = (Q - T0).A/(Eu.A)
    Q_ = np.nonzero(~point)[0]
    if Q_.size:
        Dir, T0q, Nq, Aq, Bq = D0[Q_], T0[Q_], N[Q_], A[Q_], B[Q_]
        dirs = trg[Q_] - src[Q_]
        dtol = 1e-5 * (1 + np.abs(Dir).max(axis=1))
        same = np.all(np.abs(dirs - Dir[:, None, None, :]) <= dtol[:, None, None, None], axis=(1, 2, 3))
        if not same.all():
            raise ValueError(f"view {first + int(Q_[np.argmin(same)])}: rays neither share a source nor a "
                             f"direction")
        nd = dot(Nq, Dir)
        if (np.abs(nd) < 1e-12).any():
            raise ValueError(f"view {first + int(Q_[np.argmax(np.abs(nd) < 1e-12)])}: rays parallel to the "
                             f"detector plane")
        gu = Aq - (dot(Dir, Aq) / nd)[:, None] * Nq
        gv = Bq - (dot(Dir, Bq) / nd)[:, None] * Nq
        M[Q_, 0] = np.concatenate([gu, -dot(gu, T0q)[:, None]], axis=1) / euA[Q_, None]
        M[Q_, 1] = np.concatenate([gv, -dot(gv, T0q)[:, None]], axis=1) / evB[Q_, None]
        M[Q_, 2] = np.array([0.0, 0.0, 0.0, 1.0])
    if flat_v:
        M[:, 1] = 0.0
    # every ray must project onto its own index (point sources: the target;
    # parallel bundles: the midpoint)
    X = np.where(point[:, None, None, None], trg, 0.5 * (src + trg)).reshape(nviews, nu * nv, 3)
    h = np.matmul(X, M[:, :, :3].transpose(0, 2, 1)) + M[:, None, :, 3]
    with np.errstate(divide="ignore", invalid="ignore"):
        pu = (h[..., 0] / h[..., 2]).reshape(nviews, nu, nv)
        pv = (h[..., 1] / h[..., 2]).reshape(nviews, nu, nv)
    err = np.abs(pu - np.arange(nu, dtype=np.float64)[None, :, None]).max(axis=(1, 2))
    if not flat_v:
        err = np.maximum(err, np.abs(pv - np.arange(nv, dtype=np.float64)[None, None, :]).max(axis=(1, 2)))
    bad = ~(np.isfinite(err) & (err <= tol))
    if bad.any():
        a = int(np.nonzero(bad)[0][0])
        raise ValueError(f"view {first + a}: detector is not an affine ray grid (err {err[a]:.3g})")
    return M.reshape(nviews, 12).astype(np.float32)
