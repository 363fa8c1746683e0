"""Synthetic workloads: seeded ellipse / ellipsoid phantoms and the
geometries of the benchmark configurations.

Phantoms follow the recipe of the benchmark definitions: centres
(FOV/2)*U(-0.45, 0.45), semi-axes (FOV/2)*U(0.05, 0.45), random orientation,
densities U(-0.2, 1.0), summed where they overlap.  Generated on the GPU
with torch (data preparation only; not part of any timed region).
"""
import numpy as np
import torch

from .proj_geom import geom_circular


def ellipse_params(n_ell, seed, ndim):
    rng = np.random.default_rng(seed)
    cen = rng.uniform(-0.45, 0.45, (n_ell, ndim))
    ax = rng.uniform(0.05, 0.45, (n_ell, ndim))
    rho = rng.uniform(-0.2, 1.0, n_ell)
    if ndim == 2:
        ang = rng.uniform(0, np.pi, n_ell)
        rot = np.stack([np.stack([np.cos(ang), -np.sin(ang)], -1),
                        np.stack([np.sin(ang), np.cos(ang)], -1)], -2)
    else:
        a, b, c = (rng.uniform(0, 2 * np.pi, n_ell), rng.uniform(0, np.pi, n_ell),
                   rng.uniform(0, 2 * np.pi, n_ell))

        def rz(t):
            z, o = np.zeros_like(t), np.ones_like(t)
            return np.stack([np.stack([np.cos(t), -np.sin(t), z], -1),
                             np.stack([np.sin(t), np.cos(t), z], -1),
                             np.stack([z, z, o], -1)], -2)

        def ry(t):
            z, o = np.zeros_like(t), np.ones_like(t)
            return np.stack([np.stack([np.cos(t), z, np.sin(t)], -1),
                             np.stack([z, o, z], -1),
                             np.stack([-np.sin(t), z, np.cos(t)], -1)], -2)
        rot = rz(a) @ ry(b) @ rz(c)
    return cen, ax, rot, rho


def phantom(shape, n_ell, seed, supersample=1, device="cuda"):
    """Ellipse (2-D, shape (n, n) or (n, n, 1)) or ellipsoid (3-D) phantom,
    float32, in FOV-normalised coordinates [-1, 1] per axis."""
    shape = tuple(int(s) for s in shape)
    flat2d = len(shape) == 2 or (len(shape) == 3 and shape[2] == 1)
    ndim = 2 if flat2d else 3
    cen, ax, rot, rho = ellipse_params(n_ell, seed, ndim)
    dims = shape[:ndim]
    out = torch.zeros(dims, dtype=torch.float64, device=device)
    ss = supersample
    for s_off in np.ndindex(*([ss] * ndim)):
        axes = []
        for k, n in enumerate(dims):
            c = (torch.arange(n, dtype=torch.float64, device=device) + (s_off[k] + 0.5) / ss) / n * 2 - 1
            axes.append(c)
        grids = torch.meshgrid(*axes, indexing="ij")
        pts = torch.stack(grids, -1)
        for e in range(n_ell):
            R = torch.as_tensor(rot[e], device=device)
            q = (pts - torch.as_tensor(cen[e], device=device)) @ R
            inside = ((q / torch.as_tensor(ax[e], device=device)) ** 2).sum(-1) <= 1.0
            out += inside.to(torch.float64) * float(rho[e])
    out /= ss ** ndim
    return out.to(torch.float32).reshape(shape).contiguous()


def phantom_device(shape, n_ell, seed, device="cuda", stream=None):
    """The phantom of `phantom` (same seeded ellipses, no supersampling)
    evaluated by the vir_phantom kernel in float32: one launch, no
    temporaries, enqueued on `stream` (default: the current stream)."""
    from . import _lib
    shape = tuple(int(s) for s in shape)
    flat2d = len(shape) == 2 or (len(shape) == 3 and shape[2] == 1)
    ndim = 2 if flat2d else 3
    cen, ax, rot, rho = ellipse_params(n_ell, seed, ndim)
    prm = np.zeros((n_ell, 16), np.float32)
    prm[:, :ndim] = cen
    prm[:, 3:3 + ndim] = ax
    if ndim == 2:
        prm[:, 5] = 1.0
    R = np.zeros((n_ell, 3, 3))
    R[:, :ndim, :ndim] = rot
    if ndim == 2:
        R[:, 2, 2] = 1.0
    prm[:, 6:15] = R.reshape(n_ell, 9)
    prm[:, 15] = rho
    p = torch.from_numpy(prm).to(device, non_blocking=False)
    nx, ny = shape[0], shape[1]
    nz = 1 if ndim == 2 else shape[2]
    out = torch.empty(shape, dtype=torch.float32, device=device)
    st = _lib.stream_ptr(stream)
    _lib.check(_lib.lib().vir_phantom(nx, ny, nz, n_ell, _lib.ptr(p), _lib.ptr(out), st), "vir_phantom")
    if stream is not None:
        p.record_stream(stream)   # keep the parameters alive until the kernel ran
    return out


def cone256_geometry(views=None):
    """3-D cone beam: 360 views over 2*pi, source-isocentre 500 mm,
    source-detector 1000 mm, 512 x 384 detector at 1.0 mm, 256^3 volume of
    0.5 mm voxels (pitch / magnification 2).  Returns float32 (src, trg) of
    shape (nviews, 512, 384, 3) and the grid."""
    v = np.arange(360) * 2 * np.pi / 360
    if views is not None:
        v = v[np.asarray(views)]
    du = (np.arange(512) - 255.5) * 1.0
    dv = (np.arange(384) - 191.5) * 1.0
    s, t = geom_circular(du, v, "cone", src_iso=500.0, det_iso=500.0, DetsZ=dv)
    return s.astype(np.float32), t.astype(np.float32), dict(nPixels=(256, 256, 256), dPixels=0.5)


def par128_geometry():
    """2-D parallel beam: 180 views over [0, pi), 185 bins at unit spacing,
    128 x 128 (x 1) grid of unit pixels."""
    v = np.arange(180) * np.pi / 180
    d = (np.arange(185) - 92.0) * 1.0
    s, t = geom_circular(d, v, "par")
    return s.astype(np.float32), t.astype(np.float32), dict(nPixels=(128, 128, 1), dPixels=1.0)
