"""Circular-orbit ray sets generated on the device.

`circular_rays` returns the float32 (src, trg) that
siddons(*geom_circular(...)) would see (proj_geom.py:61-148, cast to
float32 as sys_mat/siddon.py:684-690 does), computed by a HIP kernel from
the view angles and detector coordinates instead of being built on the host
and copied over PCIe (24 bytes per ray).  `circular_projections` gives the
per-view projection matrices the Siddon backprojector needs, fitted on the
corner rays of a uniform detector.
"""
import numpy as np
import torch

from . import _lib
from .device import require_gpu, stream
from .proj_geom import geom_circular, view_projections

_MODES = {"par": 0, "fan": 1, "cone": 2}


def _defaults(DetsY, geom, src_iso, det_iso):
    if geom == "par":
        if det_iso is None:
            det_iso = np.max(DetsY) * 100
        if src_iso is None:
            src_iso = np.max(DetsY) * 100
    elif src_iso is None or det_iso is None:
        raise ValueError(f"src_iso and det_iso must be provided with geom={geom}")
    return float(src_iso), float(det_iso)


def _trig(Views):
    # one scalar np.cos / np.sin per view, as geom_circular evaluates them
    c = np.array([np.cos(v) for v in Views.reshape(-1)], np.float64)
    s = np.array([np.sin(v) for v in Views.reshape(-1)], np.float64)
    return c, s


def circular_rays(DetsY, Views, geom="par", src_iso=None, det_iso=None, DetsZ=None,
                  device="cuda"):
    """Device float32 (src, trg) of shape Views.shape + DetsY.shape
    [+ DetsZ.shape] + (3,)."""
    require_gpu()
    if geom not in _MODES:
        raise ValueError("Cannot process geom: ", geom)
    DetsY = np.asarray(DetsY, np.float64).reshape(-1)
    Views = np.asarray(Views, np.float64)
    if geom == "cone" and DetsZ is None:
        raise ValueError("DetsZ must be provided with geom=cone")
    src_iso, det_iso = _defaults(DetsY, geom, src_iso, det_iso)
    c, s = _trig(Views)
    shape = Views.shape + DetsY.shape
    nv = 1
    parts = [c, s, DetsY]
    if DetsZ is not None:
        DetsZ = np.asarray(DetsZ, np.float64).reshape(-1)
        shape = shape + DetsZ.shape
        nv = DetsZ.size
        parts.append(DetsZ)
    # one upload for the cosines, sines and detector coordinates
    dev = torch.from_numpy(np.concatenate(parts)).to(device)
    na, ny = c.size, DetsY.size
    cd, sd, dy = dev[:na], dev[na:2 * na], dev[2 * na:2 * na + ny]
    dz = dev[2 * na + ny:] if DetsZ is not None else None
    src = torch.empty(shape + (3,), dtype=torch.float32, device=device)
    trg = torch.empty(shape + (3,), dtype=torch.float32, device=device)
    _lib.check(_lib.lib().vir_geom_circular(_MODES[geom], c.size, _lib.ptr(cd), _lib.ptr(sd),
                                            DetsY.size, _lib.ptr(dy), nv, _lib.ptr(dz), src_iso,
                                            det_iso, _lib.ptr(src), _lib.ptr(trg), stream()),
               "vir_geom_circular")
    return src, trg


def square_rays(BinsX, distY, Angs, device="cuda"):
    """Device float32 (src, trg) of shape (nBins, nAngs, 4, 3): the float32
    cast of geom_square(BinsX, distY, Angs) (proj_geom.py:12-57), generated
    by vir_geom_square from the bin positions and numpy's per-angle offsets
    instead of being built on the host and copied over PCIe."""
    from .proj_geom import square_steps
    require_gpu()
    BinsX = np.asarray(BinsX)
    Angs = np.asarray(Angs)
    D, E = square_steps(distY, Angs)
    nb, na = BinsX.size, Angs.size
    t = lambda a: torch.from_numpy(np.ascontiguousarray(np.asarray(a, np.float64).reshape(-1))).to(device)  # noqa: E731
    bx, nbx, dd, ee = t(BinsX), t(-BinsX), t(D), t(E)
    src = torch.empty((nb, na, 4, 3), dtype=torch.float32, device=device)
    trg = torch.empty_like(src)
    _lib.check(_lib.lib().vir_geom_square(nb, _lib.ptr(bx), _lib.ptr(nbx), na, float(distY), _lib.ptr(dd),
                                          _lib.ptr(ee), _lib.ptr(src), _lib.ptr(trg), stream()),
               "vir_geom_square")
    return src, trg


def _uniform(a):
    if a.size < 3:
        return True
    step = (a[-1] - a[0]) / (a.size - 1)
    dev = np.abs(a - (a[0] + step * np.arange(a.size))).max()
    return dev <= 1e-9 * max(abs(step), 1e-300) * a.size


def circular_projections(DetsY, Views, geom="par", src_iso=None, det_iso=None, DetsZ=None):
    """Per-view 3x4 projection matrices (float32, (nviews, 12)) of the ray
    set circular_rays returns, fitted on the four corner rays of each view.
    The detector coordinates must be uniformly spaced."""
    DetsY = np.asarray(DetsY, np.float64).reshape(-1)
    Views = np.asarray(Views, np.float64).reshape(-1)
    if not _uniform(DetsY) or (DetsZ is not None and not _uniform(np.asarray(DetsZ, np.float64))):
        raise ValueError("detector coordinates are not uniformly spaced")
    src_iso, det_iso = _defaults(DetsY, geom, src_iso, det_iso)
    cy = DetsY[[0, -1]] if DetsY.size > 1 else DetsY
    cz = None
    if DetsZ is not None:
        DetsZ = np.asarray(DetsZ, np.float64).reshape(-1)
        cz = DetsZ[[0, -1]] if DetsZ.size > 1 else DetsZ
    s, t = geom_circular(cy, Views, geom, src_iso=src_iso, det_iso=det_iso, DetsZ=cz)
    s = s.astype(np.float32).reshape(Views.size, cy.size, 1 if cz is None else cz.size, 3)
    t = t.astype(np.float32).reshape(s.shape)
    mats = view_projections(s, t).astype(np.float64).reshape(-1, 3, 4)
    if DetsY.size > 1:
        mats[:, 0] *= DetsY.size - 1
    if cz is not None and DetsZ.size > 1:
        mats[:, 1] *= DetsZ.size - 1
    return mats.reshape(-1, 12).astype(np.float32)
