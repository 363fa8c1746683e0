"""The single-ray C Siddon of sys_mat/siddons.c on the GPU, batched.

siddons.c is not equivalent to sys_mat/siddon.py: it ravels voxel indices x
fastest (nx ny z + nx y + x), its entry/exit range ignores the z slab, it
snaps nothing, and its grid lines are accumulated float sums.  These calls
reproduce it, bit for bit on rays where it stays inside its arrays (tested
against the reference's own siddons.c, compiled unmodified):

  calc_grid_lines(nPixels, dPixels)          siddons.c:137-148 (host)
  siddons(src, trg, nPixels, dPixels)        siddons.c:290-341, calc_inter :235-261
      one ray (3,) -> its rows (idx int32, length float32) up to the idx = -1
      terminator; rays (..., 3) -> (idx, length, offsets int64) CSR
  siddons_proj(iArray, src, trg, nPixels, dPixels)
                                             siddons.c:345-396, sum_inter :265-287
      ray sums in float32 of iArray (flat, x fastest: (nz, ny, nx))
  forward_project(iArray, nPixels, dPixels)  siddons.c:403-460
      the file's fixed 4-view, 4 x 4-ray demonstration, returning the 64
      values it writes into dArray
  siddon_c(src, trg, nPixels, dPixels), forward_project_c(iArray, dArray,
      nPixels, dPixels)                      sys_mat/siddon.py:887-966 (the
                                             ctypes callers of siddons.so)

Differences, each where siddons.c is undefined: a ray missing the grid has
no rows (siddons.c returns without touching its row buffer); planes outside
[0, N] are not merged and entries outside the volume add nothing to a
projection (siddons.c reads outside its arrays there: its first plane index
is -1 on rays whose entry rounds a hair before plane 0, about half of all
rays, and its range ignores the z slab); forward_project has no limit on
the merged list (its buffer holds 2 nz + 3 values, siddons.c:424).
"""
import numpy as np
import torch

from .. import _lib
from ..device import require_gpu, stream


def calc_grid_lines(nPixels, dPixels):
    """float32 grid lines (nx+1, ny+1, nz+1): line 0 at -n d / 2, then each
    the previous plus d, in float32 (siddons.c:137-148)."""
    n = np.broadcast_to(np.asarray(nPixels, np.int64), (3,))
    d = np.broadcast_to(np.asarray(dPixels, np.float32), (3,))
    out = []
    for a in range(3):
        lines = np.empty(int(n[a]) + 1, np.float32)
        lines[0] = np.float32(np.float32(-n[a]) * d[a]) / np.float32(2.0)
        for i in range(1, int(n[a]) + 1):
            lines[i] = lines[i - 1] + d[a]
        out.append(lines)
    return tuple(out)


class _Grid:
    def __init__(self, nPixels, dPixels):
        require_gpu()
        n = np.broadcast_to(np.asarray(nPixels, np.int64), (3,))
        d = np.broadcast_to(np.asarray(dPixels, np.float32), (3,))
        g = _lib.VirScGrid()
        g.nx, g.ny, g.nz = (int(v) for v in n)
        g.dx, g.dy, g.dz = (float(v) for v in d)
        self.g = g
        self.n = tuple(int(v) for v in n)
        self.lines = [torch.from_numpy(a).cuda() for a in calc_grid_lines(n, d)]

    def args(self):
        return (self.g, *(_lib.ptr(t) for t in self.lines))


def _rays(src, trg):
    trg = np.array(trg, dtype=np.float32)
    src = np.array(src, dtype=np.float32)
    single = trg.ndim == 1
    trg = np.atleast_2d(trg)
    src = np.broadcast_to(np.atleast_2d(src), trg.shape)
    if trg.shape[-1] != 3:
        raise ValueError("src and trg must have a trailing dimension of 3")
    s = torch.from_numpy(np.ascontiguousarray(src.reshape(-1, 3))).cuda()
    t = torch.from_numpy(np.ascontiguousarray(trg.reshape(-1, 3))).cuda()
    return s, t, trg.shape[:-1], single


def siddons(src, trg, nPixels, dPixels):
    G = _Grid(nPixels, dPixels)
    s, t, shape, single = _rays(src, trg)
    nr = s.shape[0]
    L = _lib.lib()
    counts = torch.empty(max(nr, 1), dtype=torch.int32, device="cuda")
    _lib.check(L.vir_siddons_c_count(_lib.ptr(s), _lib.ptr(t), nr, *G.args(), _lib.ptr(counts), stream()),
               "vir_siddons_c_count")
    off = torch.zeros(nr + 1, dtype=torch.int64, device="cuda")
    torch.cumsum(counts[:nr].to(torch.int64), 0, out=off[1:])
    nnz = int(off[-1].item())
    idx = torch.empty(max(nnz, 1), dtype=torch.int32, device="cuda")
    ln = torch.empty(max(nnz, 1), dtype=torch.float32, device="cuda")
    _lib.check(L.vir_siddons_c_fill(_lib.ptr(s), _lib.ptr(t), nr, *G.args(), _lib.ptr(off), _lib.ptr(idx),
                                    _lib.ptr(ln), stream()), "vir_siddons_c_fill")
    idx, ln, off = idx[:nnz].cpu().numpy(), ln[:nnz].cpu().numpy(), off.cpu().numpy()
    if single:
        return idx, ln
    return idx, ln, off


def siddons_proj(iArray, src, trg, nPixels, dPixels):
    """float32 ray sums, shaped like the rays (a scalar for one ray)."""
    G = _Grid(nPixels, dPixels)
    s, t, shape, single = _rays(src, trg)
    vol = torch.from_numpy(np.ascontiguousarray(np.asarray(iArray, np.float32).reshape(-1))).cuda()
    if vol.numel() != G.n[0] * G.n[1] * G.n[2]:
        raise ValueError("iArray does not match nPixels")
    out = torch.empty(s.shape[0], dtype=torch.float32, device="cuda")
    _lib.check(_lib.lib().vir_siddons_c_proj(_lib.ptr(s), _lib.ptr(t), s.shape[0], *G.args(), _lib.ptr(vol),
                                             _lib.ptr(out), None, stream()), "vir_siddons_c_proj")
    out = out.cpu().numpy()
    return out[0] if single else out.reshape(shape)


def demo_rays(nPixels, dPixels):
    """The rays of siddons.c:403-460's forward_project -- 4 views of 4 x 4
    parallel rays -- with its float/double evaluation order; (64, 3) each,
    ordered as dArray (16 angle + 4 xy + z)."""
    n = np.broadcast_to(np.asarray(nPixels, np.int64), (3,))
    d = np.broadcast_to(np.asarray(dPixels, np.float32), (3,))
    f32, f64 = np.float32, np.float64
    d_ang = f32(6.283185307179586 / 4)
    s2 = f32(f32(n[0] * n[0]) * d[0]) + f32(f32(n[1] * n[1]) * d[1])
    dist = f32(np.sqrt(f64(s2)))
    src = np.zeros((64, 3), np.float32)
    trg = np.zeros((64, 3), np.float32)
    for a in range(4):
        arg = f64(f32(f32(a) * d_ang))
        c, s = f32(np.cos(arg)), f32(np.sin(arg))
        sd, cd = f64(f32(s * dist)), f64(f32(c * dist))
        for xy in range(4):
            off = xy - 1.5
            for z in range(4):
                k = 16 * a + 4 * xy + z
                src[k] = (f32(f64(c) * off - sd), f32(f64(s) * off + cd), f32(z - 1.5))
                trg[k] = (f32(f64(c) * off + sd), f32(f64(s) * off - cd), f32(z - 1.5))
    return src, trg


def forward_project(iArray, nPixels, dPixels):
    src, trg = demo_rays(nPixels, dPixels)
    return siddons_proj(iArray, src, trg, nPixels, dPixels)


def siddon_c(src, trg, nPixels=128, dPixels=1.0):
    """sys_mat/siddon.py:887-933: the rows of one ray, as (idx, length)."""
    return siddons(src, trg, nPixels, dPixels)


def forward_project_c(iArray, dArray, nPixels=128, dPixels=1.0):
    """sys_mat/siddon.py:937-966: forward_project into dArray (64 float32
    values, written in place)."""
    out = forward_project(iArray, nPixels, dPixels)
    np.asarray(dArray).reshape(-1)[:64] = out
