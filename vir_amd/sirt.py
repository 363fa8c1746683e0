"""SIRT on the GPU: the reference's simultaneous iterative reconstruction
loop with device-resident state.

Recurrence (sys_mat/sirt_dd.py:37-76, tests/Test_SIRT/sirt_test.py:21-40,
tests/Test_SIRT/2d_sirt.py:18-35, processing/SIRT/sirt.py:14-44):

    r = (b - A x) / W_sino     where W_sino != 0   (else r = A x)
    u = A^T r / W_img          where W_img  != 0   (else u = A^T r)
    x = x + gamma * u

with W_sino = A 1 and W_img = A^T 1 (sirt_test.py:157-158).  Projectors
that offer it (fused_sirt: the Siddon brick gather from per-ray records, the
2-D distance-driven pair) write the weighted residual from the forward
projection's and the update from the backprojection's write-back, so an
iteration is two launches; otherwise the two steps are elementwise kernels
(vir_sirt_residual, vir_sirt_update) and an iteration is four launches.
Both give the same bits.  An iteration can be captured once into a HIP
graph and replayed.

Any projector with fp(x, out=) / bp(y, out=) on CUDA tensors works:
SiddonProjector, DDProjector, StoredMatrix (below).
"""
import numpy as np
import torch

from . import _lib
from .device import as_device_f32, out_tensor, require_gpu, stream


class StoredMatrix:
    """A stored system matrix (the siddons() CSR) as an operator pair:
    CSR gather forward, voxel-major transpose gather backward."""

    def __init__(self, sdlist, nPixels, ray_shape=None):
        from .projection.proj import _to_csr, _device_csr, _is_flat, check_csr
        from .sys_mat.siddon import csr_transpose
        require_gpu()
        self.img_shape = tuple(int(v) for v in np.atleast_1d(nPixels))
        nvox = int(np.prod(self.img_shape))
        ravel = True
        if not _is_flat(sdlist):
            flat = np.asarray(sdlist, dtype=object).reshape(-1)
            first = next((r for r in flat if r is not None), None)
            ravel = first is None or len(first) == 2
            if ray_shape is None:
                ray_shape = np.asarray(sdlist, dtype=object).shape
        csr = _to_csr(sdlist, self.img_shape, ravel)
        check_csr(*csr, nvox)
        self.ind, self.len, self.off = _device_csr(csr, nvox)
        self.nrays = self.off.numel() - 1
        self.sino_shape = tuple(ray_shape) if ray_shape is not None else (self.nrays,)
        if int(np.prod(self.sino_shape)) != self.nrays:
            raise ValueError(f"ray shape {self.sino_shape} does not hold {self.nrays} rays")
        self.nvox = nvox
        self.colptr, self.row, self.clen = csr_transpose(self.ind, self.len, self.off, nvox)

    def fp(self, x, out=None, accumulate=False):
        x = as_device_f32(x)
        if x.numel() != self.nvox:
            raise ValueError("volume size does not match the matrix")
        out = out_tensor(out, self.sino_shape, x.device)
        fl = _lib.VIR_ACCUMULATE if accumulate else 0
        _lib.check(_lib.lib().vir_csr_fp(_lib.ptr(self.ind), _lib.ptr(self.len), _lib.ptr(self.off),
                                         self.nrays, _lib.ptr(x), _lib.ptr(out), fl, stream()),
                   "vir_csr_fp")
        return out

    def bp(self, y, out=None, accumulate=False):
        y = as_device_f32(y)
        if y.numel() != self.nrays:
            raise ValueError("sinogram size does not match the matrix")
        out = out_tensor(out, self.img_shape, y.device)
        fl = _lib.VIR_ACCUMULATE if accumulate else 0
        _lib.check(_lib.lib().vir_csc_bp(_lib.ptr(self.colptr), _lib.ptr(self.row), _lib.ptr(self.clen),
                                         self.nvox, _lib.ptr(y), _lib.ptr(out), fl, stream()),
                   "vir_csc_bp")
        return out


def _shapes(P):
    img = getattr(P, "img_shape", None) or getattr(P, "nPixels")
    sino = getattr(P, "sino_shape", None) or getattr(P, "ray_shape")
    return tuple(img), tuple(sino)


class SIRT:
    """Device-resident SIRT for one projector.

    W_sino / W_img default to A 1 and A^T 1.  step() runs one iteration in
    place on self.x; run(n) runs n; graph=True captures one iteration into a
    HIP graph on first use and replays it.

    nbatch > 1 runs that many reconstructions on the same projector at once
    (b and x0 with a leading batch dimension, shared W_sino / W_img) through
    the projector's batched fused steps (fp_residual_batch /
    bp_update_batch): each projection traces the rays once for a group of
    items, and every item's iterates equal those of its own SIRT, bit for
    bit."""

    def __init__(self, P, b, x0=None, W_sino=None, W_img=None, gamma=1.0, graph=False, fused=True, nbatch=1):
        """fused: use the projector's fused steps when it has them."""
        require_gpu()
        self.P = P
        self.nbatch = int(nbatch)
        if self.nbatch < 1:
            raise ValueError("nbatch must be >= 1")
        self.fused = bool(fused) and bool(getattr(P, "fused_sirt", False))
        if self.nbatch > 1 and not (self.fused and hasattr(P, "fp_residual_batch")):
            raise ValueError("a batched SIRT needs a projector with batched fused steps")
        self.img_shape, self.sino_shape = _shapes(P)
        lead = (self.nbatch,) if self.nbatch > 1 else ()
        self.b = as_device_f32(b).reshape(lead + self.sino_shape).contiguous()
        dev = self.b.device
        if W_sino is None:
            W_sino = P.fp(torch.ones(self.img_shape, dtype=torch.float32, device=dev))
        if W_img is None:
            W_img = P.bp(torch.ones(self.sino_shape, dtype=torch.float32, device=dev))
        self.W_sino = as_device_f32(W_sino).reshape(self.sino_shape).contiguous()
        self.W_img = as_device_f32(W_img).reshape(self.img_shape).contiguous()
        self.x = (torch.zeros(lead + self.img_shape, dtype=torch.float32, device=dev) if x0 is None
                  else as_device_f32(x0).reshape(lead + self.img_shape).clone())
        self.r = torch.empty(lead + self.sino_shape, dtype=torch.float32, device=dev)
        self.u = torch.empty(self.img_shape, dtype=torch.float32, device=dev) if self.nbatch == 1 else None
        self.gamma = float(gamma)
        self.use_graph = graph
        self._graph = None
        self.iterations = 0

    def _iteration(self):
        if self.nbatch > 1:
            self.P.fp_residual_batch(self.x, self.b, self.W_sino, self.r)
            self.P.bp_update_batch(self.r, self.W_img, self.gamma, self.x)
            return
        if self.fused:
            self.P.fp_residual(self.x, self.b, self.W_sino, self.r)
            self.P.bp_update(self.r, self.W_img, self.gamma, self.x)
            return
        L = _lib.lib()
        st = stream()
        self.P.fp(self.x, out=self.r)
        _lib.check(L.vir_sirt_residual(_lib.ptr(self.b), _lib.ptr(self.r), _lib.ptr(self.W_sino),
                                       self.r.numel(), st), "vir_sirt_residual")
        self.P.bp(self.r, out=self.u)
        _lib.check(L.vir_sirt_update(_lib.ptr(self.x), _lib.ptr(self.u), _lib.ptr(self.W_img),
                                     self.gamma, self.x.numel(), st), "vir_sirt_update")

    def step(self):
        if not self.use_graph:
            self._iteration()
        else:
            if self._graph is None:
                # warm up on a side stream, then capture one iteration; the
                # warm-up iteration is undone by restoring x
                x_saved = self.x.clone()
                s = torch.cuda.Stream()
                s.wait_stream(torch.cuda.current_stream())
                with torch.cuda.stream(s):
                    self._iteration()
                torch.cuda.current_stream().wait_stream(s)
                self.x.copy_(x_saved)
                g = torch.cuda.CUDAGraph()
                with torch.cuda.graph(g):
                    self._iteration()
                self._graph = g
                # capture does not execute: run the first iteration now
            self._graph.replay()
        self.iterations += 1
        return self.x

    def run(self, n):
        for _ in range(int(n)):
            self.step()
        return self.x


def sirt(sino, sdlist, W_sino, W_img, nPixels, gamma=1, beta=0, iters=10, tv_kw=None):
    """Drop-in for tests/Test_SIRT/sirt_test.py:21-40: x0 = 0, `iters`
    updates, returns every iterate as float32 array nPixels + (iters,).
    beta > 0 applies total-variation denoising with weight beta after each
    update (sirt_test.py:37-38), vir_amd.tv.denoise_tv_chambolle with its
    defaults or the keyword arguments in tv_kw."""
    nPixels = tuple(int(v) for v in np.atleast_1d(nPixels))
    P = StoredMatrix(sdlist, nPixels, ray_shape=np.shape(sino))
    S = SIRT(P, sino, W_sino=W_sino, W_img=W_img, gamma=gamma)
    out = np.zeros(nPixels + (iters,), dtype=np.float32)
    for i in range(iters):
        S.step()
        if beta > 0:
            from .tv import denoise_tv_chambolle
            S.x.copy_(denoise_tv_chambolle(S.x, weight=beta, **(tv_kw or {})))
        out[..., i] = S.x.cpu().numpy()
    return out


def sirt_iterates(P, sino, iters, gamma=1.0, x0=None, W_sino=None, W_img=None, first=0):
    """Generic driver: returns the final iterate after `iters` updates of the
    recurrence starting from x0 (x0 = 1e-5 and iters-1 updates reproduces
    sys_mat/sirt_dd.py:46-72)."""
    S = SIRT(P, sino, x0=x0, W_sino=W_sino, W_img=W_img, gamma=gamma)
    S.run(iters)
    return S.x
