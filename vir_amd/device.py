"""Device plumbing: torch-ROCm tensors as HBM buffers, HIP streams.

torch provides allocation and streams only; every projector computation
runs in libvir_amd.so.  Without a GPU these helpers raise.
"""
import numpy as np
import torch

from . import _lib


def require_gpu():
    if not torch.cuda.is_available():
        raise _lib.VirError("vir_amd needs a ROCm GPU (torch.cuda.is_available() is False)")
    _lib.lib()


def as_device_f32(a, device="cuda"):
    """numpy / torch -> contiguous float32 tensor on the GPU."""
    if isinstance(a, torch.Tensor):
        t = a.to(device=device, dtype=torch.float32)
    else:
        t = torch.from_numpy(np.array(a, dtype=np.float32, order="C", copy=True)).to(device)
    return t.contiguous()


def as_device(a, dtype, device="cuda"):
    if isinstance(a, torch.Tensor):
        return a.to(device=device, dtype=dtype).contiguous()
    return torch.from_numpy(np.ascontiguousarray(a)).to(device=device, dtype=dtype).contiguous()


def out_tensor(out, shape, device, what="out"):
    """`out` checked as a kernel's output buffer (a contiguous float32 GPU
    tensor of prod(shape) elements), or a new one of `shape` when None.  A
    kernel writes every element it is given, so a short or strided buffer
    would be written out of bounds: refuse it before anything is launched."""
    if out is None:
        return torch.empty(shape, dtype=torch.float32, device=device)
    if not isinstance(out, torch.Tensor) or not out.is_cuda or out.dtype != torch.float32:
        raise TypeError(f"{what} must be a float32 CUDA tensor")
    if not out.is_contiguous():
        raise ValueError(f"{what} must be contiguous")
    n = int(np.prod(shape))
    if out.numel() != n:
        raise ValueError(f"{what} has {out.numel()} elements, expected {n}")
    return out


def require_cuda_f32(*tensors, what="argument"):
    """TypeError unless every tensor is a float32 CUDA tensor: their
    data_ptr() goes straight to a kernel, which would read a host pointer or
    reinterpret another dtype's bytes."""
    for t in tensors:
        if not isinstance(t, torch.Tensor) or not t.is_cuda or t.dtype != torch.float32:
            raise TypeError(f"{what} must be float32 CUDA tensors")


def stream():
    return _lib.stream_ptr()
