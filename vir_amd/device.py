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


def stream():
    return _lib.stream_ptr()
