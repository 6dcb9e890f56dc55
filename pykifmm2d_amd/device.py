"""Device-memory plumbing: torch tensors carry HBM buffers and streams to the C-ABI."""
import ctypes

import numpy as np


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("pykifmm2d_amd needs a ROCm GPU (torch.cuda.is_available() is False); "
                           "there is no CPU fallback")
    return torch


def as_device_f64(a):
    """float64, contiguous, on cuda:0 (copies numpy input; torch input is converted if needed)."""
    torch = _torch()
    if isinstance(a, torch.Tensor):
        return a.to(device="cuda", dtype=torch.float64).contiguous()
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).to("cuda")


def empty_f64(shape):
    torch = _torch()
    return torch.empty(shape, dtype=torch.float64, device="cuda")


def stream_handle(stream=None):
    """hipStream_t of `stream` (a torch stream; default torch's current stream on cuda:0), as a
    ctypes pointer."""
    torch = _torch()
    if stream is None:
        stream = torch.cuda.current_stream()
    return ctypes.c_void_p(stream.cuda_stream)


def ptr(t):
    return ctypes.c_void_p(t.data_ptr())


def to_host(t):
    return t.detach().cpu().numpy()
