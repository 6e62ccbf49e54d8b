"""Device-buffer plumbing: torch (ROCm) tensors hold the HIP library's inputs
and outputs; all arithmetic happens in the library's kernels."""

import numpy as np
import torch

from . import _lib

F64 = torch.float64
I32 = torch.int32


def require_gpu():
    """Raise unless a ROCm GPU is visible and the HIP library loads (no CPU fallback)."""
    if not torch.cuda.is_available():
        raise RuntimeError("simkit_amd needs a ROCm GPU (none is visible); there is no CPU fallback")
    _lib.load()


def dev(a, dtype=F64):
    """Contiguous device copy of an array-like (or pass-through for a device tensor)."""
    require_gpu()
    if isinstance(a, torch.Tensor):
        t = a.to(device="cuda", dtype=dtype)
    else:
        t = torch.from_numpy(np.ascontiguousarray(np.asarray(a), dtype=_np_dtype(dtype))).to("cuda")
    return t.contiguous()


def per_element(a, n):
    """Broadcast a scalar / (n,) / (n,1) material array to a device (n,) FP64 vector."""
    arr = np.asarray(a, dtype=np.float64).reshape(-1)
    if arr.size == 1:
        arr = np.full(n, float(arr[0]))
    if arr.size != n:
        raise ValueError(f"expected {n} per-element values, got {arr.size}")
    return dev(arr)


def empty(n, dtype=F64):
    require_gpu()
    return torch.empty(int(n), dtype=dtype, device="cuda")


def zeros(n, dtype=F64):
    require_gpu()
    return torch.zeros(int(n), dtype=dtype, device="cuda")


def ptr(t):
    return None if t is None else t.data_ptr()


def stream():
    return torch.cuda.current_stream().cuda_stream


def host(t):
    return t.detach().cpu().numpy()


def _np_dtype(dtype):
    return {torch.float64: np.float64, torch.float32: np.float32, torch.int32: np.int32,
            torch.uint8: np.uint8, torch.int64: np.int64}[dtype]
