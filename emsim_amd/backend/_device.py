"""Device-buffer plumbing for the hip backend (torch tensors as HIP memory).

Torch is used only to own device memory (its caching allocator) and to name
the current HIP stream; all arithmetic on the hot path runs in the HIP
library.  Conversions raise TypeError / ValueError for inputs the kernels
cannot take, instead of the silent forcecasts of the reference's pybind
layer (``emsim/backend/fftw_ext/src/wave_kernel.cpp:14-15``).
"""
from ctypes import c_void_p

import numpy as np
import torch


def device() -> torch.device:
    if not torch.cuda.is_available():
        raise RuntimeError("the hip backend needs a ROCm GPU (torch.cuda.is_available() is False)")
    return torch.device("cuda", torch.cuda.current_device())


def stream_handle() -> c_void_p:
    return c_void_p(torch.cuda.current_stream().cuda_stream)


def ptr(t: torch.Tensor) -> c_void_p:
    return c_void_p(t.data_ptr())


def as_device(x, dtype: torch.dtype, name: str) -> torch.Tensor:
    """numpy array or torch tensor -> contiguous tensor of ``dtype`` on the GPU."""
    dev = device()
    if isinstance(x, torch.Tensor):
        t = x
    elif isinstance(x, np.ndarray):
        t = torch.from_numpy(np.ascontiguousarray(x))
    else:
        try:
            t = torch.as_tensor(np.asarray(x))
        except Exception as exc:  # pragma: no cover - defensive
            raise TypeError(f"{name}: expected a numpy array or torch tensor, got {type(x)}") from exc
    if dtype.is_complex != t.dtype.is_complex:
        raise TypeError(f"{name}: expected a {'complex' if dtype.is_complex else 'real'} array, "
                        f"got {t.dtype}")
    if t.dtype != dtype:
        t = t.to(dtype)
    if t.device != dev:
        t = t.to(dev, non_blocking=False)
    return t.contiguous()


def workspace(nbytes: int) -> torch.Tensor:
    return torch.empty(max(int(nbytes), 1), dtype=torch.uint8, device=device())
