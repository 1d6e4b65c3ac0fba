"""Host conversion of backend arrays (``emsim/utils/array.py:10-16``).

The reference accepts numpy and cupy arrays.  The hip backend returns torch
tensors that live on the GPU, so they are accepted here as well.
"""
import numpy as np


def assure_numpy(arr):
    if type(arr) is np.ndarray:
        return arr
    try:
        import torch
    except ImportError:           # pragma: no cover - torch is a hard dependency
        torch = None
    if torch is not None and isinstance(arr, torch.Tensor):
        return arr.detach().cpu().numpy()
    raise ValueError("arr must be either a numpy array or a torch tensor.")
