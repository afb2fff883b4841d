"""Array plumbing between the caller's type and device tensors.

The reference API takes and returns NumPy arrays.  Here a NumPy input is
uploaded once and the result handed back as NumPy (the plain drop-in); a torch
tensor already on the GPU stays there and the result is a device tensor (the
device-resident form).
"""

import numpy as np
import torch

from . import _lib


def to_device(a, dtype=torch.float32, device='cuda'):
    """(tensor on the GPU, caller_was_numpy)."""
    if isinstance(a, torch.Tensor):
        if a.device.type != 'cuda':
            raise ValueError(
                "steam_amd: tensors must live on the GPU (device='cuda'); "
                f"got a tensor on {a.device}")
        t = a if a.dtype == dtype else a.to(dtype)
        return t.contiguous(), False
    arr = np.ascontiguousarray(a, dtype=torch.empty((), dtype=dtype).numpy().dtype)
    return torch.from_numpy(arr).to(device, non_blocking=False), True


def to_caller(t, was_numpy):
    if was_numpy:
        return t.cpu().numpy()
    return t


def check_device(device):
    return _lib.require_gpu(device)
