"""Argument checks of the Python bindings, run before any pointer reaches
the C ABI: a buffer of the wrong shape, type, layout or device would make a
copy or kernel read or write past its end."""
from __future__ import annotations

import numpy as np


def host_array(a, shape, dtype, name: str) -> np.ndarray:
    """A C-contiguous numpy array of exactly `shape` and `dtype`."""
    if not isinstance(a, np.ndarray):
        raise ValueError(f"{name} must be a numpy array")
    if a.shape != tuple(shape) or a.dtype != dtype or not a.flags["C_CONTIGUOUS"]:
        raise ValueError(f"{name} must be a C-contiguous {np.dtype(dtype).name} array of shape "
                         f"{tuple(shape)} (got {a.dtype.name} {a.shape})")
    return a


def host_input(a, shape, dtype, name: str) -> np.ndarray:
    """An input of exactly `shape`, converted to a C-contiguous `dtype` copy
    if needed."""
    a = np.asarray(a)
    if a.shape != tuple(shape):
        raise ValueError(f"{name} must have shape {tuple(shape)} (got {a.shape})")
    return np.ascontiguousarray(a, dtype)


def device_tensor(t, shape, dtype, name: str):
    """A contiguous torch CUDA tensor of exactly `shape` and `dtype`."""
    import torch
    if not isinstance(t, torch.Tensor):
        raise ValueError(f"{name} must be a torch tensor")
    if (not t.is_cuda or t.dtype != dtype or tuple(t.shape) != tuple(shape)
            or not t.is_contiguous()):
        raise ValueError(f"{name} must be a contiguous {dtype} CUDA tensor of shape "
                         f"{tuple(shape)} (got {t.dtype} {tuple(t.shape)} on {t.device})")
    return t
