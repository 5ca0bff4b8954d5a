"""Numeric types and the device context (global_defs.py:4-54).

The reference forces complex128 / float64 everywhere; so does this package.  One
process drives one GPU; the leading "device" axis of every array is kept with
size 1 so that shapes match jVMC's.
"""
import numpy as np
import torch

tCpx = np.complex128
tReal = np.float64

_device = None


def device():
    global _device
    if _device is None:
        if not torch.cuda.is_available():
            raise RuntimeError("vmc_jax_amd needs a ROCm GPU (gfx950); none is visible")
        _device = torch.device("cuda", torch.cuda.current_device())
    return _device


def set_pmap_devices(devices):
    """Kept for API compatibility: one GPU per process (global_defs.py:38-46)."""
    return None


def device_count():
    return 1


def get_iterable(x):
    import collections.abc
    return x if isinstance(x, collections.abc.Iterable) else (x,)
