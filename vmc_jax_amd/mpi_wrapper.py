"""Single-GPU semantics of jVMC's data-parallel layer (mpi_wrapper.py:53-243).

One process drives one GPU, so ``global_sum`` is the plain sum over the device
and batch axes.  ``distribute_sampling`` keeps the reference's rounding and sets
the module global ``globNumSamples`` that the SNR of the TDVP solve reads
(tdvp.py:181).
"""
import torch

from . import _lib

rank = 0
commSize = 1
globNumSamples = 0


def distribute_sampling(numSamples, localDevices=None, numChainsPerDevice=1):
    """mpi_wrapper.py:53-95 with commSize = 1."""
    global globNumSamples
    if localDevices is None:
        globNumSamples = numSamples
        return numSamples, globNumSamples
    perProc = localDevices * numChainsPerDevice
    spc = int((numSamples + perProc - 1) // perProc)
    globNumSamples = spc * perProc
    return spc, globNumSamples


def first_sample_id():
    return 0


def global_sum(data):
    """mpi_wrapper.py:114-147: sum over the device and batch axes."""
    return data.reshape((-1,) + tuple(data.shape[2:])).sum(dim=0)


def _weighted(data, p):
    d = data.reshape(data.shape[1], -1)
    cpx = torch.is_complex(d)
    d = d.to(torch.complex128 if cpx else torch.float64).contiguous()
    w = p.reshape(-1).to(torch.float64).contiguous()
    return d, w, cpx


def global_mean(data, p):
    """mpi_wrapper.py:150-174: sum_j p_j X_j."""
    d, w, cpx = _weighted(data, p)
    n, m = d.shape
    mean = torch.empty(m, dtype=d.dtype, device=d.device)
    cent = torch.empty_like(d)
    _lib.call("jv_center", _lib.ptr(d), _lib.ptr(w), n, m, int(cpx), _lib.ptr(mean), _lib.ptr(cent),
              _lib.stream_ptr())
    return mean


def global_variance(data, p):
    """mpi_wrapper.py:177-217: sum_j p_j |X_j - <X>|^2."""
    d, w, cpx = _weighted(data, p)
    n, m = d.shape
    mean = torch.empty(m, dtype=d.dtype, device=d.device)
    cent = torch.empty_like(d)
    _lib.call("jv_center", _lib.ptr(d), _lib.ptr(w), n, m, int(cpx), _lib.ptr(mean), _lib.ptr(cent),
              _lib.stream_ptr())
    var = torch.empty(m, dtype=torch.float64, device=d.device)
    _lib.call("jv_col_norm2", _lib.ptr(cent), n, m, int(cpx), _lib.ptr(var), _lib.stream_ptr())
    return var


def gather(data):
    """mpi_wrapper.py:278-292 with one process: drop the device axis."""
    return data.reshape((-1,) + tuple(data.shape[2:]))


def get_communication_time():
    return 0.0
