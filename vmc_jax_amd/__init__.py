"""MI355X-native implementation of jVMC's VMC inner loop.

Drop-in for jVMC's ``NQS``, ``MCSampler`` / ``ExactSampler``,
``BranchFreeOperator``, ``SampledObs``, ``TDVP`` / ``MinSR`` and steppers: the
Python classes keep jVMC's names, arguments and return shapes and call the HIP
kernels of ``libjvmc_hip.so`` (gfx950) through a C ABI.
"""
from . import _lib  # noqa: F401  (fails loudly when the library is not built)
from . import global_defs, mpi_wrapper, nets, operator, sampler, stats, util, vqs
from .vqs import NQS

__version__ = "0.1.0"
__all__ = ["global_defs", "mpi_wrapper", "nets", "operator", "sampler", "stats", "util", "vqs", "NQS"]
