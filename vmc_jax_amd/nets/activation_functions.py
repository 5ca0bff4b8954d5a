"""Activation names understood by the HIP kernels (activation_functions.py:19-22).

Only ``log_cosh`` is implemented on the device; it is passed around as this
module's function object, as in jVMC.
"""
import numpy as np


def log_cosh(x):
    """Host reference of the stable log(cosh(x)) (activation_functions.py:19-22)."""
    x = np.asarray(x)
    sgn = np.where(np.signbit(np.real(x)), -1.0, 1.0)
    y = x * sgn
    return y + np.log1p(np.exp(-2.0 * y)) - np.log(2.0)
