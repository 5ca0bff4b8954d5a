"""Network descriptors mirroring jVMC's Flax modules.

``CpxRBM`` (rbm.py:17-38), ``RBM`` (rbm.py:67-88) and ``CNN`` (cnn.py:18-81)
take the same constructor arguments as in jVMC.  They hold no parameters: the
NQS owns the flat parameter vector on the device and the HIP kernels evaluate
the network from it.  The lattice shape is taken from the first configuration
the NQS sees, as Flax does at ``init``.
"""
import numpy as np

from . import activation_functions
from .activation_functions import log_cosh
from .. import _lib
from . import sym_wrapper
from .sym_wrapper import SymNet


class _NetBase:
    is_generator = False

    def descriptor(self, sampleShape):
        raise NotImplementedError

    def init_params(self, desc, rng):
        raise NotImplementedError


class CpxRBM(_NetBase):
    """Complex RBM: log psi = sum_h log_cosh(W^T (2s-1) + b) (rbm.py:30-38).

    ``visibleBias`` adds sum_i a_i (2 s_i - 1); it is an extension (jVMC's
    CpxRBM has no visible bias) stored as the leaf ``visible_bias`` after
    ``Dense_0`` in the flat vector.
    """

    def __init__(self, numHidden=2, bias=False, visibleBias=False):
        self.numHidden = int(numHidden)
        self.bias = bool(bias)
        self.visibleBias = bool(visibleBias)

    def descriptor(self, sampleShape):
        d = _lib.JvNet()
        d.kind = _lib.JV_NET_CPXRBM
        shape = tuple(int(s) for s in sampleShape)
        d.ndim = 1
        d.shape[0] = int(np.prod(shape))
        d.shape[1] = 1
        d.num_hidden = self.numHidden
        d.hidden_bias = int(self.bias)
        d.visible_bias = int(self.visibleBias)
        return d

    def init_params(self, desc, rng):
        # cplx_init: U[0, 0.01) for real and imaginary parts (initializers.py:17-20);
        # biases zero.  Same distribution, not the same bits as Flax's PRNG.
        P = np.zeros(desc.num_real_params)
        k0 = desc.off_kernel
        n = desc.num_sites * desc.num_hidden
        P[k0:k0 + 2 * n] = 0.01 * rng.uniform(size=2 * n)
        return P


class RBM(_NetBase):
    """Real RBM: log psi = sum log cosh(W^T (2s-1) + b) (rbm.py:80-88)."""

    def __init__(self, numHidden=2, bias=False):
        self.numHidden = int(numHidden)
        self.bias = bool(bias)

    def descriptor(self, sampleShape):
        if len(sampleShape) != 1:
            raise NotImplementedError("the real RBM acts on the last axis (rbm.py:88); use 1D lattices")
        d = _lib.JvNet()
        d.kind = _lib.JV_NET_RBM
        d.ndim = 1
        d.shape[0] = int(sampleShape[0])
        d.shape[1] = 1
        d.num_hidden = self.numHidden
        d.hidden_bias = int(self.bias)
        return d

    def init_params(self, desc, rng):
        # lecun_normal kernel (truncated normal, std 1/sqrt(fan_in)), zero bias
        P = np.zeros(desc.num_real_params)
        k0 = desc.off_kernel
        n = desc.num_sites * desc.num_hidden
        std = 1.0 / np.sqrt(desc.num_sites) / 0.87962566103423978
        P[k0:k0 + n] = np.clip(rng.normal(size=n), -2, 2) * std
        return P


class CNN(_NetBase):
    """Real-parameter CNN (cnn.py:18-81): circular padding, unit strides."""

    def __init__(self, F=(8,), channels=(10,), strides=(1,), actFun=(log_cosh,), bias=True,
                 firstLayerBias=False, periodicBoundary=True):
        if not periodicBoundary:
            raise NotImplementedError("only periodicBoundary=True is implemented")
        if any(int(s) != 1 for s in strides):
            raise NotImplementedError("only unit strides are implemented")
        for f in actFun:
            if f is not log_cosh and getattr(f, "__name__", "") != "log_cosh":
                raise NotImplementedError("only the log_cosh activation is implemented")
        if len(channels) > _lib.JV_MAX_LAYERS:
            raise NotImplementedError("at most %d layers" % _lib.JV_MAX_LAYERS)
        self.F = tuple(int(f) for f in F)
        self.channels = tuple(int(c) for c in channels)
        self.strides = tuple(strides)
        self.bias = bool(bias)
        self.firstLayerBias = bool(firstLayerBias)

    def descriptor(self, sampleShape):
        shape = tuple(int(s) for s in sampleShape)
        if len(shape) != len(self.F) or len(shape) not in (1, 2):
            raise ValueError("filter rank must match a 1D or 2D lattice")
        d = _lib.JvNet()
        d.kind = _lib.JV_NET_CNN
        d.ndim = len(shape)
        d.shape[0] = shape[0]
        d.shape[1] = shape[1] if len(shape) == 2 else 1
        d.num_layers = len(self.channels)
        d.filter[0] = self.F[0]
        d.filter[1] = self.F[1] if len(self.F) == 2 else 1
        for l, c in enumerate(self.channels):
            d.channels[l] = c
            d.layer_bias[l] = int(self.bias if l > 0 else self.firstLayerBias)
        return d

    def init_params(self, desc, rng):
        # variance_scaling(1.0, "fan_avg", "uniform") kernels, zero biases (cnn.py:132)
        P = np.zeros(desc.num_real_params)
        FF = desc.filter[0] * desc.filter[1]
        for l in range(desc.num_layers):
            cin = 1 if l == 0 else desc.channels[l - 1]
            cout = desc.channels[l]
            n = FF * cin * cout
            lim = np.sqrt(3.0 / ((FF * cin + FF * cout) / 2.0))
            o = desc.off_conv_kernel[l]
            P[o:o + n] = rng.uniform(-lim, lim, size=n)
        return P

    def he_normal(self, desc, rng):
        """He-normal kernels N(0, 2/fan_in), fan_in = prod(F) * C_in; zero biases."""
        P = np.zeros(desc.num_real_params)
        FF = desc.filter[0] * desc.filter[1]
        for l in range(desc.num_layers):
            cin = 1 if l == 0 else desc.channels[l - 1]
            n = FF * cin * desc.channels[l]
            o = desc.off_conv_kernel[l]
            P[o:o + n] = rng.normal(size=n) * np.sqrt(2.0 / (FF * cin))
        return P


__all__ = ["CpxRBM", "RBM", "CNN", "SymNet", "sym_wrapper", "activation_functions", "log_cosh"]
