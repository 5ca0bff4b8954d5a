"""Symmetrisation wrapper (jVMC nets/sym_wrapper.py:1-83).

``SymNet(orbit, net, avgFun)`` describes log psi(s) = avgFun({log psi_net(g s)},
orbit.factor) over the orbit of lattice symmetries.  The NQS evaluates it with
the HIP entry points jv_sym_forward / jv_sym_gradients / jv_sym_eloc /
jv_sym_mc_sample; the averaging functions below name the reduction the device
applies (``SymNet`` dispatches on their identity) and compute the same value for
an array of coefficients of one configuration (last axis = orbit).
"""
import torch

from .. import _lib


def avgFun_Coefficients_Exp(coeffs, sym_factor):
    """sym_wrapper.py:7-9: logsumexp(coeffs, b=sym_factor)."""
    c = torch.as_tensor(coeffs)
    b = torch.as_tensor(sym_factor, device=c.device)
    cpx = c.is_complex() or b.is_complex()
    a = c.to(torch.complex128 if cpx else torch.float64)
    a = torch.where(b != 0, a, torch.full_like(a, float("-inf")))
    amax = torch.max(a.real, dim=-1, keepdim=True).values
    amax = torch.where(torch.isfinite(amax), amax, torch.zeros_like(amax))
    S = (torch.exp(a - amax) * b).sum(dim=-1)
    if cpx:
        return torch.log(S) + amax[..., 0]
    out = torch.log(torch.abs(S)) + amax[..., 0]
    return torch.where(S < 0, torch.full_like(out, float("nan")), out)


def avgFun_Coefficients_Log(coeffs, sym_factor):
    """sym_wrapper.py:12-13: the mean of the coefficients."""
    return torch.mean(torch.as_tensor(coeffs), dim=-1)


def avgFun_Coefficients_Sep(coeffs, sym_factor):
    """sym_wrapper.py:16-19: amplitudes and phases averaged separately."""
    c = torch.as_tensor(coeffs).to(torch.complex128)
    re, im = c.real, c.imag
    return 0.5 * torch.log(torch.mean(torch.exp(2 * re), dim=-1)) + 1j * torch.angle(
        torch.mean(torch.exp(1j * im), dim=-1))


_AVG_CODES = {avgFun_Coefficients_Exp: _lib.JV_AVG_EXP, avgFun_Coefficients_Log: _lib.JV_AVG_LOG,
              avgFun_Coefficients_Sep: _lib.JV_AVG_SEP}


class SymNet:
    """sym_wrapper.py:22-83: the network ``net`` averaged over ``orbit``
    (a ``util.symmetries.LatticeSymmetry``) with ``avgFun``."""

    is_generator = False

    def __init__(self, orbit, net, avgFun=avgFun_Coefficients_Exp):
        if isinstance(net, SymNet):
            raise NotImplementedError("nested SymNet")
        if avgFun not in _AVG_CODES:
            raise NotImplementedError("avgFun must be one of avgFun_Coefficients_Exp / _Log / _Sep")
        self.orbit = orbit
        self.net = net
        self.avgFun = avgFun

    @property
    def avg_code(self):
        return _AVG_CODES[self.avgFun]

    def descriptor(self, sampleShape):
        return self.net.descriptor(sampleShape)

    def init_params(self, desc, rng):
        return self.net.init_params(desc, rng)


__all__ = ["SymNet", "avgFun_Coefficients_Exp", "avgFun_Coefficients_Log", "avgFun_Coefficients_Sep"]
