"""Samplers (sampler.py:67-534) driving the HIP Metropolis kernels.

``MCSampler`` keeps its walkers on the GPU between calls.  Each ``sample()``
re-initialises the acceptance counters, thermalises for
``thermalizationSweeps * sweepSteps`` proposals and emits the walkers after each
of ``samplesPerChain`` sweeps (sample index t*C + c), exactly the schedule of
sampler.py:301-323.  Proposals come from Philox4x32-10 keyed by ``key`` with the
call index as a counter word (JAX threefry streams cannot be reproduced here).
"""
import math

import numpy as np
import torch

from . import _lib
from . import global_defs
from . import mpi_wrapper as mpi


def propose_spin_flip(key, s, info):
    """Single-spin-flip proposer (sampler.py:15-19).  The HIP kernel implements
    it; the function is the identity tag passed as ``updateProposer``."""
    raise RuntimeError("propose_spin_flip runs inside the Metropolis kernel")


class MCSampler:
    def __init__(self, net, sampleShape, key, updateProposer=None, numChains=1, updateProposerArg=None,
                 numSamples=100, thermalizationSweeps=10, sweepSteps=10, initState=None, mu=2,
                 logProbFactor=0.5):
        if updateProposer is None:
            raise RuntimeError("Instantiation of MCSampler: `updateProposer` is `None` and cannot be used for "
                               "MCMC sampling.")
        if updateProposer is not propose_spin_flip:
            raise NotImplementedError("only propose_spin_flip is implemented on the device")
        if mu < 0 or mu > 2:
            raise ValueError("mu must be in the range [0, 2]")
        self.sampleShape = tuple(sampleShape)
        self.net = net
        self.numChains = int(numChains)
        self.numSamples = numSamples
        self.thermalizationSweeps = thermalizationSweeps
        self.sweepSteps = sweepSteps
        self.mu = mu
        self.logProbFactor = logProbFactor
        self.updateProposer = updateProposer
        self.key = int(key) if not hasattr(key, "__len__") else int(np.asarray(key).ravel()[-1])
        self.callIndex = 0
        L = int(np.prod(self.sampleShape))
        init = np.zeros(L, np.int32) if initState is None else np.asarray(initState, np.int32).ravel()
        self.states = torch.as_tensor(np.repeat(init[None, :], self.numChains, axis=0),
                                      device=global_defs.device()).contiguous()
        self.net(self.states.reshape((1, self.numChains) + self.sampleShape))
        self.counters = torch.zeros(2, dtype=torch.int64, device=global_defs.device())
        self.globNumSamples = 0

    def set_number_of_samples(self, N):
        self.numSamples = N

    def set_random_key(self, key):
        self.key = int(key)

    def get_last_number_of_samples(self):
        return self.globNumSamples

    # sample() accepts logpsi_stream (see there)
    defers_logpsi = True

    def sample(self, parameters=None, numSamples=None, multipleOf=1, *, logpsi_stream=None):
        """sampler.py:195-235.

        logpsi_stream (not in the reference): for the mixed form (FP32
        chains, whose emitted samples' FP64 log psi is a separate forward
        pass) with |psi|^2 sampling (mu = 2, uniform weights), the forward
        pass runs on that stream after the chains, and the returned log psi
        is ready in its order only; the configurations and weights are ready
        on the current stream.  The network's parameters must not change
        before that stream has run the forward pass.  Ignored otherwise, and
        when ``parameters`` is given."""
        if numSamples is None:
            numSamples = self.numSamples
        tmp = None
        if parameters is not None:
            tmp = self.net.get_parameters()
            self.net.set_parameters(parameters)
        spc, self.globNumSamples = mpi.distribute_sampling(
            numSamples, localDevices=1, numChainsPerDevice=int(np.lcm(self.numChains, multipleOf)))
        C = self.numChains
        L = int(np.prod(self.sampleShape))
        dev = global_defs.device()
        configs = torch.empty((spc * C, L), dtype=torch.int32, device=dev)
        logpsi = torch.empty(spc * C, dtype=torch.complex128, device=dev)
        a = _lib.JvMcArgs()
        a.num_chains = C
        a.therm_steps = int(self.thermalizationSweeps * self.sweepSteps)
        a.samples_per_chain = spc
        a.sweep_steps = int(self.sweepSteps)
        a.mu = float(self.mu)
        a.seed = self.key & 0xFFFFFFFFFFFFFFFF
        a.call_index = self.callIndex & 0xFFFFFFFF
        a.stream = 0
        defer = (logpsi_stream is not None and parameters is None and self.net._sym is None and self.net.fp32_gradients
                 and 1.0 / self.logProbFactor - self.mu == 0.0 and spc * C > 0)
        if defer:
            _lib.call("jv_mc_sample", self.net.desc, _lib.ptr(self.net.params), a, _lib.ptr(self.states),
                      _lib.ptr(configs), None, _lib.ptr(self.counters), _lib.stream_ptr())
            self.callIndex += 1
            n = spc * C
            logpsi_stream.wait_stream(torch.cuda.current_stream())
            with torch.cuda.stream(logpsi_stream):
                configs.record_stream(logpsi_stream)
                logpsi.record_stream(logpsi_stream)
                _lib.call("jv_net_forward", self.net.desc, _lib.ptr(self.net.params), _lib.ptr(configs), n,
                          _lib.ptr(logpsi), _lib.stream_ptr())
                out = self.net._out(logpsi.reshape(1, -1))
            # mu = 2: exp(0 * Re log psi) / n, as jv_mc_weights gives it
            p = torch.full((n,), 1.0 / n, dtype=torch.float64, device=dev)
            configs = configs.reshape((1, n) + self.sampleShape)
            return configs, out, p.reshape(1, -1)
        if self.net._sym is not None:
            _lib.call("jv_sym_mc_sample", self.net.desc, _lib.ptr(self.net.params), self.net._sym[0], a,
                      _lib.ptr(self.states), _lib.ptr(configs), _lib.ptr(logpsi), _lib.ptr(self.counters),
                      _lib.stream_ptr())
        else:
            _lib.call("jv_mc_sample", self.net.desc, _lib.ptr(self.net.params), a, _lib.ptr(self.states),
                      _lib.ptr(configs), _lib.ptr(logpsi), _lib.ptr(self.counters), _lib.stream_ptr())
        self.callIndex += 1
        p = torch.empty(spc * C, dtype=torch.float64, device=dev)
        if spc * C > 0:
            _lib.call("jv_mc_weights", _lib.ptr(logpsi), spc * C, float(self.mu), float(self.logProbFactor),
                      _lib.ptr(p), _lib.stream_ptr())
        if tmp is not None:
            self.net.set_parameters(tmp)
        configs = configs.reshape((1, spc * C) + self.sampleShape)
        return configs, self.net._out(logpsi.reshape(1, -1)), p.reshape(1, -1)

    def acceptance_ratio(self):
        """sampler.py:369-380."""
        c = self.counters.cpu().numpy()
        if c[0] > 0:
            return torch.tensor([c[1] / c[0]], dtype=torch.float64)
        return torch.tensor([0.0], dtype=torch.float64)


class ExactSampler:
    """sampler.py:385-532: the full basis, bit i of k is site i."""

    def __init__(self, net, sampleShape, lDim=2, logProbFactor=0.5):
        if lDim != 2:
            raise NotImplementedError("only lDim = 2 is implemented")
        if isinstance(sampleShape, int):
            sampleShape = (sampleShape,)
        self.psi = net
        self.sampleShape = tuple(sampleShape)
        self.N = int(np.prod(self.sampleShape))
        self.lDim = lDim
        self.logProbFactor = logProbFactor
        _, self.globNumSamples = mpi.distribute_sampling(self.lDim ** self.N)
        k = torch.arange(2 ** self.N, device=global_defs.device(), dtype=torch.int64)
        bits = torch.arange(self.N, device=global_defs.device(), dtype=torch.int64)
        self.basis = ((k[:, None] >> bits[None, :]) & 1).to(torch.int32).reshape(
            (1, 2 ** self.N) + self.sampleShape).contiguous()
        self.psi(self.basis)
        self.lastNorm = 0.

    def sample(self, parameters=None, numSamples=None, multipleOf=None):
        if parameters is not None:
            tmp = self.psi.get_parameters()
            self.psi.set_parameters(parameters)
        flat = self.basis.reshape(-1, self.N)
        logpsi = self.psi.logpsi_complex(flat)
        if parameters is not None:
            self.psi.set_parameters(tmp)
        n = flat.shape[0]
        p = torch.empty(n, dtype=torch.float64, device=flat.device)
        nrm = _lib.ctypes.c_double(0.0)
        _lib.call("jv_exact_weights", _lib.ptr(logpsi), n, float(self.lastNorm), float(self.logProbFactor),
                  _lib.ptr(p), _lib.ctypes.byref(nrm), _lib.stream_ptr())
        self.lastNorm += self.logProbFactor * math.log(nrm.value)
        return self.basis, self.psi._out(logpsi.reshape(1, -1)), p.reshape(1, -1)

    def set_number_of_samples(self, N):
        pass

    def get_last_number_of_samples(self):
        return np.inf
