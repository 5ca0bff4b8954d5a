"""Metropolis and exact samplers.

Oracle (test infrastructure only).  Follows
* ``distribute_sampling``  mpi_wrapper.py:53-95 (one process, one device)
* ``MCSampler``            sampler.py:103-380: walkers persist between calls;
                           every ``sample()`` re-initialises logAccProb and the
                           counters (sampler.py:360-367), thermalises
                           ``thermalizationSweeps * sweepSteps`` steps
                           (sampler.py:309-310), then emits the walkers after
                           each of ``samplesPerChain`` sweeps with sample index
                           t*C + c (sampler.py:313-323); log psi of the samples
                           is re-evaluated (sampler.py:293-296) and
                           p = exp((1/logProbFactor - mu) Re log psi), normalised
                           (sampler.py:234-235).
* ``propose_spin_flip``    sampler.py:15-19 (one uniformly drawn site flipped)
* ``ExactSampler``         sampler.py:385-532 (bit i of k is site i; p from
                           |psi|^(1/logProbFactor) with a running lastNorm)

The proposal stream is Philox (see ``oracle.philox``) instead of JAX threefry.
"""
import math

import numpy as np

from .philox import proposal_stream


def distribute_sampling(numSamples, localDevices=None, numChainsPerDevice=1):
    """mpi_wrapper.py:53-95 with commSize = 1, rank = 0."""
    if localDevices is None:
        return numSamples, numSamples
    perProc = localDevices * numChainsPerDevice
    spc = (numSamples + perProc - 1) // perProc
    return int(spc), int(spc * perProc)


class MCSampler:
    def __init__(self, psi, sampleShape, seed, numChains=1, numSamples=100,
                 thermalizationSweeps=10, sweepSteps=10, initState=None, mu=2,
                 logProbFactor=0.5):
        if mu < 0 or mu > 2:
            raise ValueError("mu must be in the range [0, 2]")
        self.psi = psi
        self.sampleShape = tuple(sampleShape)
        self.L = int(np.prod(self.sampleShape))
        self.seed = int(seed)
        self.numChains = int(numChains)
        self.numSamples = numSamples
        self.thermalizationSweeps = thermalizationSweeps
        self.sweepSteps = sweepSteps
        self.mu = mu
        self.logProbFactor = logProbFactor
        init = np.zeros(self.L, np.int32) if initState is None else np.asarray(initState, np.int32).ravel()
        self.states = np.repeat(init[None, :], self.numChains, axis=0)
        self.callIndex = 0
        self.numProposed = 0
        self.numAccepted = 0
        self.globNumSamples = 0

    def _logacc(self, s):
        return self.mu * np.real(self.psi(s.reshape((-1,) + self.sampleShape)))

    def sample(self, numSamples=None, multipleOf=1):
        if numSamples is None:
            numSamples = self.numSamples
        spc, self.globNumSamples = distribute_sampling(
            numSamples, localDevices=1, numChainsPerDevice=int(np.lcm(self.numChains, multipleOf)))
        C = self.numChains
        states = self.states.copy()
        logAcc = self._logacc(states)
        self.numProposed = 0
        self.numAccepted = 0
        step = 0
        configs = np.zeros((spc, C, self.L), np.int32)

        def sweep(n):
            nonlocal step, states, logAcc
            for _ in range(n):
                site, u = proposal_stream(self.seed, self.callIndex, step, C, self.L)
                prop = states.copy()
                prop[np.arange(C), site] = (prop[np.arange(C), site] + 1) % 2
                newLog = self._logacc(prop)
                P = np.exp(newLog - logAcc)
                acc = u < P
                self.numProposed += C
                self.numAccepted += int(acc.sum())
                states = np.where(acc[:, None], prop, states)
                logAcc = np.where(acc, newLog, logAcc)
                step += 1

        sweep(self.thermalizationSweeps * self.sweepSteps)
        for t in range(spc):
            sweep(self.sweepSteps)
            configs[t] = states
        self.states = states
        self.callIndex += 1
        configs = configs.reshape((spc * C,) + self.sampleShape)
        logPsi = self.psi(configs)
        p = np.exp((1.0 / self.logProbFactor - self.mu) * np.real(logPsi))
        return configs, logPsi, p / p.sum()

    def acceptance_ratio(self):
        if self.numProposed > 0:
            return self.numAccepted / self.numProposed
        return 0.0

    def get_last_number_of_samples(self):
        return self.globNumSamples


class ExactSampler:
    def __init__(self, psi, sampleShape, lDim=2, logProbFactor=0.5):
        if lDim != 2:
            raise NotImplementedError("only lDim = 2")
        if isinstance(sampleShape, int):
            sampleShape = (sampleShape,)
        self.psi = psi
        self.sampleShape = tuple(sampleShape)
        self.N = int(np.prod(self.sampleShape))
        self.logProbFactor = logProbFactor
        k = np.arange(2 ** self.N)
        self.basis = ((k[:, None] >> np.arange(self.N)[None, :]) & 1).astype(np.int32).reshape(
            (-1,) + self.sampleShape)
        self.lastNorm = 0.0
        self.globNumSamples = 2 ** self.N

    def sample(self, parameters=None, numSamples=None, multipleOf=None):
        logPsi = self.psi(self.basis)
        p = np.exp(np.real(logPsi - self.lastNorm) / self.logProbFactor)
        nrm = p.sum()
        p = p / nrm
        self.lastNorm += self.logProbFactor * math.log(nrm)
        return self.basis, logPsi, p

    def get_last_number_of_samples(self):
        return np.inf
