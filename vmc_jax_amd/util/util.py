"""Drivers (util.py:91-212): ``measure`` and ``ground_state_search``."""
import time

import numpy as np
import torch

from .. import mpi_wrapper as mpi
from . import stepper as jVMCstepper


def measure(observables, psi, sampler, numSamples=None):
    """util.py:91-157."""
    sampleConfigs, sampleLogPsi, p = sampler.sample(numSamples=numSamples)
    result = {}
    for name, ops in observables.items():
        if not isinstance(ops, (list, tuple)):
            ops = [ops]
        means, variances, errors = [], [], []
        for op in ops:
            args = ()
            if isinstance(op, tuple):
                args = tuple(op[1:])
                op = op[0]
            Oloc = op.get_O_loc(sampleConfigs, psi, sampleLogPsi, *args)
            m = mpi.global_mean(Oloc[..., None], p)[0]
            v = mpi.global_variance(Oloc[..., None], p)[0]
            means.append(complex(m.item()))
            variances.append(float(v.item()))
            errors.append(np.sqrt(variances[-1]) / np.sqrt(sampler.get_last_number_of_samples()))
        result[name] = {"mean": np.real(np.array(means)), "variance": np.real(np.array(variances)),
                        "MC_error": np.real(np.array(errors))}
    return result


def ground_state_search(psi, ham, tdvpEquation, sampler, numSteps=200, varianceTol=1e-10, stepSize=1e-2,
                        observables=None, outp=None):
    """util.py:160-212."""
    delta = getattr(tdvpEquation, "diagonalShift", 0.0)
    stepper = jVMCstepper.Euler(timeStep=stepSize)
    n = 0
    varE = 1.0
    while n < numSteps and varE > varianceTol:
        dp, _ = stepper.step(0, tdvpEquation, psi.get_parameters(), hamiltonian=ham, psi=psi, numSamples=None,
                             outp=outp)
        psi.set_parameters(dp)
        n += 1
        varE = tdvpEquation.get_energy_variance()
        if hasattr(tdvpEquation, "set_diagonal_shift"):
            delta = 0.95 * delta
            tdvpEquation.set_diagonal_shift(delta)
    return n
