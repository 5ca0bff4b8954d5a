"""TDVP and MinSR solves, ODE steppers and the drivers.

Oracle (test infrastructure only).  Follows
* TDVP          util/tdvp.py:14-326 (equation, eigh, SNR, regularised pinv loop,
                metadata incl. tdvp_error)
* MinSR         util/minsr.py:53-165
* pinv          ``jnp.linalg.pinv(T, rcond, hermitian=True)`` used at
                minsr.py:61,73: eigendecomposition, singular values |lambda|,
                keep |lambda| > rcond * max|lambda|, invert the rest.
* Euler, Heun, AdaptiveHeun   util/stepper.py:6-182
* measure, ground_state_search  util/util.py:91-212
"""
import numpy as np

from .sampler import ExactSampler
from .stats import SampledObs, global_mean, global_variance


def realFun(x):
    return np.real(x)


def imagFun(x):
    return 0.5 * (x - np.conj(x))


def pinv_hermitian(T, rcond):
    ev, V = np.linalg.eigh(T)
    s = np.abs(ev)
    cutoff = rcond * s.max()
    inv = np.where(s > cutoff, 1.0 / np.where(s > cutoff, ev, 1.0), 0.0)
    return (V * inv[None, :]) @ np.conj(V).T


class TDVP:
    def __init__(self, sampler, snrTol=2, pinvTol=1e-14, pinvCutoff=1e-8, makeReal='imag',
                 rhsPrefactor=1.j, diagonalShift=0., crossValidation=False):
        self.sampler = sampler
        self.snrTol = snrTol
        self.pinvTol = pinvTol
        self.pinvCutoff = pinvCutoff
        self.diagonalShift = diagonalShift
        self.rhsPrefactor = rhsPrefactor
        self.crossValidation = crossValidation
        self.makeReal = imagFun if makeReal == 'imag' else realFun
        self.metaData = None

    def set_diagonal_shift(self, delta):
        self.diagonalShift = delta

    def trafo_helper(self, x):
        return self.makeReal((-self.rhsPrefactor) * x)

    def get_tdvp_equation(self, Eloc, gradients):
        """tdvp.py:134-148."""
        self.ElocMean = Eloc.mean()[0]
        self.ElocVar = Eloc.var()[0]
        self.F0 = (-self.rhsPrefactor) * gradients.covar(Eloc).ravel()
        F = self.makeReal(self.F0)
        self.S0 = gradients.covar()
        S = self.makeReal(self.S0)
        if self.diagonalShift > 1e-10:
            S = S + np.diag(self.diagonalShift * np.diag(S))
        return S, F

    def _get_tdvp_error(self, update):
        """tdvp.py:102-104."""
        return np.abs(1. + np.real(update.dot(self.S0.dot(update)) - 2. * np.real(update.dot(self.F0))) / self.ElocVar0)

    def solve(self, Eloc, gradients):
        """tdvp.py:183-213."""
        self.S, F = self.get_tdvp_equation(Eloc, gradients)
        self.ev, self.V = np.linalg.eigh(self.S)
        self.VtF = np.conj(self.V).T @ F
        # _get_snr, tdvp.py:173-181
        EO = gradients.covar_data(Eloc).transform(linearFun=np.conj(self.V).T,
                                                 nonLinearFun=self.trafo_helper)
        self.rhoVar = EO.var().ravel()
        self.snr = np.sqrt(np.abs(self.sampler.globNumSamples * (np.conj(self.VtF) * self.VtF) / self.rhoVar)).ravel()
        self.invEv = np.where(np.abs(self.ev / self.ev[-1]) > 1e-14, 1. / self.ev, 0.)
        residual = 1.0
        cutoff = 1e-2
        F_norm = np.linalg.norm(F)
        pinvEv = self.invEv
        while residual > self.pinvTol and cutoff > self.pinvCutoff:
            cutoff *= 0.8
            regularizer = 1. / (1. + (max(cutoff, self.pinvCutoff) / np.abs(self.ev / self.ev[-1])) ** 6)
            if not isinstance(self.sampler, ExactSampler):
                regularizer = regularizer * (1. / (1. + (self.snrTol / self.snr) ** 6))
            pinvEv = self.invEv * regularizer
            residual = np.linalg.norm((pinvEv * self.ev - np.ones_like(pinvEv)) * self.VtF) / F_norm
        update = np.real(self.V @ (pinvEv * self.VtF))
        return update, residual, max(cutoff, self.pinvCutoff)

    def __call__(self, netParameters, t, *, psi, hamiltonian, **rhsArgs):
        """tdvp.py:219-326."""
        tmp = psi.get_parameters()
        psi.set_parameters(netParameters)
        numSamples = rhsArgs.get("numSamples", None)
        configs, logPsi, p = self.sampler.sample(numSamples=numSamples)
        Eloc = SampledObs(hamiltonian.get_O_loc(configs, psi, logPsi, t), p)
        grads = SampledObs(psi.gradients(configs), p)
        update, res, cut = self.solve(Eloc, grads)
        psi.set_parameters(tmp)
        if rhsArgs.get("intStep", None) == 0:
            self.ElocMean0 = self.ElocMean
            self.ElocVar0 = self.ElocVar
            self.metaData = {"tdvp_error": self._get_tdvp_error(update), "tdvp_residual": res,
                             "pinv_cutoff": cut, "SNR": self.snr, "spectrum": self.ev}
            if self.crossValidation:
                update, res = self.cross_validate(Eloc, grads)
        return update

    def cross_validate(self, Eloc, grads):
        """tdvp.py:306-324: solve on the even samples, validate on the odd ones.

        The tdvp_error of the even-sample update is taken with S0, F0 of the
        odd samples (the last get_tdvp_equation call) and the full sample's
        ElocVar0; the residual ratio divides by the full solve's residual.  The
        full solve is repeated, so the returned update is the full one."""
        E1, G1 = Eloc.subset(start=0, step=2), grads.subset(start=0, step=2)
        E2, G2 = Eloc.subset(start=1, step=2), grads.subset(start=1, step=2)
        update_1, _, _ = self.solve(E1, G1)
        S2, F2 = self.get_tdvp_equation(E2, G2)
        validation_tdvpErr = self._get_tdvp_error(update_1)
        update, res, _ = self.solve(Eloc, grads)
        validation_residual = (np.linalg.norm(S2.dot(update_1) - F2) / np.linalg.norm(F2)) / res
        self.crossValidationFactor_residual = validation_residual
        self.crossValidationFactor_tdvpErr = validation_tdvpErr / self.metaData["tdvp_error"]
        self.metaData["tdvp_residual_cross_validation_ratio"] = self.crossValidationFactor_residual
        self.metaData["tdvp_error_cross_validation_ratio"] = self.crossValidationFactor_tdvpErr
        self.S, _ = self.get_tdvp_equation(Eloc, grads)
        return update, res

    def get_energy_variance(self):
        return self.ElocVar0

    def get_energy_mean(self):
        return np.real(self.ElocMean0)


class MinSR:
    def __init__(self, sampler, pinvTol=1e-14, diagonalShift=0.):
        self.sampler = sampler
        self.pinvTol = pinvTol
        self.diagonalShift = diagonalShift
        self.metaData = None

    def solve(self, eloc, gradients, holomorphic):
        """minsr.py:53-80."""
        if holomorphic:
            T = gradients.tangent_kernel()
            Tinv = pinv_hermitian(T, self.pinvTol)
            e = eloc._data.reshape(-1)
            return -np.conj(gradients._data).T @ (Tinv @ e)
        G = gradients._data
        G = np.concatenate([np.real(G), np.imag(G)], axis=0)
        T = G @ G.T + self.diagonalShift * np.eye(G.shape[0])
        Tinv = pinv_hermitian(T, self.pinvTol)
        e = eloc._data.reshape(-1)
        e = np.concatenate([np.real(e), np.imag(e)])
        return -G.T @ (Tinv @ e)

    def __call__(self, netParameters, t, *, psi, hamiltonian, **rhsArgs):
        tmp = psi.get_parameters()
        psi.set_parameters(netParameters)
        numSamples = rhsArgs.get("numSamples", None)
        configs, logPsi, p = self.sampler.sample(numSamples=numSamples)
        Eloc = SampledObs(hamiltonian.get_O_loc(configs, psi, logPsi, t), p)
        grads = SampledObs(psi.gradients(configs), p)
        update = self.solve(Eloc, grads, psi.holomorphic)
        psi.set_parameters(tmp)
        if rhsArgs.get("intStep", None) == 0:
            self.ElocMean0 = Eloc.mean()[0]
            self.ElocVar0 = Eloc.var()[0]
            self.metaData = {}
        return update

    def get_energy_variance(self):
        return self.ElocVar0

    def get_energy_mean(self):
        return np.real(self.ElocMean0)


class Euler:
    """stepper.py:6-40."""

    def __init__(self, timeStep=1e-3):
        self.dt = timeStep

    def step(self, t, f, yInitial, **rhsArgs):
        dy = f(yInitial, t, **rhsArgs, intStep=0)
        return yInitial + self.dt * dy, self.dt


class Heun:
    """stepper.py:45-93."""

    def __init__(self, timeStep=1e-3):
        self.dt = timeStep

    def step(self, t, f, yInitial, **rhsArgs):
        dt = rhsArgs.get("dt", self.dt)
        k0 = f(yInitial, t, **rhsArgs, intStep=0)
        y = yInitial + dt * k0
        k1 = f(y, t + dt, **rhsArgs, intStep=1)
        return yInitial + 0.5 * dt * (k0 + k1), dt


class AdaptiveHeun:
    """stepper.py:98-182."""

    def __init__(self, timeStep=1e-3, tol=1e-8, maxStep=1):
        self.dt = timeStep
        self.tolerance = tol
        self.maxStep = maxStep

    def step(self, t, f, y, normFunction=np.linalg.norm, **rhsArgs):
        fe = 0.5
        dt = self.dt
        yInitial = y.copy()
        while fe < 1.:
            y = yInitial.copy()
            k0 = f(y, t, **rhsArgs, intStep=0)
            y = y + dt * k0
            k1 = f(y, t + dt, **rhsArgs, intStep=1)
            dy0 = 0.5 * dt * (k0 + k1)
            y = y - 0.5 * dt * k0
            k10 = f(y, t + 0.5 * dt, **rhsArgs, intStep=2)
            dy1 = 0.25 * dt * (k0 + k10)
            y = yInitial + dy1
            k01 = f(y, t + 0.5 * dt, **rhsArgs, intStep=3)
            y = y + 0.5 * dt * k01
            k11 = f(y, t + dt, **rhsArgs, intStep=4)
            dy1 = dy1 + 0.25 * dt * (k01 + k11)
            updateDiff = normFunction(dy1 - dy0)
            fe = self.tolerance / updateDiff
            tmp = 0.2 if 0.2 > 0.9 * fe ** 0.33333 else 0.9 * fe ** 0.33333
            if tmp > 2.:
                tmp = 2.
            realDt = dt
            dt *= tmp
            if dt > self.maxStep:
                dt = self.maxStep
        self.dt = dt
        return yInitial + dy1, realDt


def measure(observables, psi, sampler, numSamples=None):
    """util.py:91-157."""
    configs, logPsi, p = sampler.sample(numSamples=numSamples)
    result = {}
    for name, ops in observables.items():
        if not isinstance(ops, (list, tuple)):
            ops = [ops]
        means, vars_, errs = [], [], []
        for op in ops:
            args = ()
            if isinstance(op, tuple):
                args = tuple(op[1:])
                op = op[0]
            Oloc = op.get_O_loc(configs, psi, logPsi, *args)
            means.append(global_mean(Oloc[:, None], p)[0])
            vars_.append(global_variance(Oloc[:, None], p)[0])
            errs.append(np.sqrt(vars_[-1]) / np.sqrt(sampler.get_last_number_of_samples()))
        result[name] = {"mean": np.real(np.array(means)), "variance": np.real(np.array(vars_)),
                        "MC_error": np.real(np.array(errs))}
    return result


def ground_state_search(psi, ham, tdvpEquation, sampler, numSteps=200, varianceTol=1e-10, stepSize=1e-2):
    """util.py:160-212 (without observables / output)."""
    delta = getattr(tdvpEquation, "diagonalShift", 0.0)
    stepper = Euler(timeStep=stepSize)
    n = 0
    varE = 1.0
    while n < numSteps and varE > varianceTol:
        dp, _ = stepper.step(0, tdvpEquation, psi.get_parameters(), hamiltonian=ham, psi=psi, numSamples=None)
        psi.set_parameters(dp)
        n += 1
        varE = tdvpEquation.get_energy_variance()
        if hasattr(tdvpEquation, "set_diagonal_shift"):
            delta = 0.95 * delta
            tdvpEquation.set_diagonal_shift(delta)
    return n
