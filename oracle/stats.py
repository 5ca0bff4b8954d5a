"""SampledObs and the single-process reductions of mpi_wrapper.

Oracle (test infrastructure only).  Follows stats.py:136-336 and
mpi_wrapper.py:114-243 with one process and the device axis dropped:
observations are (N, ...) and weights (N,).
"""
import numpy as np


def global_mean(data, p):
    """mpi_wrapper.py:150-174."""
    return np.tensordot(p, data, axes=(0, 0))


def global_variance(data, p):
    """mpi_wrapper.py:177-217: sum_j p_j |X_j - <X>|^2."""
    m = global_mean(data, p)
    d = data - m[None, ...]
    return np.einsum("ij,ij,i->j", np.conj(d), d, p)


class SampledObs:
    def __init__(self, observations=None, weights=None):
        self._weights = None if weights is None else np.asarray(weights)
        self._data = None
        self._mean = None
        if observations is not None and weights is not None:
            obs = np.asarray(observations)
            if obs.ndim == 1:
                obs = obs[:, None]
            # stats.py:197-205
            self._mean = np.tensordot(self._weights, obs, axes=(0, 0))
            self._data = np.sqrt(self._weights)[:, None] * (obs - self._mean[None, :])

    def mean(self):
        return self._mean

    def covar(self, other=None):
        """stats.py:235-245: sum_n conj(d1_n)^T d2_n."""
        other = self if other is None else other
        return np.conj(self._data).T @ other._data

    def var(self):
        """stats.py:248-252."""
        return np.sum(np.abs(self._data) ** 2, axis=0)

    def covar_data(self, other=None):
        """stats.py:255-265: per-sample outer(conj d1, d2) / w."""
        other = self if other is None else other
        d = np.einsum("ni,nj->nij", np.conj(self._data), other._data) / self._weights[:, None, None]
        return SampledObs(d.reshape(len(d), -1), self._weights)

    def covar_var(self, other=None):
        """stats.py:268-279."""
        other = self if other is None else other
        outer = np.einsum("ni,nj->nij", np.conj(self._data), other._data)
        s = np.sum(np.abs(outer) ** 2 / self._weights[:, None, None], axis=0)
        return s - np.abs(self.covar(other)) ** 2

    def transform(self, nonLinearFun=lambda x: x, linearFun=None):
        """stats.py:282-292: recover per-sample values, apply f, then V."""
        raw = self._data / np.sqrt(self._weights)[:, None] + self._mean[None, :]
        t = nonLinearFun(raw)
        if linearFun is not None:
            t = t @ np.asarray(linearFun).T
        return SampledObs(t, self._weights)

    def subset(self, start=None, end=None, step=None):
        """stats.py:310-329."""
        sl = slice(start, end, step)
        w = self._weights[sl]
        nrm = w.sum()
        new = SampledObs()
        new._weights = w / nrm
        d = self._data[sl] / np.sqrt(nrm)
        new._mean = np.tensordot(np.sqrt(new._weights), d, axes=(0, 0)) + self._mean
        new._data = d + np.sqrt(new._weights)[:, None] * (self._mean - new._mean)[None, :]
        return new

    def tangent_kernel(self):
        """stats.py:332-336."""
        return self._data @ np.conj(self._data).T
