"""SampledObs (stats.py:136-336) on the GPU.

Centring (weighted mean and sqrt(w)(O - mean)), covariances, variances and the
tangent kernel run as HIP kernels; observations keep the reference's leading
device axis.
"""
import torch

from . import _lib


def _as2d(obs):
    o = obs.reshape(obs.shape[1], -1) if obs.dim() >= 2 else obs.reshape(-1, 1)
    if torch.is_complex(o):
        return o.to(torch.complex128).contiguous(), True
    if o.dtype == torch.float32:  # the mixed form's FP32 gradients stay FP32
        return o.contiguous(), False
    return o.to(torch.float64).contiguous(), False


class SampledObs:
    def __init__(self, observations=None, weights=None):
        self._weights = None if weights is None else weights.reshape(-1).to(torch.float64).contiguous()
        self._data = None
        self._mean = None
        if observations is not None and weights is not None:
            self._compute_data_and_mean(observations)

    def _compute_data_and_mean(self, observations):
        """stats.py:197-205."""
        o, cpx = _as2d(observations)
        n, m = o.shape
        self._cpx = cpx
        self._data = torch.empty_like(o)
        if o.dtype == torch.float32:
            # mean accumulated in FP64, centred data rounded to FP32
            self._mean = torch.empty(m, dtype=torch.float64, device=o.device)
            _lib.call("jv_center_f32", _lib.ptr(o), _lib.ptr(self._weights), n, m, 0, _lib.ptr(self._mean),
                      _lib.ptr(self._data), _lib.stream_ptr())
            return
        self._mean = torch.empty(m, dtype=o.dtype, device=o.device)
        _lib.call("jv_center", _lib.ptr(o), _lib.ptr(self._weights), n, m, int(cpx), _lib.ptr(self._mean),
                  _lib.ptr(self._data), _lib.stream_ptr())

    def mean(self):
        return self._mean

    def covar(self, other=None):
        """stats.py:235-245: sum_n conj(d1_n)^T d2_n."""
        other = self if other is None else other
        a, b = self._data, other._data
        cpx = torch.is_complex(a) or torch.is_complex(b)
        if cpx:
            a = a.to(torch.complex128).contiguous()
            b = b.to(torch.complex128).contiguous()
        n, p1 = a.shape
        p2 = b.shape[1]
        C = torch.empty((p1, p2), dtype=a.dtype, device=a.device)
        _lib.call("jv_gemm", p1, p2, n, _lib.ptr(a), 1, p1, int(cpx), _lib.ptr(b), p2, 1, 0, _lib.ptr(C), p2,
                  int(cpx), 1.0, _lib.stream_ptr())
        return C

    def var(self):
        """stats.py:248-252."""
        n, m = self._data.shape
        out = torch.empty(m, dtype=torch.float64, device=self._data.device)
        _lib.call("jv_col_norm2", _lib.ptr(self._data), n, m, int(torch.is_complex(self._data)), _lib.ptr(out),
                  _lib.stream_ptr())
        return out

    def covar_data(self, other=None):
        """stats.py:255-265."""
        other = self if other is None else other
        d = torch.einsum("ni,nj->nij", self._data.conj(), other._data) / self._weights[:, None, None]
        return SampledObs(d.reshape(1, d.shape[0], -1), self._weights)

    def covar_var(self, other=None):
        """stats.py:268-279."""
        other = self if other is None else other
        outer = torch.einsum("ni,nj->nij", self._data.conj(), other._data)
        s = torch.sum(torch.abs(outer) ** 2 / self._weights[:, None, None], dim=0)
        return s - torch.abs(self.covar(other)) ** 2

    def transform(self, nonLinearFun=lambda x: x, linearFun=None):
        """stats.py:282-292."""
        raw = self._data / torch.sqrt(self._weights)[:, None] + self._mean[None, :]
        t = nonLinearFun(raw)
        if linearFun is not None:
            t = t @ linearFun.T
        return SampledObs(t.reshape(1, t.shape[0], -1), self._weights)

    def subset(self, start=None, end=None, step=None):
        """stats.py:310-329 (jv_subset: renormalised weights, re-centred data)."""
        n, m = self._data.shape
        rows = range(n)[slice(start, end, step)]
        if len(rows) == 0:
            raise ValueError("subset selects no samples")
        cnt = len(rows)
        new = SampledObs()
        new._weights = torch.empty(cnt, dtype=torch.float64, device=self._data.device)
        new._mean = torch.empty_like(self._mean)
        new._data = torch.empty((cnt, m), dtype=self._data.dtype, device=self._data.device)
        new._cpx = torch.is_complex(self._data)
        _lib.call("jv_subset", _lib.ptr(self._data), _lib.ptr(self._weights), _lib.ptr(self._mean), n, m,
                  int(new._cpx), rows.start, rows.step, cnt, _lib.ptr(new._data), _lib.ptr(new._weights),
                  _lib.ptr(new._mean), _lib.stream_ptr())
        return new

    def tangent_kernel(self):
        """stats.py:332-336: D D^dagger."""
        n, p = self._data.shape
        T = torch.empty((n, n), dtype=self._data.dtype, device=self._data.device)
        _lib.call("jv_gram_nn", _lib.ptr(self._data), n, p, int(torch.is_complex(self._data)), _lib.ptr(T),
                  _lib.stream_ptr())
        return T
