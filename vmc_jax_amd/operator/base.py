"""Operator base class (base.py:20-317) on the GPU.

``get_O_loc`` runs the fused HIP local estimator: s' are enumerated in-kernel
and never materialised, and log psi(s') comes from an incremental (RBM) or
receptive-field-limited (CNN) re-evaluation.  ``get_s_primes`` still
materialises (sp, matEl) with the reference's nonzero compaction for callers
that want them.
"""
import abc

import torch

from .. import _lib


class Operator(metaclass=abc.ABCMeta):
    def __init__(self, ElocBatchSize=-1):
        self.compiled = False
        self.ElocBatchSize = ElocBatchSize

    @abc.abstractmethod
    def compile(self):
        """Build the device operator table."""

    def get_s_primes(self, s, *args):
        """base.py:116-160: (sp (1, N*M, *shape), matEl (1, N, M)) with
        M = max nonzero count, padded with the last string's s' and m = 0."""
        if not self.compiled:
            self.compile()
        pref = self.prefactors(*args)
        shape = tuple(s.shape[2:])
        flat = s.reshape(-1, int(torch.tensor(shape).prod()) if shape else 1).to(torch.int32).contiguous()
        n, L = flat.shape
        nops = self.table.num_ops
        sp = torch.empty((n, nops, L), dtype=torch.int32, device=flat.device)
        me = torch.empty((n, nops), dtype=torch.complex128, device=flat.device)
        cnt = torch.empty(n, dtype=torch.int32, device=flat.device)
        _lib.call("jv_get_s_primes", self.table, _lib.ptr(pref), _lib.ptr(flat), n, L, _lib.ptr(sp), _lib.ptr(me),
                  _lib.ptr(cnt), _lib.stream_ptr())
        M = int(cnt.max().item()) if n > 0 else 0  # host sync, as base.py:157-158
        self.numNonzero = cnt.reshape(1, -1)
        self.matEl = me[:, :M].contiguous().reshape(1, n, M)
        self.sp = sp[:, :M].contiguous().reshape((1, n * M) + shape)
        return self.sp, self.matEl

    def get_O_loc_unbatched(self, logPsiS, logPsiSP):
        """base.py:194-212 (uses the matEl of the last get_s_primes call)."""
        m = self.matEl
        lp = logPsiSP.reshape(m.shape).to(torch.complex128)
        return torch.sum(m * torch.exp(lp - logPsiS.to(torch.complex128)[..., None]), dim=-1)

    def get_O_loc(self, samples, psi, logPsiS=None, *args):
        """base.py:166-192, fused on the device."""
        if not self.compiled:
            self.compile()
        if logPsiS is None:
            logPsiS = psi(samples)
        if self.ElocBatchSize > 0:
            return self.get_O_loc_batched(samples, psi, logPsiS, self.ElocBatchSize, *args)
        return self._eloc(samples, psi, logPsiS, *args)

    def _eloc(self, samples, psi, logPsiS, *args):
        pref = self.prefactors(*args)
        flat = samples.reshape(-1, psi.desc.num_sites).to(torch.int32).contiguous()
        n = flat.shape[0]
        lps = logPsiS.reshape(-1).to(torch.complex128).contiguous()
        out = torch.empty(n, dtype=torch.complex128, device=flat.device)
        if psi._sym is not None:
            # symmetrised networks: s' materialised per chunk, log psi(s') over the orbit
            _lib.call("jv_sym_eloc", psi.desc, _lib.ptr(psi.params), psi._sym[0], self.table, _lib.ptr(pref),
                      _lib.ptr(flat), _lib.ptr(lps), n, _lib.ptr(out), _lib.stream_ptr())
        else:
            _lib.call("jv_eloc", psi.desc, _lib.ptr(psi.params), self.table, _lib.ptr(pref), _lib.ptr(flat),
                      _lib.ptr(lps), n, _lib.ptr(out), _lib.stream_ptr())
        return out.reshape(1, n)

    def get_O_loc_batched(self, samples, psi, logPsiS, batchSize, *args):
        """base.py:214-278: same sums, evaluated in sample batches."""
        n = samples.shape[1]
        outs = []
        for b in range(0, n, batchSize):
            outs.append(self._eloc(samples[:, b:b + batchSize], psi, logPsiS[:, b:b + batchSize], *args))
        return torch.cat(outs, dim=1)
