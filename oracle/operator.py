"""Branch-free operator strings, s' enumeration and local estimators.

Oracle (test infrastructure only).  Follows
* local-operator tables   branch_free.py:17-191 (Id, Sx, Sy, Sz, Sp, Sm, and the
                          fermionic number, creation, annihilation)
* OpStr / scal_opstr       branch_free.py:195-276
* compile (padding, diag)  branch_free.py:348-441
* _get_s_primes            branch_free.py:443-487 (factors applied right to left,
                           Jordan-Wigner sign mask of fermionic factors,
                           diagonal strings summed into the first diagonal slot)
* nonzero compaction       base.py:91-114, get_s_primes base.py:116-160
* O_loc                    base.py:162-212
"""
import copy
import functools

import numpy as np


class LocalOp(dict):
    def __init__(self, **kw):
        for k, v in kw.items():
            self[k] = v

    def __mul__(self, other):
        if isinstance(other, OpStr):
            return OpStr(self, *other)
        return OpStr(self, other)

    def __rmul__(self, other):
        return other * OpStr(self)


def Id(idx=0, lDim=2):
    return LocalOp(idx=idx, map=np.arange(lDim, dtype=np.int32),
                   matEls=np.ones(lDim, dtype=np.complex128), diag=True)


def Sx(idx):
    return LocalOp(idx=idx, map=np.array([1, 0], np.int32),
                   matEls=np.array([1.0, 1.0], np.complex128), diag=False)


def Sy(idx):
    return LocalOp(idx=idx, map=np.array([1, 0], np.int32),
                   matEls=np.array([1.j, -1.j], np.complex128), diag=False)


def Sz(idx):
    return LocalOp(idx=idx, map=np.array([0, 1], np.int32),
                   matEls=np.array([-1.0, 1.0], np.complex128), diag=True)


def Sp(idx):
    return LocalOp(idx=idx, map=np.array([1, 0], np.int32),
                   matEls=np.array([1.0, 0.0], np.complex128), diag=False)


def Sm(idx):
    return LocalOp(idx=idx, map=np.array([0, 0], np.int32),
                   matEls=np.array([0.0, 1.0], np.complex128), diag=False)


def number(idx):
    """branch_free.py:127-146."""
    return LocalOp(idx=idx, map=np.array([0, 1], np.int32), matEls=np.array([0., 1.], np.complex128), diag=True,
                   fermionic=False)


def creation(idx):
    """branch_free.py:149-168."""
    return LocalOp(idx=idx, map=np.array([1, 0], np.int32), matEls=np.array([1., 0.], np.complex128), diag=False,
                   fermionic=True)


def annihilation(idx):
    """branch_free.py:171-191."""
    return LocalOp(idx=idx, map=np.array([1, 0], np.int32), matEls=np.array([0., 1.], np.complex128), diag=False,
                   fermionic=True)


def _id_prefactor(*args, val=1.0, **kwargs):
    return val


def _prod_fun(*args, f1, f2, **kwargs):
    return f1(*args) * f2(*args)


class OpStr(tuple):
    """branch_free.py:203-255: callables (prefactors) first, then local ops."""

    def __new__(cls, *args):
        factors, ops = [], []
        for o in args:
            if isinstance(o, (LocalOp, dict)):
                ops.append(o)
            elif callable(o):
                factors.append(o)
            else:
                factors.append(functools.partial(_id_prefactor, val=o))
        while len(factors) > 1:
            factors[0] = functools.partial(_prod_fun, f1=factors[0], f2=factors.pop())
        return super().__new__(cls, tuple(factors + ops))

    def __mul__(self, other):
        if not isinstance(other, (tuple, OpStr)):
            other = OpStr(other)
        if callable(other[0]):
            return OpStr(*(other[0] * self), *(other[1:]))
        return OpStr(*self, *other)

    def __rmul__(self, a):
        if isinstance(a, dict):
            return OpStr(LocalOp(**a), *self)
        newOp = [copy.deepcopy(o) for o in self]
        if not callable(a):
            a = functools.partial(_id_prefactor, val=a)
        if callable(newOp[0]):
            newOp[0] = functools.partial(_prod_fun, f1=a, f2=newOp[0])
        else:
            newOp = [a] + newOp
        return OpStr(*tuple(newOp))


def scal_opstr(a, op):
    if not isinstance(op, (tuple, OpStr)):
        raise RuntimeError("Can add prefactors only to OpStr or tuple objects.")
    if not isinstance(op, OpStr):
        op = OpStr(*op)
    return a * op


class BranchFreeOperator:
    def __init__(self, lDim=2, ElocBatchSize=-1):
        self.ops = []
        self.lDim = lDim
        self.ElocBatchSize = ElocBatchSize
        self.compiled = False

    def add(self, opDescr):
        if not isinstance(opDescr, tuple):
            opDescr = (opDescr,)
        self.ops.append(opDescr)
        self.compiled = False

    def __iadd__(self, opDescr):
        self.add(opDescr)
        return self

    def compile(self):
        """branch_free.py:348-414: pad every string to maxOpStrLength with Id on
        site 0; factors stored in application order (right to left)."""
        maxLen = 0
        for op in self.ops:
            n = len(op) - (1 if callable(op[0]) else 0)
            maxLen = max(maxLen, n)
        IdOp = Id(lDim=self.lDim)
        idx, mp, me, diag, pref, fermi = [], [], [], [], [], []
        for o, op in enumerate(self.ops):
            k0 = 1 if callable(op[0]) else 0
            if k0:
                pref.append((o, op[0]))
            isDiag = True
            ri, rm, re, rf = [], [], [], []
            for k in range(maxLen):
                kRev = len(op) - k - 1
                if kRev >= k0:
                    if not op[kRev]["diag"]:
                        isDiag = False
                    ri.append(op[kRev]["idx"]); rm.append(op[kRev]["map"]); re.append(op[kRev]["matEls"])
                    rf.append(1 if op[kRev].get("fermionic", False) else 0)
                else:
                    ri.append(IdOp["idx"]); rm.append(IdOp["map"]); re.append(IdOp["matEls"])
                    rf.append(0)
            idx.append(ri); mp.append(rm); me.append(re); fermi.append(rf)
            if isDiag:
                diag.append(o)
        self.idxC = np.array(idx, dtype=np.int32)
        self.mapC = np.array(mp, dtype=np.int32)
        self.matElsC = np.array(me, dtype=np.complex128)
        self.diag = np.array(diag, dtype=np.int32)
        self.fermiC = np.array(fermi, dtype=np.int64)
        self.prefactor = pref
        self.maxOpStrLength = maxLen
        self.compiled = True

    def prefactors(self, *args):
        """arg_fun (branch_free.py:416-438): ones, overwritten by f(*args)."""
        if not self.compiled:
            self.compile()
        res = np.ones(len(self.ops), dtype=np.complex128)
        for i, f in self.prefactor:
            res[i] = f(*args)
        return res

    def s_primes_full(self, s, *args):
        """All numOps connected configs before compaction (branch_free.py:443-487)."""
        if not self.compiled:
            self.compile()
        s = np.asarray(s)
        N = len(s)
        shape = s.shape[1:]
        flat = s.reshape(N, -1)
        nOps = len(self.ops)
        sp = np.repeat(flat[:, None, :], nOps, axis=1).copy()  # (N, nOps, L)
        m = np.broadcast_to(self.prefactors(*args), (N, nOps)).copy()
        rows = np.arange(N)[:, None]
        opsr = np.arange(nOps)[None, :]
        L = flat.shape[1]
        # mask[i][j] = 1 for j > i (jnp.tril(ones, -1).T, branch_free.py:453)
        mask = np.tril(np.ones((L, L), dtype=np.int64), -1).T
        for k in range(self.maxOpStrLength):
            site = self.idxC[:, k]  # (nOps,)
            cur = sp[rows, opsr, site[None, :]]  # (N, nOps)
            # branch_free.py:463-466: prod_j ((1-2f)(2f mask[idx, j] + 1 - 2f) c_j + 1 - |c_j|)
            m = m * self.matElsC[opsr, k, cur]
            if self.fermiC[:, k].any():  # the product is 1 for non-fermionic factors
                f = self.fermiC[:, k][None, :, None]  # (1, nOps, 1)
                c = sp.astype(np.int64)
                term = (1 - 2 * f) * (2 * f * mask[site][None, :, :] + (1 - 2 * f)) * c + (1 - np.abs(c))
                m = m * np.prod(term, axis=-1)
            sp[rows, opsr, site[None, :]] = self.mapC[opsr, k, cur]
        if len(self.diag) > 1:
            m[:, self.diag[0]] = np.sum(m[:, self.diag], axis=1)
            m[:, self.diag[1:]] = 0.0
        return sp.reshape((N, nOps) + shape), m

    def get_s_primes(self, s, *args):
        """base.py:116-160: keep |m| > 1e-6 in string order; pad to the maximum
        count over samples with the last string's s' and m = 0."""
        sp, m = self.s_primes_full(s, *args)
        N, nOps = m.shape
        nz = np.abs(m) > 1e-6
        choice = np.where(nz, np.arange(nOps)[None, :], nOps - 1)
        idx = np.sort(choice, axis=1)
        cnt = nz.sum(axis=1)
        M = int(cnt.max()) if N else 0
        idx = idx[:, :M]
        mm = np.take_along_axis(m, idx, axis=1)
        mm[np.arange(M)[None, :] >= cnt[:, None]] = 0.0
        spc = np.take_along_axis(sp.reshape(N, nOps, -1), idx[:, :, None], axis=1)
        self.sp = spc.reshape((N * M,) + s.shape[1:] if hasattr(s, "shape") else (N * M, -1))
        self.matEl = mm
        self.numNonzero = cnt
        return self.sp, self.matEl

    def get_O_loc_unbatched(self, logPsiS, logPsiSP):
        """base.py:162-164, 194-212."""
        m = self.matEl
        return np.sum(m * np.exp(np.asarray(logPsiSP).reshape(m.shape) - np.asarray(logPsiS)[:, None]), axis=1)

    def get_O_loc(self, samples, psi, logPsiS=None, *args):
        """base.py:166-192 (batched and unbatched give the same sums)."""
        samples = np.asarray(samples)
        if logPsiS is None:
            logPsiS = psi(samples)
        sp, _ = self.get_s_primes(samples, *args)
        return self.get_O_loc_unbatched(logPsiS, psi(sp))

    def dense_matrix(self, L, *args):
        """Brute-force dense matrix <s|H|s'> over the 2^L basis (bit i of k is
        site i, sampler.py:442-455).  Used to pin s' enumeration by ED."""
        basis = ((np.arange(2 ** L)[:, None] >> np.arange(L)[None, :]) & 1).astype(np.int32)
        sp, m = self.s_primes_full(basis, *args)
        D = 2 ** L
        Hm = np.zeros((D, D), dtype=np.complex128)
        ks = (sp.reshape(D, -1, L) * (1 << np.arange(L))[None, None, :]).sum(-1)
        for r in range(D):
            np.add.at(Hm[r], ks[r], m[r])
        return Hm
