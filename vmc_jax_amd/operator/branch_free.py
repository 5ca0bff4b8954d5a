"""Branch-free operator strings (branch_free.py:17-487) on the GPU.

The local-operator tables, ``OpStr`` / ``scal_opstr`` algebra and
``BranchFreeOperator.compile`` padding rules follow the reference; the compiled
tables are uploaded once and the s' enumeration runs inside the HIP kernels.
Prefactor functions are evaluated on the host at every call, as the reference's
``arg_fun`` does (branch_free.py:416-438).
"""
import copy
import functools

import numpy as np
import torch

from .. import _lib
from .. import global_defs
from .base import Operator

opDtype = global_defs.tCpx


class LocalOp(dict):
    """branch_free.py:279-308."""

    def __init__(self, **kwargs):
        for k in kwargs.keys():
            self[k] = kwargs[k]

    def __mul__(self, other):
        if isinstance(other, dict):
            return OpStr(self, LocalOp(**other))
        if isinstance(other, OpStr):
            return OpStr(self, *other)
        return OpStr(self, other)

    def __rmul__(self, other):
        return other * OpStr(self)


def Id(idx=0, lDim=2):
    return LocalOp(idx=idx, map=np.array([j for j in range(lDim)], dtype=np.int32),
                   matEls=np.array([1. for j in range(lDim)], dtype=opDtype), diag=True)


def Sx(idx):
    return LocalOp(idx=idx, map=np.array([1, 0], dtype=np.int32), matEls=np.array([1.0, 1.0], dtype=opDtype),
                   diag=False)


def Sy(idx):
    return LocalOp(idx=idx, map=np.array([1, 0], dtype=np.int32), matEls=np.array([1.j, -1.j], dtype=opDtype),
                   diag=False)


def Sz(idx):
    return LocalOp(idx=idx, map=np.array([0, 1], dtype=np.int32), matEls=np.array([-1.0, 1.0], dtype=opDtype),
                   diag=True)


def Sp(idx):
    return LocalOp(idx=idx, map=np.array([1, 0], dtype=np.int32), matEls=np.array([1.0, 0.0], dtype=opDtype),
                   diag=False)


def Sm(idx):
    return LocalOp(idx=idx, map=np.array([0, 0], dtype=np.int32), matEls=np.array([0.0, 1.0], dtype=opDtype),
                   diag=False)


def number(idx):
    """Fermionic number operator c^dagger c (branch_free.py:127-146)."""
    return LocalOp(idx=idx, map=np.array([0, 1], dtype=np.int32), matEls=np.array([0., 1.], dtype=opDtype),
                   diag=True, fermionic=False)


def creation(idx):
    """Fermionic creation operator c^dagger (branch_free.py:149-168)."""
    return LocalOp(idx=idx, map=np.array([1, 0], dtype=np.int32), matEls=np.array([1., 0.], dtype=opDtype),
                   diag=False, fermionic=True)


def annihilation(idx):
    """Fermionic annihilation operator c (branch_free.py:171-191)."""
    return LocalOp(idx=idx, map=np.array([1, 0], dtype=np.int32), matEls=np.array([0., 1.], dtype=opDtype),
                   diag=False, fermionic=True)


def _id_prefactor(*args, val=1.0, **kwargs):
    return val


def _prod_fun(*args, f1, f2, **kwargs):
    return f1(*args) * f2(*args)


class OpStr(tuple):
    """branch_free.py:203-255."""

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
    """branch_free.py:258-276."""
    if not isinstance(op, (tuple, OpStr)):
        raise RuntimeError("Can add prefactors only to OpStr or tuple objects.")
    if not isinstance(op, OpStr):
        op = OpStr(*op)
    return a * op


class BranchFreeOperator(Operator):
    """branch_free.py:311-487."""

    def __init__(self, lDim=2, **kwargs):
        self.ops = []
        self.lDim = lDim
        super().__init__(**kwargs)

    def add(self, opDescr):
        if not isinstance(opDescr, tuple):
            opDescr = (opDescr,)
        self.ops.append(opDescr)
        self.compiled = False

    def __iadd__(self, opDescr):
        self.add(opDescr)
        return self

    def compile(self):
        """branch_free.py:348-441: pad to maxOpStrLength with Id(0), factors
        stored right to left, diagonal strings listed."""
        if self.lDim != 2:
            raise NotImplementedError("only lDim = 2 is implemented on the device")
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
        # upper bound on the sites one string changes: factors whose map is not
        # the identity (the kernels pick single-flip code paths when it is 1)
        ident = np.arange(self.lDim)
        self.max_sites_changed = int(max([sum(int(not np.array_equal(np.asarray(f), ident)) for f in row)
                                          for row in mp] + [0]))
        # host tables (numpy); uploaded to the device on first use
        self.idxH = np.array(idx, dtype=np.int32)
        self.mapH = np.array(mp, dtype=np.int32)
        self.matElsH = np.array(me, dtype=np.complex128)
        self.diagH = np.array(diag, dtype=np.int32)
        self.fermionicH = np.array(fermi, dtype=np.int32)
        self.prefactor = pref
        self.maxOpStrLength = maxLen
        self._table = None
        self._pref_cache = None
        self.compiled = True
        return self

    @property
    def table(self):
        """Device operator table (jv_optable) of the compiled strings."""
        if not self.compiled:
            self.compile()
        if self._table is None:
            dev = global_defs.device()
            self.idxC = torch.as_tensor(self.idxH, device=dev).contiguous()
            self.mapC = torch.as_tensor(self.mapH, device=dev).contiguous()
            self.matElsC = torch.as_tensor(self.matElsH, device=dev).contiguous()
            dg = self.diagH if len(self.diagH) else np.zeros(1, np.int32)
            self.diagC = torch.as_tensor(dg, device=dev).contiguous()
            t = _lib.JvOpTable()
            t.num_ops = len(self.ops)
            t.max_len = self.maxOpStrLength
            t.l_dim = 2
            t.num_diag = len(self.diagH)
            t.max_sites_changed = self.max_sites_changed
            t.idx = self.idxC.data_ptr()
            t.map = self.mapC.data_ptr()
            t.matel = self.matElsC.data_ptr()
            t.diag = self.diagC.data_ptr()
            # the sign mask is passed only when a string has a fermionic factor
            if self.fermionicH.any():
                self.fermiC = torch.as_tensor(self.fermionicH, device=dev).contiguous()
                t.fermi = self.fermiC.data_ptr()
            else:
                t.fermi = None
            self._table = t
        return self._table

    def prefactors(self, *args):
        """arg_fun (branch_free.py:416-438) -> device complex [numOps]."""
        if not self.compiled:
            self.compile()
        key = tuple(args)
        if self._pref_cache is not None and self._pref_cache[0] == key:
            return self._pref_cache[1]
        res = np.ones(len(self.ops), dtype=np.complex128)
        for i, f in self.prefactor:
            res[i] = f(*args)
        d = torch.as_tensor(res, device=global_defs.device()).contiguous()
        try:
            hash(key)
            self._pref_cache = (key, d)
        except TypeError:
            self._pref_cache = None
        return d
