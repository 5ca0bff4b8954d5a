"""Fermionic operator strings in the oracle (CPU).

Pinned by the reference's golden value (operator_test.py:199-248: the impurity
Hamiltonian on the state stored in tests/data_ref/fermion_ref.txt, copied to
tests/golden/fermion_ref.txt, has energy -9.95314531), by the canonical
anticommutation relations {c_i, c_j^dagger} = delta_ij on every configuration
(operator_test.py:174-197), and by a Jordan-Wigner matrix built with np.kron.
"""
import os

import numpy as np

from oracle import operator as oop
from oracle import nets as onets

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "fermion_ref.txt")


def impurity_hamiltonian(mod, t=-1.0, mu=-2.0, V=4.0, L=4):
    """operator_test.py:214-241 (two flavours on a chain of L sites)."""
    H = mod.BranchFreeOperator()
    fl = 2 * L
    up, do = 0, fl - 1
    for i in range(L):
        H.add(mod.scal_opstr(V, (mod.number(up + i), mod.number(do - i))))
        H.add(mod.scal_opstr(mu, (mod.number(up + i),)))
        H.add(mod.scal_opstr(mu, (mod.number(do - i),)))
        if i == L - 1:
            continue
        H.add(mod.scal_opstr(t, (mod.creation(up + i + 1), mod.annihilation(up + i))))
        H.add(mod.scal_opstr(t, (mod.creation(up + i), mod.annihilation(up + i + 1))))
        H.add(mod.scal_opstr(t, (mod.creation(do - i - 1), mod.annihilation(do - i))))
        H.add(mod.scal_opstr(t, (mod.creation(do - i), mod.annihilation(do - i - 1))))
    return H


class TablePsi:
    """The reference test's Target network (operator_test.py:17-38): log of a
    stored amplitude, the configuration read with site 0 as the most
    significant bit."""

    def __init__(self, amps, L, delta=1e-15):
        self.amps = np.asarray(amps)
        self.L = L
        self.delta = delta

    def index(self, s):
        s = np.asarray(s).reshape(-1, self.L)
        return (s[:, ::-1] * (2 ** np.arange(self.L))).sum(axis=1)

    def __call__(self, s):
        a = self.amps[self.index(s)]
        return np.log(np.abs(a + self.delta)) + 1j * np.angle(a)


def golden_energy(sp_fn=None):
    L = 8
    b = np.loadtxt(GOLD, dtype=np.complex128)
    psi = TablePsi(b, L)
    k = np.arange(2 ** L)
    basis = ((k[:, None] >> np.arange(L)[None, :]) & 1).astype(np.int32)
    lp = psi(basis)
    p = np.exp(2 * np.real(lp))
    p /= p.sum()
    return psi, basis, lp, p


def test_golden_impurity_energy():
    H = impurity_hamiltonian(oop)
    psi, basis, lp, p = golden_energy()
    E = np.sum(p * H.get_O_loc(basis, psi, lp))
    assert np.allclose(E, -9.95314531)


def test_anticommutators_on_every_configuration():
    """{c_i, c_j^dagger} O_loc is 1 (i = j) or 0 (i != j) exactly."""
    L = 5
    rng = np.random.default_rng(0)
    onet = onets.CpxRBM((L,), numHidden=3, bias=True)
    psi = onets.NQS(onet, 0.4 * rng.normal(size=onet.num_real_params()))
    k = np.arange(2 ** L)
    basis = ((k[:, None] >> np.arange(L)[None, :]) & 1).astype(np.int32)
    for i in range(L):
        for j in range(L):
            C = oop.BranchFreeOperator()
            C.add(oop.scal_opstr(1., (oop.creation(j), oop.annihilation(i))))
            C.add(oop.scal_opstr(1., (oop.annihilation(i), oop.creation(j))))
            e = C.get_O_loc(basis, psi, psi(basis))
            np.testing.assert_array_equal(e, np.full(len(basis), 1.0 if i == j else 0.0))


def _jw_dense(L):
    """c_i = Z_{i+1} ... Z_{L-1} a_i on the occupation basis (bit i of the
    index is site i), Z = diag(1, -1): the sign counts occupied sites after i.
    As <s|.|s'> matrices, jVMC's creation(i) (s_i: 0 -> 1 with element 1) is c_i."""
    a = np.array([[0., 1.], [0., 0.]])  # |0><1| on one site: occupation 1 -> 0
    Z = np.diag([1., -1.])
    I2 = np.eye(2)
    cs = []
    for i in range(L):
        # np.kron puts site L-1 first so that bit i of the index is site i
        m = np.ones((1, 1))
        for j in reversed(range(L)):
            m = np.kron(m, a if j == i else (Z if j > i else I2))
        cs.append(m)
    return cs


def test_hopping_matches_jordan_wigner_matrix():
    L = 5
    rng = np.random.default_rng(1)
    cs = _jw_dense(L)
    H = oop.BranchFreeOperator()
    Hd = np.zeros((2 ** L, 2 ** L), dtype=np.complex128)
    for i in range(L):
        for j in range(L):
            t = rng.normal() + 1j * rng.normal()
            H.add(oop.scal_opstr(t, (oop.creation(i), oop.annihilation(j))))
            # strings act on the bra, factors right to left: <s|(A, B)|s'> = (M_B M_A)[s, s']
            # with M_creation(i) = c_i and M_annihilation(j) = c_j^T in this basis
            Hd += t * cs[j].T @ cs[i]
        H.add(oop.scal_opstr(0.3 * (i + 1), (oop.number(i), oop.number((i + 2) % L))))
        Hd += 0.3 * (i + 1) * (cs[i].T @ cs[i]) @ (cs[(i + 2) % L].T @ cs[(i + 2) % L])
    psi = rng.normal(size=2 ** L) + 1j * rng.normal(size=2 ** L)
    k = np.arange(2 ** L)
    basis = ((k[:, None] >> np.arange(L)[None, :]) & 1).astype(np.int32)
    lpfun = lambda s: np.log(psi[(np.asarray(s).reshape(-1, L) * (1 << np.arange(L))).sum(1)])
    e = H.get_O_loc(basis, lpfun, lpfun(basis))
    # O_loc(s) = sum_s' <s|H|s'> psi(s') / psi(s)
    np.testing.assert_allclose(e, (Hd @ psi) / psi, rtol=1e-12, atol=1e-12)
