"""GPU parity of fermionic operator strings (Jordan-Wigner sign masks).

s' and matrix elements are bit-exact against the oracle; the reference's
golden impurity energy (operator_test.py:199-248) is reproduced from the
device's s' and local-estimator sum; the anticommutators hold to 1e-10 through
the fused local estimator (operator_test.py:174-197); fused E_loc of an RBM and
a CNN agree with the oracle to 1e-12.
"""
import numpy as np
import pytest
import torch

from oracle import nets as onets
from oracle import operator as oop

from helpers import to_np
from test_fermions import TablePsi, golden_energy, impurity_hamiltonian

pytestmark = pytest.mark.gpu

jv = pytest.importorskip("vmc_jax_amd")
from vmc_jax_amd import operator as dop  # noqa: E402


def dev(x):
    return torch.as_tensor(np.asarray(x), device="cuda")


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def test_fermionic_s_primes_bit_exact():
    H, Ho = impurity_hamiltonian(dop), impurity_hamiltonian(oop)
    s = np.random.default_rng(60).integers(0, 2, size=(200, 8)).astype(np.int32)
    sp, m = H.get_s_primes(dev(s[None]))
    spo, mo = Ho.get_s_primes(s)
    np.testing.assert_array_equal(to_np(sp)[0], spo)
    np.testing.assert_array_equal(to_np(m)[0], mo)
    np.testing.assert_array_equal(to_np(H.numNonzero)[0], Ho.numNonzero)


def test_golden_impurity_energy_on_device():
    H = impurity_hamiltonian(dop)
    psi, basis, lp, p = golden_energy()
    sp, _ = H.get_s_primes(dev(basis[None]))
    lpsp = psi(to_np(sp)[0])
    e = to_np(H.get_O_loc_unbatched(dev(lp[None]), dev(lpsp[None])))[0]
    assert np.allclose(np.sum(p * e), -9.95314531)


def test_anticommutators_fused_eloc():
    L = 5
    rng = np.random.default_rng(61)
    onet = onets.CpxRBM((L,), numHidden=3, bias=True)
    P = 0.4 * rng.normal(size=onet.num_real_params())
    psi = jv.NQS(jv.nets.CpxRBM(numHidden=3, bias=True))
    psi(torch.zeros((1, 1, L), dtype=torch.int32))
    psi.set_parameters(P)
    k = np.arange(2 ** L)
    basis = dev(((k[:, None] >> np.arange(L)[None, :]) & 1).astype(np.int32)[None])
    lp = psi(basis)
    for i in range(L):
        for j in range(L):
            C = dop.BranchFreeOperator()
            C.add(dop.scal_opstr(1., (dop.creation(j), dop.annihilation(i))))
            C.add(dop.scal_opstr(1., (dop.annihilation(i), dop.creation(j))))
            e = to_np(C.get_O_loc(basis, psi, lp))[0]
            # the two strings reach the same s' by flipping sites in opposite
            # orders, so the incremental log psi(s') differ in the last bit
            # (the reference test compares with allclose, atol 1e-8)
            np.testing.assert_allclose(e, np.full(2 ** L, 1.0 if i == j else 0.0), rtol=0, atol=1e-10)


@pytest.mark.parametrize("kind", ["cpxrbm", "cnn"])
def test_fermionic_fused_eloc(kind):
    rng = np.random.default_rng(62)
    if kind == "cpxrbm":
        onet = onets.CpxRBM((8,), numHidden=6, bias=True, visibleBias=True)
        net = jv.nets.CpxRBM(numHidden=6, bias=True, visibleBias=True)
    else:
        onet = onets.CNN((8,), F=(3,), channels=(4, 3), strides=(1,))
        net = jv.nets.CNN(F=(3,), channels=(4, 3), strides=(1,))
    P = 0.3 * rng.normal(size=onet.num_real_params())
    psi = jv.NQS(net)
    psi(torch.zeros((1, 1, 8), dtype=torch.int32))
    psi.set_parameters(P)
    opsi = onets.NQS(onet, P)
    H, Ho = impurity_hamiltonian(dop), impurity_hamiltonian(oop)
    s = rng.integers(0, 2, size=(120, 8)).astype(np.int32)
    e = to_np(H.get_O_loc(dev(s[None]), psi, psi(dev(s[None]))))[0]
    eo = Ho.get_O_loc(s, opsi, opsi(s))
    assert rel(e, eo) < 1e-12
