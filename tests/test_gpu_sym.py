"""GPU parity of the symmetrised networks (SymNet) against the FP64 oracle.

log psi and O_k agree to 1e-12 relative (summation order only), E_loc to
1e-11; Metropolis configurations are equal given the same Philox stream; the
MCMC histogram matches the exact probabilities to 2e-3 (sampler_test.py:32-82
with the SymNet of that test); TDVP / MinSR updates to 1e-7 relative.
"""
import warnings

import numpy as np
import pytest
import torch

from oracle import nets as onets
from oracle import sampler as osamp
from oracle import solvers as osolv
from oracle import symmetries as osym

from helpers import WEIGHTS16, to_np, tfim_pair

pytestmark = pytest.mark.gpu

jv = pytest.importorskip("vmc_jax_amd")
from vmc_jax_amd.util import symmetries as sym  # noqa: E402
from vmc_jax_amd.nets import sym_wrapper as sw  # noqa: E402

warnings.simplefilter("ignore", UserWarning)

AVG = {"exp": sw.avgFun_Coefficients_Exp, "log": sw.avgFun_Coefficients_Log, "sep": sw.avgFun_Coefficients_Sep}


def dev(x):
    return torch.as_tensor(np.asarray(x), device="cuda")


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def well_conditioned(lpo, span=12.0):
    """Samples whose symmetrised sum does not cancel: finite log psi within
    `span` of the largest.  Where the sum cancels (a configuration with no
    weight in the chosen symmetry sector) log psi is rounding noise in any
    implementation and is left out of the comparison."""
    re = np.real(lpo)
    fin = np.isfinite(re)
    return fin & (re > np.max(re[fin]) - span)


def orbits(shape, names, **kw):
    if len(shape) == 1:
        return sym.get_orbit_1D(shape[0], *names, **kw), osym.orbit_1d(shape[0], *names, **kw)
    return sym.get_orbit_2D_square(shape[0], *names, **kw), osym.orbit_2d_square(shape[0], *names, **kw)


def pair(kind, shape, names, avg, seed, **kw):
    """(psi, opsi, P) for the same network, orbit and parameters."""
    rng = np.random.default_rng(seed)
    orb, (d, f) = orbits(shape, names, **kw)
    if kind == "cpxrbm":
        net, onet = jv.nets.CpxRBM(numHidden=6, bias=True, visibleBias=True), \
            onets.CpxRBM(shape, numHidden=6, bias=True, visibleBias=True)
    elif kind == "rbm":
        net, onet = jv.nets.RBM(numHidden=5, bias=True), onets.RBM(shape, numHidden=5, bias=True)
    else:
        F = (3,) * len(shape)
        net, onet = jv.nets.CNN(F=F, channels=(3, 2), strides=(1,) * len(shape)), \
            onets.CNN(shape, F=F, channels=(3, 2), strides=(1,) * len(shape))
    psi = jv.NQS(net, orbit=orb, avgFun=AVG[avg])
    psi(torch.zeros((1, 1) + tuple(shape), dtype=torch.int32))
    P = 0.3 * rng.normal(size=onet.num_real_params())
    psi.set_parameters(P)
    opsi = onets.NQS(osym.SymNet(onet, d, f, avg), P)
    return psi, opsi, P


CASES = [
    ("cpxrbm", (8,), ("translation", "reflection", "spinflip"), "exp", {}),
    ("cpxrbm", (8,), ("translation",), "exp", {"translation_factor": np.exp(2j * np.pi / 8)}),
    ("cpxrbm", (3, 3), ("rotation", "reflection"), "log", {}),
    ("rbm", (8,), ("translation", "spinflip"), "exp", {"translation_factor": np.exp(2j * np.pi * 3 / 8)}),
    ("rbm", (8,), ("translation",), "log", {}),
    ("rbm", (7,), ("translation", "reflection"), "sep", {}),
    ("cnn", (4, 4), ("rotation", "reflection", "translation", "spinflip"), "exp", {}),
    ("cnn", (4, 4), ("translation",), "sep", {}),
    ("cnn", (4, 4), ("translation", "reflection"), "exp", {"reflection_factor": -1.0}),
]


@pytest.mark.parametrize("kind,shape,names,avg,kw", CASES)
def test_sym_forward_gradients(kind, shape, names, avg, kw):
    psi, opsi, P = pair(kind, shape, names, avg, 40, **kw)
    rng = np.random.default_rng(41)
    s = rng.integers(0, 2, size=(61,) + shape).astype(np.int32)
    lp = to_np(psi(dev(s[None])))[0]
    lpo = opsi(s)
    assert lp.dtype == (np.complex128 if np.iscomplexobj(lpo) else np.float64)
    np.testing.assert_array_equal(np.isnan(lp), np.isnan(lpo))
    ok = well_conditioned(lpo)
    assert ok.sum() >= len(s) // 3
    assert rel(lp[ok], lpo[ok]) < 1e-12
    G = to_np(psi.gradients(dev(s[None])))[0]
    Go = opsi.gradients(s)
    assert rel(G[ok], Go[ok]) < 1e-12
    Gc = to_np(psi.gradients(dev(s[None]), layout=0))[0]
    leaf = np.concatenate(opsi.net.leaf_grads(opsi.net.unflatten(P), s), axis=1)
    assert rel(Gc[ok], leaf[ok]) < 1e-12
    if not np.iscomplexobj(leaf):
        assert Gc.dtype == np.float64


def test_sym_real_exp_negative_sums():
    """Real network, real factor -1: NaN where the signed sum is negative and
    zero gradients there (jax logsumexp / where semantics)."""
    psi, opsi, P = pair("rbm", (8,), ("reflection",), "exp", 42, reflection_factor=-1.0)
    s = np.random.default_rng(43).integers(0, 2, size=(80, 8)).astype(np.int32)
    lp = to_np(psi(dev(s[None])))[0]
    lpo = opsi(s)
    assert np.isnan(lpo).any() and np.isfinite(lpo).any()
    np.testing.assert_array_equal(np.isnan(lp), np.isnan(lpo))
    G = to_np(psi.gradients(dev(s[None])))[0]
    assert np.all(G[np.isnan(lpo)] == 0)
    ok = well_conditioned(lpo)
    assert rel(lp[ok], lpo[ok]) < 1e-12
    assert rel(G[ok], opsi.gradients(s)[ok]) < 1e-12


def test_sym_full_group_10x10_cnn():
    """The 1600-element space group of the 10 x 10 lattice with spin flip."""
    psi, opsi, P = pair("cnn", (10, 10), ("rotation", "reflection", "translation", "spinflip"), "exp", 44)
    assert psi._sym[0].num_elements == 1600
    s = np.random.default_rng(45).integers(0, 2, size=(6, 10, 10)).astype(np.int32)
    assert rel(to_np(psi(dev(s[None])))[0], opsi(s)) < 1e-12
    assert rel(to_np(psi.gradients(dev(s[None])))[0], opsi.gradients(s)) < 1e-12


@pytest.mark.parametrize("kind,shape,names,avg,kw", [CASES[0], CASES[3], CASES[6]])
def test_sym_eloc(kind, shape, names, avg, kw):
    psi, opsi, P = pair(kind, shape, names, avg, 46, **kw)
    H, Ho = tfim_pair(shape, -1.0, -0.7)
    s = np.random.default_rng(47).integers(0, 2, size=(50,) + shape).astype(np.int32)
    lp = psi(dev(s[None]))
    e = to_np(H.get_O_loc(dev(s[None]), psi, lp))[0]
    lpo = opsi(s)
    eo = Ho.get_O_loc(s, opsi, lpo)
    ok = well_conditioned(lpo)
    assert rel(e[ok], eo[ok]) < 1e-11
    # materialised route: get_s_primes + psi(s') + the reference's sum
    sp, _ = H.get_s_primes(dev(s[None]))
    e2 = to_np(H.get_O_loc_unbatched(lp, psi(sp)))[0]
    assert rel(e2[ok], eo[ok]) < 1e-11


def test_sym_mc_equals_oracle_chain():
    psi, opsi, P = pair("cpxrbm", (8,), ("translation", "reflection"), "exp", 48)
    ms = jv.sampler.MCSampler(psi, (8,), 3, updateProposer=jv.sampler.propose_spin_flip, numChains=24,
                              thermalizationSweeps=2, sweepSteps=8)
    mo = osamp.MCSampler(opsi, (8,), 3, numChains=24, thermalizationSweeps=2, sweepSteps=8)
    for call in range(2):
        s, lp, p = ms.sample(numSamples=96)
        so, lpo, po = mo.sample(numSamples=96)
        np.testing.assert_array_equal(to_np(s)[0], so)
        assert rel(to_np(lp)[0], lpo) < 1e-12
        assert rel(to_np(p)[0], po) < 1e-12
        assert float(ms.acceptance_ratio()[0]) == mo.acceptance_ratio()


def test_sym_mc_histogram_matches_exact():
    """sampler_test.py:32-82: CpxRBM in a SymNet with translation, reflection and
    spin flip; 777 chains, 500000 samples."""
    orb = sym.get_orbit_1D(4, "translation", "reflection", "spinflip")
    psi = jv.NQS(jv.nets.SymNet(net=jv.nets.CpxRBM(numHidden=2, bias=False), orbit=orb))
    psi(torch.zeros((1, 1, 4), dtype=torch.int32))
    psi.set_parameters(WEIGHTS16)
    pex = to_np(jv.sampler.ExactSampler(psi, 4).sample()[2])[0]
    ms = jv.sampler.MCSampler(psi, (4,), 0, updateProposer=jv.sampler.propose_spin_flip, numChains=777)
    s, _, p = ms.sample(numSamples=500000)
    k = (to_np(s)[0] * (1 << np.arange(4))).sum(1)
    pmc = np.bincount(k, weights=to_np(p)[0], minlength=16)
    assert np.max(np.abs(pmc - pex)) < 2e-3


def test_sym_holomorphic_recognition():
    """vqs_test.py:90-103: a CpxRBM in a SymNet stays holomorphic."""
    for k in range(4):
        psi = jv.NQS(jv.nets.SymNet(net=jv.nets.CpxRBM(numHidden=2 ** k, bias=True), orbit=sym.get_orbit_1D(3)))
        psi(torch.zeros((1, 4, 3), dtype=torch.int32))
        assert psi.holomorphic


@pytest.mark.parametrize("mode", ["real", "imag"])
def test_sym_tdvp_exact(mode):
    psi, opsi, P = pair("cpxrbm", (8,), ("translation", "spinflip"), "exp", 49)
    H, Ho = tfim_pair((8,), -1.0, -0.7)
    es, eso = jv.sampler.ExactSampler(psi, 8), osamp.ExactSampler(opsi, 8)
    x = 1.0 if mode == "real" else 1.j
    eq = jv.util.TDVP(es, snrTol=2, pinvTol=1e-8, rhsPrefactor=x, makeReal=mode)
    eqo = osolv.TDVP(eso, snrTol=2, pinvTol=1e-8, rhsPrefactor=x, makeReal=mode)
    upd = to_np(eq(psi.get_parameters(), 0.0, psi=psi, hamiltonian=H, intStep=0))
    ref = eqo(opsi.get_parameters(), 0.0, psi=opsi, hamiltonian=Ho, intStep=0)
    assert rel(to_np(eq.ev), eqo.ev) < 1e-10
    assert rel(upd, ref) < 1e-7


def test_sym_minsr_real_params_complex_output():
    """Real RBM with a complex translation factor: the stacked [Re O; Im O]
    system of minsr.py:68-78."""
    psi, opsi, P = pair("rbm", (8,), ("translation",), "exp", 50, translation_factor=np.exp(2j * np.pi / 8))
    H, Ho = tfim_pair((8,), -1.0, -0.7)
    es, eso = jv.sampler.ExactSampler(psi, 8), osamp.ExactSampler(opsi, 8)
    eq, eqo = jv.util.MinSR(es, pinvTol=1e-8), osolv.MinSR(eso, pinvTol=1e-8)
    upd = to_np(eq(psi.get_parameters(), 0.0, psi=psi, hamiltonian=H, intStep=0))
    ref = eqo(opsi.get_parameters(), 0.0, psi=opsi, hamiltonian=Ho, intStep=0)
    assert rel(upd, ref) < 1e-7
