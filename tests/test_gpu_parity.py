"""GPU parity: every HIP entry point against the FP64 oracle on the same inputs.

Tolerances (FP64 throughout): log psi / O_k / E_loc agree to 1e-11 relative
(different summation order only); s' enumeration, compaction counts, Metropolis
configurations and acceptance counts are bit-exact given the same Philox
stream; linear-algebra outputs (S, F, T, eigenvalues) to 1e-10 relative and
solver updates to 1e-7 relative (the pseudo-inverse amplifies rounding by the
condition number kept under the cutoff).
"""
import ctypes

import numpy as np
import pytest
import torch

from oracle import nets as onets
from oracle import operator as oop
from oracle import sampler as osamp
from oracle import solvers as osolv
from oracle.stats import SampledObs as OObs

from helpers import REF_ZZ, WEIGHTS16, to_np, tfim_pair, xxz_pair

pytestmark = pytest.mark.gpu

jv = pytest.importorskip("vmc_jax_amd")


def dev(x, dtype=None):
    t = torch.as_tensor(np.asarray(x), device="cuda")
    return t if dtype is None else t.to(dtype)


def make_psi(net_dev, shape, P):
    psi = jv.NQS(net_dev)
    psi(torch.zeros((1, 1) + tuple(shape), dtype=torch.int32))
    psi.set_parameters(P)
    return psi


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


@pytest.fixture(params=["rocsolver", "dc"])
def stedc(request, monkeypatch):
    """Tridiagonal stage of the eigensolvers: rocSOLVER dstedc or the library's
    own divide and conquer (read on every solve)."""
    monkeypatch.setenv("VMC_JAX_AMD_STEDC", request.param)
    return request.param


# ---------------------------------------------------------------- networks
@pytest.mark.parametrize("bias,vbias", [(False, False), (True, False), (True, True)])
def test_cpxrbm_forward_gradients(bias, vbias):
    rng = np.random.default_rng(10)
    L, H = 10, 20
    onet = onets.CpxRBM((L,), numHidden=H, bias=bias, visibleBias=vbias)
    P = 0.3 * rng.normal(size=onet.num_real_params())
    psi = make_psi(jv.nets.CpxRBM(numHidden=H, bias=bias, visibleBias=vbias), (L,), P)
    s = rng.integers(0, 2, size=(257, L)).astype(np.int32)
    opsi = onets.NQS(onet, P)
    lp = to_np(psi(dev(s[None])))[0]
    assert rel(lp, opsi(s)) < 1e-13
    G = to_np(psi.gradients(dev(s[None])))[0]
    assert rel(G, opsi.gradients(s)) < 1e-13
    Gc = to_np(psi.gradients(dev(s[None]), layout=0))[0]
    leaf = [g for g in onet.leaf_grads(onet.unflatten(P), s)]
    assert rel(Gc, np.concatenate(leaf, axis=1)) < 1e-13


def test_real_rbm_forward_gradients():
    rng = np.random.default_rng(11)
    L, H = 12, 7
    onet = onets.RBM((L,), numHidden=H, bias=True)
    P = 0.3 * rng.normal(size=onet.num_real_params())
    psi = make_psi(jv.nets.RBM(numHidden=H, bias=True), (L,), P)
    s = rng.integers(0, 2, size=(100, L)).astype(np.int32)
    opsi = onets.NQS(onet, P)
    assert rel(to_np(psi(dev(s[None])))[0], opsi(s)) < 1e-13
    assert rel(to_np(psi.gradients(dev(s[None])))[0], opsi.gradients(s)) < 1e-13


@pytest.mark.parametrize("shape,F,ch", [((10, 10), (3, 3), (8, 8, 8)), ((6,), (3,), (3, 2)),
                                        ((4, 5), (2, 3), (3, 4))])
def test_cnn_forward_gradients(shape, F, ch):
    rng = np.random.default_rng(12)
    onet = onets.CNN(shape, F=F, channels=ch, strides=(1,) * len(shape))
    net = jv.nets.CNN(F=F, channels=ch, strides=(1,) * len(shape))
    psi = jv.NQS(net)
    psi(torch.zeros((1, 1) + shape, dtype=torch.int32))
    P = net.he_normal(psi.desc, rng)
    psi.set_parameters(P)
    assert psi.numRealParameters == onet.num_real_params()
    s = rng.integers(0, 2, size=(37,) + shape).astype(np.int32)
    opsi = onets.NQS(onet, P)
    assert rel(to_np(psi(dev(s[None])))[0], opsi(s)) < 1e-12
    assert rel(to_np(psi.gradients(dev(s[None])))[0], opsi.gradients(s)) < 1e-12
    assert rel(to_np(psi.gradients(dev(s[None]), layout=0))[0], np.real(opsi.gradients(s))) < 1e-12


# ---------------------------------------------------------------- operators
def test_s_primes_bit_exact():
    rng = np.random.default_rng(13)
    L = 6
    from vmc_jax_amd import operator as dop
    H = dop.BranchFreeOperator()
    Ho = oop.BranchFreeOperator()
    for i in range(L):
        H.add(dop.scal_opstr(2., (dop.Sx(i),)))
        H.add(dop.scal_opstr(2., (dop.Sy(i), dop.Sz((i + 1) % L))))
        H.add(dop.scal_opstr(-1., (dop.Sz(i), dop.Sz((i + 2) % L))))
        H.add(dop.scal_opstr(0.5, (dop.Sp(i),)))
        Ho.add(oop.scal_opstr(2., (oop.Sx(i),)))
        Ho.add(oop.scal_opstr(2., (oop.Sy(i), oop.Sz((i + 1) % L))))
        Ho.add(oop.scal_opstr(-1., (oop.Sz(i), oop.Sz((i + 2) % L))))
        Ho.add(oop.scal_opstr(0.5, (oop.Sp(i),)))
    s = rng.integers(0, 2, size=(50, L)).astype(np.int32)
    sp, m = H.get_s_primes(dev(s[None]))
    spo, mo = Ho.get_s_primes(s)
    np.testing.assert_array_equal(to_np(sp)[0], spo)
    np.testing.assert_array_equal(to_np(m)[0], mo)
    np.testing.assert_array_equal(to_np(H.numNonzero)[0], Ho.numNonzero)


def test_time_dependent_prefactor():
    """operator_test.py:70-93 on the device."""
    rng = np.random.default_rng(3)
    s = rng.integers(0, 2, size=(24, 4)).astype(np.int32)
    from vmc_jax_amd import operator as dop
    h = dop.BranchFreeOperator()
    f = lambda t: 2.0 * t
    for i in range(3):
        h.add(dop.scal_opstr(f, (dop.Sp(i),)))
    for t in [0.5, 2, 13.9]:
        sp, m = h.get_s_primes(dev(s[None]), t)
        tmp = to_np(h.get_O_loc_unbatched(torch.ones(1, 24, device="cuda"), torch.ones(sp.shape[:2], device="cuda")))
        assert np.sum(np.abs(tmp - f(t) * np.sum(-(s[:, :3] - 1), axis=-1))) < 1e-7


@pytest.mark.parametrize("kind", ["tfim", "xxz"])
def test_eloc_rbm(kind):
    rng = np.random.default_rng(14)
    L, Hh = 10, 20
    onet = onets.CpxRBM((L,), numHidden=Hh, bias=True, visibleBias=True)
    P = 0.2 * rng.normal(size=onet.num_real_params())
    psi = make_psi(jv.nets.CpxRBM(numHidden=Hh, bias=True, visibleBias=True), (L,), P)
    opsi = onets.NQS(onet, P)
    H, Ho = tfim_pair((L,), -1.0, -0.7) if kind == "tfim" else xxz_pair(L, 0.6)
    s = rng.integers(0, 2, size=(300, L)).astype(np.int32)
    lp = psi(dev(s[None]))
    e = to_np(H.get_O_loc(dev(s[None]), psi, lp))[0]
    eo = Ho.get_O_loc(s, opsi, opsi(s))
    assert rel(e, eo) < 1e-12


def test_eloc_rbm_matches_dense_ed():
    rng = np.random.default_rng(15)
    L = 8
    onet = onets.CpxRBM((L,), numHidden=6, bias=True)
    P = 0.3 * rng.normal(size=onet.num_real_params())
    psi = make_psi(jv.nets.CpxRBM(numHidden=6, bias=True), (L,), P)
    H, Ho = tfim_pair((L,), -1.0, -0.7)
    es = jv.sampler.ExactSampler(psi, L)
    basis, lp, p = es.sample()
    e = to_np(H.get_O_loc(basis, psi, lp))[0]
    v = np.exp(to_np(lp)[0])
    ref = (Ho.dense_matrix(L) @ v) / v
    assert rel(e, ref) < 1e-12


@pytest.mark.parametrize("shape", [(10, 10), (6, 6), (12,)])
def test_eloc_cnn(shape):
    rng = np.random.default_rng(16)
    F = (3,) * len(shape)
    ch = (8, 8, 8) if len(shape) == 2 else (4, 3)
    onet = onets.CNN(shape, F=F, channels=ch, strides=(1,) * len(shape))
    net = jv.nets.CNN(F=F, channels=ch, strides=(1,) * len(shape))
    psi = jv.NQS(net)
    psi(torch.zeros((1, 1) + shape, dtype=torch.int32))
    P = net.he_normal(psi.desc, rng)
    psi.set_parameters(P)
    opsi = onets.NQS(onet, P)
    H, Ho = tfim_pair(shape, -1.0, -3.0)
    s = rng.integers(0, 2, size=(48,) + shape).astype(np.int32)
    e = to_np(H.get_O_loc(dev(s[None]), psi, psi(dev(s[None]))))[0]
    eo = Ho.get_O_loc(s, opsi, opsi(s))
    assert rel(e, eo) < 1e-11


@pytest.mark.parametrize("shape,ch", [((10, 10), (8, 8, 8)), ((6, 6), (4, 4, 4)), ((7, 9), (8, 8)),
                                      ((7, 9), (4, 8, 4))])  # last: a 7-wide window on 7 rows
def test_eloc_cnn_incremental_windows_fp64(shape, ch, monkeypatch):
    """FP64 E_loc with incremental windows (VMC_JAX_AMD_ELOC_DELTA=1: stored
    pre-activations plus the changed inputs) against window recomputation (the
    FP64 default) and the oracle: the two orders of summation agree to a few
    ulps of log psi."""
    rng = np.random.default_rng(sum(shape) + 41)
    net = jv.nets.CNN(F=(3, 3), channels=ch, strides=(1, 1))
    psi = jv.NQS(net)
    psi(torch.zeros((1, 1) + shape, dtype=torch.int32))
    P = net.he_normal(psi.desc, rng)
    psi.set_parameters(P)
    opsi = onets.NQS(onets.CNN(shape, F=(3, 3), channels=ch, strides=(1, 1)), P)
    H, Ho = tfim_pair(shape, -1.0, -3.0)
    cs = rng.integers(0, 2, size=(128,) + shape).astype(np.int32)
    s = dev(cs[None])
    lp = psi(s)
    monkeypatch.setenv("VMC_JAX_AMD_ELOC_DELTA", "0")
    e0 = to_np(H.get_O_loc(s, psi, lp))[0]
    monkeypatch.setenv("VMC_JAX_AMD_ELOC_DELTA", "1")
    e1 = to_np(H.get_O_loc(s, psi, lp))[0]
    monkeypatch.delenv("VMC_JAX_AMD_ELOC_DELTA")
    assert rel(e1, e0) < 1e-13
    assert rel(e1, Ho.get_O_loc(cs, opsi, opsi(cs))) < 1e-11


def test_eloc_window_kernel_equals_generic_kernel():
    """Single-flip window kernel and the generic masked kernel on the same C3
    inputs (max_sites_changed = 0 forces the generic path)."""
    rng = np.random.default_rng(28)
    shape = (10, 10)
    net = jv.nets.CNN(F=(3, 3), channels=(8, 8, 8), strides=(1, 1))
    psi = jv.NQS(net)
    psi(torch.zeros((1, 1) + shape, dtype=torch.int32))
    psi.set_parameters(net.he_normal(psi.desc, rng))
    H, _ = tfim_pair(shape, -1.0, -3.0)
    H2, _ = tfim_pair(shape, -1.0, -3.0)
    H2.compile()
    assert H.compile().max_sites_changed == 1
    H2.max_sites_changed = 0
    s = dev(rng.integers(0, 2, size=(1, 512) + shape).astype(np.int32))
    lp = psi(s)
    assert rel(to_np(H.get_O_loc(s, psi, lp)), to_np(H2.get_O_loc(s, psi, lp))) < 1e-13


def test_eloc_fused_equals_materialised_c3():
    """Full C3 size (10x10, 8192 configurations): the fused kernel equals the
    materialised s' + forward path (base.py:166-192) on the device."""
    rng = np.random.default_rng(17)
    shape = (10, 10)
    net = jv.nets.CNN(F=(3, 3), channels=(8, 8, 8), strides=(1, 1))
    psi = jv.NQS(net)
    psi(torch.zeros((1, 1) + shape, dtype=torch.int32))
    psi.set_parameters(net.he_normal(psi.desc, rng))
    H, _ = tfim_pair(shape, -1.0, -3.0)
    s = dev(rng.integers(0, 2, size=(1, 8192) + shape).astype(np.int32))
    lp = psi(s)
    e1 = H.get_O_loc(s, psi, lp)
    sp, _ = H.get_s_primes(s)
    e2 = H.get_O_loc_unbatched(lp, psi(sp))
    assert rel(to_np(e1), to_np(e2)) < 1e-12


# ---------------------------------------------------------------- samplers
def _mc_pair(onet, net_dev, shape, P, chains, seed, therm, sweep):
    psi = make_psi(net_dev, shape, P)
    opsi = onets.NQS(onet, P)
    ms = jv.sampler.MCSampler(psi, shape, seed, updateProposer=jv.sampler.propose_spin_flip, numChains=chains,
                              thermalizationSweeps=therm, sweepSteps=sweep)
    mo = osamp.MCSampler(opsi, shape, seed, numChains=chains, thermalizationSweeps=therm, sweepSteps=sweep)
    return psi, opsi, ms, mo


def test_mc_rbm_bit_exact():
    rng = np.random.default_rng(18)
    L = 12
    onet = onets.CpxRBM((L,), numHidden=8, bias=True)
    P = 0.5 * rng.normal(size=onet.num_real_params())
    psi, opsi, ms, mo = _mc_pair(onet, jv.nets.CpxRBM(numHidden=8, bias=True), (L,), P, 40, 7, 3, L)
    for call in range(2):  # walkers persist between calls
        s, lp, p = ms.sample(numSamples=200)
        so, lpo, po = mo.sample(numSamples=200)
        np.testing.assert_array_equal(to_np(s)[0], so)
        assert rel(to_np(lp)[0], lpo) < 1e-13
        assert rel(to_np(p)[0], po) < 1e-13
        assert float(ms.acceptance_ratio()[0]) == mo.acceptance_ratio()


@pytest.mark.parametrize("ch", [(4, 4), (4, 4, 4)])
def test_mc_cnn_bit_exact(ch):
    """(4, 4): receptive-field windows fit the 6x6 lattice (window kernel);
    (4, 4, 4): they wrap (generic masked kernel)."""
    rng = np.random.default_rng(19)
    shape = (6, 6)
    onet = onets.CNN(shape, F=(3, 3), channels=ch, strides=(1, 1))
    net = jv.nets.CNN(F=(3, 3), channels=ch, strides=(1, 1))
    psi0 = jv.NQS(net)
    psi0(torch.zeros((1, 1) + shape, dtype=torch.int32))
    P = net.he_normal(psi0.desc, rng)
    psi, opsi, ms, mo = _mc_pair(onet, net, shape, P, 24, 3, 2, 10)
    for call in range(2):
        s, lp, p = ms.sample(numSamples=96)
        so, lpo, po = mo.sample(numSamples=96)
        np.testing.assert_array_equal(to_np(s)[0], so)
        assert rel(to_np(lp)[0], lpo) < 1e-12
        assert float(ms.acceptance_ratio()[0]) == mo.acceptance_ratio()


def test_mc_histogram_matches_exact():
    """sampler_test.py:32-82: MCMC histogram against exact probabilities."""
    psi = make_psi(jv.nets.CpxRBM(numHidden=2), (4,), WEIGHTS16)
    pex = to_np(jv.sampler.ExactSampler(psi, 4).sample()[2])[0]
    ms = jv.sampler.MCSampler(psi, (4,), 0, updateProposer=jv.sampler.propose_spin_flip, numChains=777)
    s, _, p = ms.sample(numSamples=500000)
    k = (to_np(s)[0] * (1 << np.arange(4))).sum(1)
    pmc = np.bincount(k, weights=to_np(p)[0], minlength=16)
    assert np.max(np.abs(pmc - pex)) < 2e-3


def test_exact_sampler():
    psi = make_psi(jv.nets.CpxRBM(numHidden=2), (4,), WEIGHTS16)
    opsi = onets.NQS(onets.CpxRBM((4,), numHidden=2), WEIGHTS16)
    b, lp, p = jv.sampler.ExactSampler(psi, 4).sample()
    bo, lpo, po = osamp.ExactSampler(opsi, 4).sample()
    np.testing.assert_array_equal(to_np(b)[0], bo)
    assert rel(to_np(p)[0], po) < 1e-13


# ---------------------------------------------------------------- statistics / linear algebra
@pytest.mark.parametrize("cpx", [False, True])
def test_center_and_grams(cpx):
    rng = np.random.default_rng(20)
    n, p = 300, 70
    O = rng.normal(size=(n, p)) + (1j * rng.normal(size=(n, p)) if cpx else 0)
    w = rng.uniform(size=n); w /= w.sum()
    o = jv.stats.SampledObs(dev(O[None]), dev(w[None]))
    oo = OObs(O, w)
    assert rel(to_np(o.mean()), oo.mean()) < 1e-13
    assert rel(to_np(o._data), oo._data) < 1e-13
    assert rel(to_np(o.covar()), oo.covar()) < 1e-12
    assert rel(to_np(o.tangent_kernel()), oo.tangent_kernel()) < 1e-12
    assert rel(to_np(o.var()), oo.var()) < 1e-13


@pytest.mark.parametrize("cpx", [False, True])
def test_large_products_through_blas(cpx):
    """Plain products above 2^31 flops run on rocBLAS (deterministic mode): a
    GEMM with transposed second operand (the SNR projection's layout) and the
    covariance product, against numpy; the Hermitian Grams at the same sizes
    stay on the MFMA kernel."""
    rng = np.random.default_rng(33)
    z = (lambda *s: rng.normal(size=s) + 1j * rng.normal(size=s)) if cpx else (lambda *s: rng.normal(size=s))
    M, N, K = (1536, 768, 1024) if not cpx else (1024, 512, 512)
    A, B = z(M, K), z(N, K)
    C = torch.empty((M, N) + ((2,) if cpx else ()), dtype=torch.float64, device="cuda")
    Ad, Bd = dev(A).contiguous(), dev(B).contiguous()
    if cpx:
        Ad, Bd = torch.view_as_real(Ad).contiguous(), torch.view_as_real(Bd).contiguous()
    jv._lib.call("jv_gemm", M, N, K, jv._lib.ptr(Ad), K, 1, 0, jv._lib.ptr(Bd), 1, K, 0, jv._lib.ptr(C), N,
                 int(cpx), 0.5, jv._lib.stream_ptr())
    got = to_np(torch.view_as_complex(C) if cpx else C)
    assert rel(got, 0.5 * A @ B.T) < 1e-13
    n, pp = (4096, 512) if not cpx else (2048, 512)
    O = z(n, pp)
    o = jv.stats.SampledObs(dev(O[None]), dev(np.full((1, n), 1.0 / n)))
    oo = OObs(O, np.full(n, 1.0 / n))
    assert rel(to_np(o.covar()), oo.covar()) < 1e-12
    assert rel(to_np(o.tangent_kernel()), oo.tangent_kernel()) < 1e-12
    # deterministic: a second evaluation is bit-identical
    assert np.array_equal(to_np(o.covar()), to_np(o.covar()))


@pytest.mark.parametrize("cpx,n", [(False, 100), (True, 100), (False, 200), (True, 200), (False, 300), (False, 777),
                                   (True, 300), (True, 777)])
def test_eigh(cpx, n, stedc):
    """rocSOLVER below 128; from 128 up the Householder tridiagonalisation
    (real: dsytrd conventions; complex Hermitian: zhetrd conventions, real T)
    + divide and conquer + dormtr / zunmtr."""
    rng = np.random.default_rng(21)
    A = rng.normal(size=(n, n)) + (1j * rng.normal(size=(n, n)) if cpx else 0)
    A = A + A.conj().T
    Ad = dev(A).contiguous()
    w = torch.empty(n, dtype=torch.float64, device="cuda")
    jv._lib.call("jv_eigh", n, int(cpx), jv._lib.ptr(Ad), jv._lib.ptr(w), jv._lib.stream_ptr())
    V = to_np(Ad)
    ev = to_np(w)
    assert rel(ev, np.linalg.eigvalsh(A)) < 1e-12
    assert np.linalg.norm(A @ V - V * ev[None, :]) / np.linalg.norm(A) < 1e-12


# ---------------------------------------------------------------- solvers
def _cnn_case(shape, ch, N, seed):
    rng = np.random.default_rng(seed)
    onet = onets.CNN(shape, F=(3, 3), channels=ch, strides=(1, 1))
    net = jv.nets.CNN(F=(3, 3), channels=ch, strides=(1, 1))
    psi = jv.NQS(net)
    psi(torch.zeros((1, 1) + shape, dtype=torch.int32))
    P = net.he_normal(psi.desc, rng)
    psi.set_parameters(P)
    opsi = onets.NQS(onet, P)
    H, Ho = tfim_pair(shape, -1.0, -3.0)
    ms = jv.sampler.MCSampler(psi, shape, seed, updateProposer=jv.sampler.propose_spin_flip, numChains=64)
    s, lp, p = ms.sample(numSamples=N)
    return psi, opsi, H, Ho, s, lp, p


def test_minsr_real_cnn():
    psi, opsi, H, Ho, s, lp, p = _cnn_case((6, 6), (4, 4, 4), 512, 22)
    eq = jv.util.MinSR(None, pinvTol=1e-10)
    El = jv.stats.SampledObs(H.get_O_loc(s, psi, lp, 0.0), p)
    G = jv.stats.SampledObs(psi.gradients(s, layout=0), p)
    upd = to_np(eq.solve(El, G, False))
    cs, pw = to_np(s)[0], to_np(p)[0]
    Elo = OObs(Ho.get_O_loc(cs, opsi, opsi(cs), 0.0), pw)
    Go = OObs(opsi.gradients(cs), pw)
    ref = osolv.MinSR(None, pinvTol=1e-10).solve(Elo, Go, False)
    assert abs(complex(El.mean()[0].item()) - Elo.mean()[0]) <= 1e-12 * abs(Elo.mean()[0])
    assert rel(upd, ref) < 1e-7


def test_minsr_holomorphic_rbm():
    rng = np.random.default_rng(23)
    L = 8
    onet = onets.CpxRBM((L,), numHidden=6, bias=True)
    P = 0.1 * rng.normal(size=onet.num_real_params())
    psi = make_psi(jv.nets.CpxRBM(numHidden=6, bias=True), (L,), P)
    opsi = onets.NQS(onet, P)
    H, Ho = tfim_pair((L,), -1.0, -0.7)
    es, eso = jv.sampler.ExactSampler(psi, L), osamp.ExactSampler(opsi, L)
    eq, eqo = jv.util.MinSR(es, pinvTol=1e-8), osolv.MinSR(eso, pinvTol=1e-8)
    upd = to_np(eq(psi.get_parameters(), 0.0, psi=psi, hamiltonian=H, intStep=0))
    ref = eqo(opsi.get_parameters(), 0.0, psi=opsi, hamiltonian=Ho, intStep=0)
    assert rel(upd, ref) < 1e-7
    assert abs(eq.get_energy_mean() - eqo.get_energy_mean()) < 1e-11


@pytest.mark.parametrize("mode,H", [("real", 8), ("imag", 8), ("real", 16), ("imag", 32)])
def test_tdvp_exact(mode, H):
    """H = 8 gives Pc = 80 (rocSOLVER for G^H G); H = 16 gives Pc = 152 and
    H = 32 Pc = 296: G^H G goes through the complex Hermitian tridiagonal
    path in both modes."""
    rng = np.random.default_rng(24)
    L = 8
    onet = onets.CpxRBM((L,), numHidden=H, bias=True, visibleBias=True)
    P = 0.1 * rng.normal(size=onet.num_real_params())
    psi = make_psi(jv.nets.CpxRBM(numHidden=H, bias=True, visibleBias=True), (L,), P)
    opsi = onets.NQS(onet, P)
    H, Ho = tfim_pair((L,), -1.0, -0.7)
    es, eso = jv.sampler.ExactSampler(psi, L), osamp.ExactSampler(opsi, L)
    x = 1.0 if mode == "real" else 1.j
    eq = jv.util.TDVP(es, snrTol=2, pinvTol=1e-8, rhsPrefactor=x, makeReal=mode)
    eqo = osolv.TDVP(eso, snrTol=2, pinvTol=1e-8, rhsPrefactor=x, makeReal=mode)
    upd = to_np(eq(psi.get_parameters(), 0.0, psi=psi, hamiltonian=H, intStep=0))
    ref = eqo(opsi.get_parameters(), 0.0, psi=opsi, hamiltonian=Ho, intStep=0)
    assert rel(to_np(eq.ev), eqo.ev) < 1e-10
    assert rel(to_np(eq.S), eqo.S) < 1e-12
    assert rel(upd, ref) < 1e-7
    md, mo = eq.get_metadata(), eqo.metaData
    assert abs(md["tdvp_residual"] - mo["tdvp_residual"]) < 1e-9
    assert md["pinv_cutoff"] == mo["pinv_cutoff"]
    assert abs(md["tdvp_error"] - mo["tdvp_error"]) < 1e-8 * max(1.0, mo["tdvp_error"])
    assert abs(eq.get_energy_variance() - eqo.get_energy_variance()) < 1e-10


@pytest.mark.parametrize("mode", ["imag", "real"])
def test_tdvp_holomorphic_solve_equals_general(mode, monkeypatch, stedc):
    """jv_tdvp_solve_holomorphic (S through the Pc x Pc Gram G^H G) against
    jv_tdvp_solve on the same centred data: 'imag' with Metropolis samples and
    the SNR regulariser (eigenvalues +-lambda, non-degenerate), 'real' with the
    exact sampler (doubly degenerate spectrum; the update is basis independent)."""
    rng = np.random.default_rng(26)
    L = 8
    onet = onets.CpxRBM((L,), numHidden=16, bias=True, visibleBias=True)
    P = 0.1 * rng.normal(size=onet.num_real_params())
    psi = make_psi(jv.nets.CpxRBM(numHidden=16, bias=True, visibleBias=True), (L,), P)
    H, _ = tfim_pair((L,), -1.0, -0.7)
    if mode == "imag":
        sm = jv.sampler.MCSampler(psi, (L,), 6, updateProposer=jv.sampler.propose_spin_flip, numChains=64)
        s, lp, p = sm.sample(numSamples=2048)
    else:
        sm = jv.sampler.ExactSampler(psi, L)
        s, lp, p = sm.sample()
    x = 1.j if mode == "imag" else 1.0
    El = jv.stats.SampledObs(H.get_O_loc(s, psi, lp, 0.0), p)
    G = jv.stats.SampledObs(psi.gradients(s), p)
    eq = jv.util.TDVP(sm, snrTol=2, pinvTol=1e-8, rhsPrefactor=x, makeReal=mode)
    u1, r1, c1 = eq.solve(El, G)
    ev1, S1, snr1 = to_np(eq.ev), to_np(eq.S), to_np(eq.snr)
    eq._holo_cols = psi.holomorphic_columns()
    monkeypatch.setitem(jv.util.tdvp.HOLOMORPHIC_SOLVE, mode, 1)  # route Pc = 152 through the pair solve
    u2, r2, c2 = eq.solve(El, G)
    assert psi.numRealParameters == 2 * eq._holo_cols[0].numel() == 304
    assert rel(to_np(eq.ev), ev1) < 1e-12
    assert rel(to_np(eq.S), S1) < 1e-13
    if mode == "imag":
        assert rel(to_np(eq.snr), snr1) < 1e-8
    assert c1 == c2 and abs(r1 - r2) < 1e-10
    assert rel(to_np(u2), to_np(u1)) < 1e-9


def test_tdvp_mc_snr_real_rbm():
    """MC mode with the SNR regulariser.  A real RBM has a non-degenerate
    spectrum, so the eigenbasis (and hence SNR) is unique up to sign."""
    rng = np.random.default_rng(25)
    L = 10
    onet = onets.RBM((L,), numHidden=6, bias=True)
    P = 0.3 * rng.normal(size=onet.num_real_params())
    psi = make_psi(jv.nets.RBM(numHidden=6, bias=True), (L,), P)
    opsi = onets.NQS(onet, P)
    H, Ho = tfim_pair((L,), -1.0, -0.7)
    ms = jv.sampler.MCSampler(psi, (L,), 5, updateProposer=jv.sampler.propose_spin_flip, numChains=64)
    s, lp, p = ms.sample(numSamples=2048)
    eq = jv.util.TDVP(ms, snrTol=2, pinvTol=1e-6, rhsPrefactor=1.0, makeReal="real")
    El = jv.stats.SampledObs(H.get_O_loc(s, psi, lp, 0.0), p)
    G = jv.stats.SampledObs(psi.gradients(s), p)
    upd, res, cut = eq.solve(El, G)
    cs, pw = to_np(s)[0], to_np(p)[0]
    eso = osamp.MCSampler(opsi, (L,), 5, numChains=64)
    eso.globNumSamples = ms.globNumSamples
    eqo = osolv.TDVP(eso, snrTol=2, pinvTol=1e-6, rhsPrefactor=1.0, makeReal="real")
    ref, reso, cuto = eqo.solve(OObs(Ho.get_O_loc(cs, opsi, opsi(cs), 0.0), pw), OObs(opsi.gradients(cs), pw))
    assert rel(np.sort(to_np(eq.snr)), np.sort(eqo.snr)) < 1e-8
    assert cut == cuto
    assert rel(to_np(upd), ref) < 1e-7


# ---------------------------------------------------------------- end to end (reference tests)
def test_time_evolution_zz_golden():
    """tdvp_test.py:59-128 on the device: ZZ(t) within 1e-3 of the golden curve."""
    from scipy.interpolate import interp1d
    from vmc_jax_amd import operator as dop
    L = 4
    psi = make_psi(jv.nets.CpxRBM(numHidden=2), (L,), WEIGHTS16)
    H, _ = tfim_pair((L,), -1.0, -0.3)
    ZZ = dop.BranchFreeOperator()
    for l in range(L):
        ZZ.add((dop.Sz(l), dop.Sz((l + 1) % L)))
    es = jv.sampler.ExactSampler(psi, L)
    stepper = jv.util.stepper.AdaptiveHeun(timeStep=1e-3, tol=1e-5)
    eq = jv.util.TDVP(es, snrTol=1, pinvTol=0.0, pinvCutoff=1e-8, rhsPrefactor=1.j, diagonalShift=0.,
                      makeReal='imag')
    t = 0
    times = [0]
    m = jv.util.measure({'E': H, 'ZZ': ZZ}, psi, es)
    obs = [[m['E']['mean'][0], m['ZZ']['mean'][0]]]
    while t < 0.5:
        dp, dt = stepper.step(0, eq, psi.get_parameters(), hamiltonian=H, psi=psi, numSamples=0)
        psi.set_parameters(dp)
        t += dt
        times.append(t)
        m = jv.util.measure({'E': [(H, t)], 'ZZ': ZZ}, psi, es)
        obs.append([m['E']['mean'][0], m['ZZ']['mean'][0]])
    obs = np.array(obs)
    assert np.max(np.abs((obs[:, 0] - obs[0, 0]) / obs[0, 0])) < 1e-3
    zz = interp1d(np.array(times), obs[:, 1])(np.arange(0, 0.5, 0.05))
    assert np.max(np.abs(zz - REF_ZZ[:len(zz)])) < 1e-3


@pytest.mark.parametrize("solver", ["tdvp", "minsr"])
def test_gs_search_ed(solver):
    """tdvp_test.py:26-56 / minsr_test.py:14-43: ground-state energy to 1e-3."""
    L = 4
    for hx, exE in [(-1.3, -6.10160339), (-0.3, -4.09296160)]:
        nh = 6 if solver == "tdvp" else 8
        psi = jv.NQS(jv.nets.CpxRBM(numHidden=nh), seed=1)
        psi(torch.zeros((1, 1, L), dtype=torch.int32))
        H, _ = tfim_pair((L,), -1.0, hx)
        es = jv.sampler.ExactSampler(psi, L)
        if solver == "tdvp":
            eq = jv.util.TDVP(es, snrTol=1, pinvTol=0.0, pinvCutoff=1e-8, rhsPrefactor=1., diagonalShift=2,
                              makeReal='real')
            jv.util.ground_state_search(psi, H, eq, es, numSteps=100, stepSize=5e-2)
        else:
            eq = jv.util.MinSR(es, pinvTol=1e-6, diagonalShift=0.)
            jv.util.ground_state_search(psi, H, eq, es, numSteps=200, stepSize=1e-2)
        E = jv.util.measure({"energy": H}, psi, es)["energy"]["mean"][0]
        assert abs((E - exE) / exE) < 1e-3


def test_minsr_dual_equals_reference_form():
    """The P x P dual solve gives the reference T-form update."""
    psi, opsi, H, Ho, s, lp, p = _cnn_case((6, 6), (4, 4, 4), 512, 26)
    El = jv.stats.SampledObs(H.get_O_loc(s, psi, lp, 0.0), p)
    G = jv.stats.SampledObs(psi.gradients(s, layout=0), p)
    a = to_np(jv.util.MinSR(None, pinvTol=1e-10, solver="T").solve(El, G, False))
    b = to_np(jv.util.MinSR(None, pinvTol=1e-10, solver="dual").solve(El, G, False))
    assert rel(b, a) < 1e-7
    c = to_np(jv.util.MinSR(None, pinvTol=1e-10, diagonalShift=1e-3, solver="T").solve(El, G, False))
    d = to_np(jv.util.MinSR(None, pinvTol=1e-10, diagonalShift=1e-3, solver="dual").solve(El, G, False))
    assert rel(d, c) < 1e-9


@pytest.mark.parametrize("shift,shape,ch,N,P,tol", [(0.0, (6, 6), (4, 4, 4), 1024, 332, 1e-10),
                                                   (1e-3, (6, 6), (4, 4, 4), 1024, 332, 1e-10),
                                                   (0.0, (10, 10), (8, 8, 8), 4096, 1240, 1e-9)])
def test_minsr_dual_deferred_equals_dual(shift, shape, ch, N, P, tol, stedc):
    """The deferred dual solve (E_loc produced on another stream, centred and
    projected on the library's auxiliary stream) gives the dual update.  At
    P = 1240 the auxiliary stream's GEMMs (explicit Q) run while the main
    stream's tridiagonal eigensolver runs its own.  The two paths centre E_loc
    in different places (library kernel / SampledObs) and, on the
    eigendecomposition route, apply Q explicitly or by reflectors; the
    rounding differences are amplified by the condition number (up to 1e10
    with pinvTol 1e-10; 5e-12 measured at P = 1240)."""
    psi, opsi, H, Ho, s, lp, p = _cnn_case(shape, ch, N, 27)
    El = jv.stats.SampledObs(H.get_O_loc(s, psi, lp, 0.0), p)
    G = jv.stats.SampledObs(psi.gradients(s, layout=0), p)
    eq = jv.util.MinSR(None, pinvTol=1e-10, diagonalShift=shift, solver="dual")
    u1 = to_np(eq.solve(El, G, False))
    w = p.reshape(-1).contiguous()
    # the Gram first (jv_minsr_dual_gram), then E_loc beside the solve
    # (jv_minsr_solve_dual_deferred_gram), as MinSR.__call__ runs them
    D, S = eq._dual_gram(G)
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        e2 = H.get_O_loc(s, psi, lp, 0.0).reshape(-1).contiguous()
        ready = torch.cuda.Event()
        ready.record(side)
    u2 = to_np(eq._solve_deferred(e2, ready, D, S, w))
    torch.cuda.current_stream().wait_stream(side)
    assert psi.numParameters == P and eq.last_info.rank > 0
    assert rel(u2, u1) < tol
    # the call above is the non-blocking form (info copied to pinned memory in
    # stream order, read by last_info); the blocking form on the same inputs
    # gives the same update and the same info
    i2 = eq.last_info
    D4, S4 = eq._dual_gram(G)
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        e4 = H.get_O_loc(s, psi, lp, 0.0).reshape(-1).contiguous()
        ready4 = torch.cuda.Event()
        ready4.record(side)
    u4 = torch.empty(P, dtype=torch.float64, device="cuda")
    i4 = jv._lib.JvSolveInfo()
    jv._lib.call("jv_minsr_solve_dual_deferred_gram", jv._lib.ptr(D4), 0, jv._lib.ptr(S4), jv._lib.ptr(e4),
                 jv._lib.ptr(w), ready4._as_parameter_, D4.shape[0], P, 0, 1e-10, float(shift),
                 int(i2.iterations == 1), jv._lib.ptr(u4), ctypes.byref(i4), jv._lib.stream_ptr())
    torch.cuda.current_stream().wait_stream(side)
    assert rel(to_np(u4), u2) < 1e-13
    assert (i4.rank, i4.iterations) == (i2.rank, i2.iterations)
    assert abs(i4.e_mean[0] - i2.e_mean[0]) <= 1e-13 * abs(i4.e_mean[0]) and i4.e_mean[1] == i2.e_mean[1]
    assert abs(i4.e_var - i2.e_var) <= 1e-12 * abs(i4.e_var)
    em = El.mean()[0].item()
    assert abs(i2.e_mean[0] - em.real) <= 1e-12 * abs(em)
    # the one-call deferred form (Gram inside, E_loc already in flight)
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        e3 = H.get_O_loc(s, psi, lp, 0.0).reshape(-1).contiguous()
        ready3 = torch.cuda.Event()
        ready3.record(side)
    u3 = torch.empty(P, dtype=torch.float64, device="cuda")
    info = jv._lib.JvSolveInfo()
    jv._lib.call("jv_minsr_solve_dual_deferred", jv._lib.ptr(D), jv._lib.ptr(e3), jv._lib.ptr(w),
                 ready3._as_parameter_, D.shape[0], P, 0, 1e-10, float(shift), eq._cholesky_hint(), jv._lib.ptr(u3),
                 ctypes.byref(info), jv._lib.stream_ptr())
    torch.cuda.current_stream().wait_stream(side)
    assert rel(to_np(u3), u1) < tol


def test_minsr_dual_overlapped_call_matches_oracle(stedc):
    """MinSR.__call__ with the dual form runs E_loc beside the solve; exact
    sampler on a 3x3 lattice (512 states > P = 184)."""
    rng = np.random.default_rng(28)
    shape = (3, 3)
    onet = onets.CNN(shape, F=(3, 3), channels=(4, 4), strides=(1, 1))
    net = jv.nets.CNN(F=(3, 3), channels=(4, 4), strides=(1, 1))
    psi = jv.NQS(net)
    psi(torch.zeros((1, 1) + shape, dtype=torch.int32))
    P = net.he_normal(psi.desc, rng)
    psi.set_parameters(P)
    opsi = onets.NQS(onet, P)
    H, Ho = tfim_pair(shape, -1.0, -3.0)
    es, eso = jv.sampler.ExactSampler(psi, shape), osamp.ExactSampler(opsi, shape)
    eq, eqo = jv.util.MinSR(es, pinvTol=1e-10, solver="dual"), osolv.MinSR(eso, pinvTol=1e-10)
    assert eq._overlapped(psi, 512, None)
    upd = to_np(eq(psi.get_parameters(), 0.0, psi=psi, hamiltonian=H, intStep=0))
    ref = eqo(opsi.get_parameters(), 0.0, psi=opsi, hamiltonian=Ho, intStep=0)
    assert rel(upd, ref) < 1e-7
    assert abs(eq.get_energy_mean() - eqo.get_energy_mean()) < 1e-11 * abs(eqo.get_energy_mean())
    assert abs(eq.get_energy_variance() - eqo.get_energy_variance()) < 1e-10 * abs(eqo.get_energy_variance())


@pytest.mark.parametrize("shape,ch,N,seed", [((10, 10), (8, 8, 8), 2048, 31), ((6, 6), (4, 4, 4), 128, 32)])
def test_minsr_dual_tridiagonal(shape, ch, N, seed, stedc):
    """The dual solve through the Householder tridiagonalisation (P >= 256) equals
    the reference T form: full rank (the 10x10 workload's 1240 parameters) and
    rank deficient (N < P, most eigenvalues of O^T O at round-off level)."""
    psi, opsi, H, Ho, s, lp, p = _cnn_case(shape, ch, N, seed)
    El = jv.stats.SampledObs(H.get_O_loc(s, psi, lp, 0.0), p)
    G = jv.stats.SampledObs(psi.gradients(s, layout=0), p)
    assert psi.desc.num_real_params >= 256
    for tol, shift in ((1e-10, 0.0), (1e-8, 1e-3)):
        a = to_np(jv.util.MinSR(None, pinvTol=tol, diagonalShift=shift, solver="T").solve(El, G, False))
        b = to_np(jv.util.MinSR(None, pinvTol=tol, diagonalShift=shift, solver="dual").solve(El, G, False))
        assert rel(b, a) < 1e-7, (shape, N, tol, shift, rel(b, a))


def test_minsr_dual_cholesky_route(monkeypatch):
    """Real dual MinSR: when S has no eigenvalue at or below the pseudo-inverse
    cutoff the update comes from a Cholesky solve, and equals the
    eigendecomposition route; a cutoff above the smallest eigenvalue, or a
    singular S (N < P), must send the solve through the eigendecomposition."""
    psi, opsi, H, Ho, s, lp, p = _cnn_case((10, 10), (8, 8, 8), 4096, 33)
    El = jv.stats.SampledObs(H.get_O_loc(s, psi, lp, 0.0), p)
    G = jv.stats.SampledObs(psi.gradients(s, layout=0), p)
    D = to_np(G._data)
    lam = np.linalg.eigvalsh(D.T @ D)
    P = psi.numParameters
    assert P == 1240 and lam[0] > 1e-10 * lam[-1]
    eq = jv.util.MinSR(None, pinvTol=1e-10, solver="dual")
    uc = to_np(eq.solve(El, G, False))
    assert eq.last_info.iterations == 1 and eq.last_info.rank == P
    monkeypatch.setenv("VMC_JAX_AMD_MINSR_CHOLESKY", "0")
    ue = to_np(eq.solve(El, G, False))
    assert eq.last_info.iterations == 0 and eq.last_info.rank == P
    assert rel(uc, ue) < 1e-9
    monkeypatch.delenv("VMC_JAX_AMD_MINSR_CHOLESKY")
    # cutoff between the two smallest eigenvalues: one is cut, no Cholesky
    tol = np.sqrt(lam[0] * lam[1]) / lam[-1]
    eq2 = jv.util.MinSR(None, pinvTol=float(tol), solver="dual")
    eq2.solve(El, G, False)
    assert eq2.last_info.iterations == 0 and eq2.last_info.rank == P - 1
    # singular S
    psi, opsi, H, Ho, s, lp, p = _cnn_case((6, 6), (4, 4, 4), 128, 32)
    El = jv.stats.SampledObs(H.get_O_loc(s, psi, lp, 0.0), p)
    G = jv.stats.SampledObs(psi.gradients(s, layout=0), p)
    eq.solve(El, G, False)
    assert eq.last_info.iterations == 0 and eq.last_info.rank <= 128


def test_minsr_dual_holomorphic():
    rng = np.random.default_rng(27)
    L = 8
    onet = onets.CpxRBM((L,), numHidden=4, bias=True)
    P = 0.1 * rng.normal(size=onet.num_real_params())
    psi = make_psi(jv.nets.CpxRBM(numHidden=4, bias=True), (L,), P)
    opsi = onets.NQS(onet, P)
    H, Ho = tfim_pair((L,), -1.0, -0.7)
    es, eso = jv.sampler.ExactSampler(psi, L), osamp.ExactSampler(opsi, L)
    eq = jv.util.MinSR(es, pinvTol=1e-8, solver="dual")
    upd = to_np(eq(psi.get_parameters(), 0.0, psi=psi, hamiltonian=H, intStep=0))
    ref = osolv.MinSR(eso, pinvTol=1e-8)(opsi.get_parameters(), 0.0, psi=opsi, hamiltonian=Ho, intStep=0)
    assert rel(upd, ref) < 1e-7


def test_checkpoint_resume_bit_exact(tmp_path):
    """A run resumed from a checkpoint draws the same samples as the
    uninterrupted run (walkers and Philox counters are saved)."""
    rng = np.random.default_rng(29)
    L = 10
    P = 0.3 * rng.normal(size=onets.CpxRBM((L,), numHidden=6, bias=True).num_real_params())

    def fresh():
        psi = make_psi(jv.nets.CpxRBM(numHidden=6, bias=True), (L,), P)
        return psi, jv.sampler.MCSampler(psi, (L,), 11, updateProposer=jv.sampler.propose_spin_flip, numChains=32)
    psi, ms = fresh()
    ms.sample(numSamples=256)
    jv.util.checkpoint.save(str(tmp_path / "ck.npz"), psi, ms, time=0.5)
    ref = to_np(ms.sample(numSamples=256)[0])
    psi2, ms2 = fresh()
    t = jv.util.checkpoint.load(str(tmp_path / "ck.npz"), psi2, ms2)
    assert t == 0.5
    np.testing.assert_array_equal(to_np(ms2.sample(numSamples=256)[0]), ref)
