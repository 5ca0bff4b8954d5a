"""GPU parity at the benchmark configurations and the reference's Metropolis
time-evolution test.

* tdvp_test.py:131-198 (MCSampler, snrTol = 1, cross-validation) through the
  device path, plus the cross-validation metadata against the oracle on the
  same samples;
* the 10x10 / 8-channel CNN's Metropolis kernel bit-exact against the oracle;
* MinSR.__call__ on the 10x10 CNN (P = 1240, N = 2048) through the overlapped
  Cholesky route against the oracle's reference T form;
* one TDVP step of the L = 10 CpxRBM exact-sampler workload and one Heun
  right-hand side of the L = 32 CpxRBM / 16384-sample workload against the
  oracle on the same samples.

Tolerances as in test_gpu_parity: log psi, E_loc, S, F to 1e-10 relative
(summation order), configurations bit-exact, updates to 1e-7 relative.
"""
import numpy as np
import pytest
import torch

from oracle import nets as onets
from oracle import operator as oop
from oracle import sampler as osamp
from oracle import solvers as osolv
from oracle.stats import SampledObs as OObs

from helpers import REF_ZZ, WEIGHTS16, to_np, tfim_pair

pytestmark = pytest.mark.gpu

jv = pytest.importorskip("vmc_jax_amd")


def make_psi(net_dev, shape, P):
    psi = jv.NQS(net_dev)
    psi(torch.zeros((1, 1) + tuple(shape), dtype=torch.int32))
    psi.set_parameters(P)
    return psi


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


class _Recording:
    """Wraps a device sampler and keeps the last samples it returned."""

    def __init__(self, sampler):
        self.s = sampler
        self.last = None

    def __getattr__(self, k):
        return getattr(self.s, k)

    def sample(self, *a, **kw):
        self.last = self.s.sample(*a, **kw)
        return self.last


# ---------------------------------------------------------------- tdvp_test.py:131-198
def test_time_evolution_mc_zz_golden():
    """tdvp_test.py:131-198 on the device: Metropolis samples (mu = 1, 500
    chains), SNR-regularised real-time TDVP with cross-validation, AdaptiveHeun;
    energy drift < 1e-1 and ZZ(t) within 2e-2 of the golden curve."""
    from scipy.interpolate import interp1d
    from vmc_jax_amd import operator as dop
    L = 4
    psi = make_psi(jv.nets.CpxRBM(numHidden=2), (L,), WEIGHTS16)
    H, _ = tfim_pair((L,), -1.0, -0.3)
    ZZ = dop.BranchFreeOperator()
    for l in range(L):
        ZZ.add((dop.Sz(l), dop.Sz((l + 1) % L)))
    mc = jv.sampler.MCSampler(psi, (L,), 0, numSamples=50000, updateProposer=jv.sampler.propose_spin_flip, mu=1,
                              numChains=500)
    stepper = jv.util.stepper.AdaptiveHeun(timeStep=1e-3, tol=1e-4)
    eq = jv.util.TDVP(mc, snrTol=1, pinvTol=1e-8, rhsPrefactor=1.j, diagonalShift=0., makeReal='imag',
                      crossValidation=True)
    t = 0
    times = [0]
    m = jv.util.measure({'E': H, 'ZZ': ZZ}, psi, mc)
    obs = [[m['E']['mean'][0], m['ZZ']['mean'][0]]]
    while t < 0.2:
        dp, dt = stepper.step(0, eq, psi.get_parameters(), hamiltonian=H, psi=psi, numSamples=5000)
        psi.set_parameters(dp)
        t += dt
        times.append(t)
        m = jv.util.measure({'E': [(H, t)], 'ZZ': ZZ}, psi, mc)
        obs.append([m['E']['mean'][0], m['ZZ']['mean'][0]])
    obs = np.array(obs)
    assert np.max(np.abs((obs[:, 0] - obs[0, 0]) / obs[0, 0])) < 1e-1
    zz = interp1d(np.array(times), obs[:, 1])(np.arange(0, 0.2, 0.05))
    assert np.max(np.abs(zz - REF_ZZ[:len(zz)])) < 2e-2
    assert np.isfinite(eq.get_metadata()["tdvp_residual_cross_validation_ratio"])
    assert np.isfinite(eq.get_metadata()["tdvp_error_cross_validation_ratio"])


@pytest.mark.parametrize("mode,shift", [("imag", 0.0), ("real", 1e-2)])
def test_tdvp_cross_validation_matches_oracle(mode, shift):
    """TDVP.__call__ with crossValidation (tdvp.py:306-324) on Metropolis samples
    identical to the oracle's (bit-exact sampler): update, tdvp_error, residual
    and both cross-validation ratios.  The spectrum stays non-degenerate, so the
    SNR-regularised update is unique: 'real' uses a real RBM; 'imag' a CpxRBM
    (S = i Im S0 has the eigenvalues +-lambda(G^H G); a real network's
    Im S0 vanishes, leaving S = 0 and a NaN update in the reference too)."""
    rng = np.random.default_rng(41)
    L = 10
    if mode == "imag":
        onet = onets.CpxRBM((L,), numHidden=6, bias=True)
        dnet = jv.nets.CpxRBM(numHidden=6, bias=True)
    else:
        onet = onets.RBM((L,), numHidden=6, bias=True)
        dnet = jv.nets.RBM(numHidden=6, bias=True)
    P = 0.3 * rng.normal(size=onet.num_real_params())
    psi = make_psi(dnet, (L,), P)
    opsi = onets.NQS(onet, P)
    H, Ho = tfim_pair((L,), -1.0, -0.7)
    x = 1.j if mode == "imag" else 1.0
    ms = jv.sampler.MCSampler(psi, (L,), 9, updateProposer=jv.sampler.propose_spin_flip, numChains=64,
                              numSamples=2048)
    mo = osamp.MCSampler(opsi, (L,), 9, numChains=64, numSamples=2048)
    kw = dict(snrTol=2, pinvTol=1e-6, rhsPrefactor=x, makeReal=mode, diagonalShift=shift, crossValidation=True)
    eq, eqo = jv.util.TDVP(ms, **kw), osolv.TDVP(mo, **kw)
    upd = to_np(eq(psi.get_parameters(), 0.0, psi=psi, hamiltonian=H, intStep=0))
    ref = eqo(opsi.get_parameters(), 0.0, psi=opsi, hamiltonian=Ho, intStep=0)
    assert rel(upd, ref) < 1e-7
    md, mo_ = eq.get_metadata(), eqo.metaData
    for k in ("tdvp_error", "tdvp_residual", "tdvp_residual_cross_validation_ratio",
              "tdvp_error_cross_validation_ratio"):
        assert abs(float(md[k]) - float(mo_[k])) <= 1e-6 * abs(float(mo_[k])) + 1e-14, (k, md[k], mo_[k])
    assert md["pinv_cutoff"] == mo_["pinv_cutoff"]


def test_subset_matches_oracle():
    """SampledObs.subset (stats.py:310-329) through jv_subset."""
    rng = np.random.default_rng(42)
    n, p = 301, 17
    for cpx in (False, True):
        O = rng.normal(size=(n, p)) + (1j * rng.normal(size=(n, p)) if cpx else 0)
        w = rng.uniform(size=n)
        w /= w.sum()
        o = jv.stats.SampledObs(torch.as_tensor(O[None], device="cuda"), torch.as_tensor(w[None], device="cuda"))
        oo = OObs(O, w)
        for sl in ((0, None, 2), (1, None, 2), (5, 200, 3), (0, 10, 1)):
            a, b = o.subset(*sl), oo.subset(*sl)
            assert rel(to_np(a._data), b._data) < 1e-13
            assert rel(to_np(a._mean), b._mean) < 1e-13
            assert rel(to_np(a._weights), b._weights) < 1e-14


# ---------------------------------------------------------------- 10x10 CNN workload
def _c3_net(seed):
    rng = np.random.default_rng(seed)
    shape = (10, 10)
    net = jv.nets.CNN(F=(3, 3), channels=(8, 8, 8), strides=(1, 1))
    psi = jv.NQS(net)
    psi(torch.zeros((1, 1) + shape, dtype=torch.int32))
    P = net.he_normal(psi.desc, rng)
    psi.set_parameters(P)
    opsi = onets.NQS(onets.CNN(shape, F=(3, 3), channels=(8, 8, 8), strides=(1, 1)), P)
    return shape, psi, opsi, P


def test_mc_cnn_bit_exact_tfim10x10():
    """The Metropolis kernel instance of the 10x10 / 3 x 8-channel CNN (window
    kernel, 8 channels, 4 waves per chain): 64 chains, two calls (walkers
    persist), configurations, log psi and acceptance counts against the oracle."""
    shape, psi, opsi, P = _c3_net(43)
    ms = jv.sampler.MCSampler(psi, shape, 17, updateProposer=jv.sampler.propose_spin_flip, numChains=64,
                              thermalizationSweeps=2, sweepSteps=100)
    mo = osamp.MCSampler(opsi, shape, 17, numChains=64, thermalizationSweeps=2, sweepSteps=100)
    for call in range(2):
        s, lp, p = ms.sample(numSamples=128)
        so, lpo, po = mo.sample(numSamples=128)
        np.testing.assert_array_equal(to_np(s)[0], so)
        assert rel(to_np(lp)[0], lpo) < 1e-12
        assert float(ms.acceptance_ratio()[0]) == mo.acceptance_ratio()
        assert 0.05 < mo.acceptance_ratio() < 0.95
        # the sampling kernel's log psi (from its maintained activations) equals
        # a fresh forward pass of the emitted configurations exactly
        assert torch.equal(lp, psi(s))


def test_minsr_overlapped_cholesky_route_tfim10x10():
    """MinSR.__call__ on the 10x10 CNN (P = 1240) with 2048 Metropolis samples:
    the default solver takes the P x P dual form with E_loc overlapped and, the
    spectrum being full rank, the Cholesky route (info.iterations == 1); the
    update equals the oracle's reference T form (minsr.py:53-80, the 4096 x 4096
    real-stacked T) on the same samples and E_loc.  The device E_loc is checked
    against the oracle on 128 of the samples (the whole set costs the oracle
    ~40 s)."""
    shape, psi, opsi, P = _c3_net(44)
    H, Ho = tfim_pair(shape, -1.0, -3.0)
    ms = _Recording(jv.sampler.MCSampler(psi, shape, 23, updateProposer=jv.sampler.propose_spin_flip,
                                         numChains=512, numSamples=2048))
    eq = jv.util.MinSR(ms, pinvTol=1e-10)
    assert eq.solver == "auto" and eq._overlapped(psi, 2048, None)
    upd = to_np(eq(psi.get_parameters(), 0.0, psi=psi, hamiltonian=H, intStep=0))
    assert eq.last_info.iterations == 1 and eq.last_info.rank == psi.numParameters == 1240
    s, lp, p = ms.last
    cs, pw = to_np(s)[0], to_np(p)[0]
    El = to_np(H.get_O_loc(s, psi, lp, 0.0))[0]
    sub = np.random.default_rng(0).choice(len(cs), 128, replace=False)
    assert rel(El[sub], Ho.get_O_loc(cs[sub], opsi, opsi(cs[sub]), 0.0)) < 1e-11
    Elo, Go = OObs(El, pw), OObs(opsi.gradients(cs), pw)
    ref = osolv.MinSR(None, pinvTol=1e-10).solve(Elo, Go, False)
    assert rel(upd, ref) < 1e-7
    assert abs(eq.get_energy_mean() - np.real(Elo.mean()[0])) <= 1e-11 * abs(Elo.mean()[0])
    assert abs(eq.get_energy_variance() - Elo.var()[0]) <= 1e-9 * Elo.var()[0]


# ---------------------------------------------------------------- 1D CpxRBM workloads
def _cpxrbm_case(L, H_, seed, visible=False):
    rng = np.random.default_rng(seed)
    onet = onets.CpxRBM((L,), numHidden=H_, bias=True, visibleBias=visible)
    P = 0.01 * rng.normal(size=onet.num_real_params())
    psi = make_psi(jv.nets.CpxRBM(numHidden=H_, bias=True, visibleBias=visible), (L,), P)
    return psi, onets.NQS(onet, P), P


def test_tdvp_step_l10_cpxrbm_exact():
    """One Euler step of the L = 10, 20-hidden CpxRBM (visible bias, N(0, 0.01^2)
    init) imaginary-time TDVP with the exact sampler (snrTol 2, pinvTol 1e-8,
    'real'): update, energy and metadata against the oracle, and the stepped
    parameters."""
    psi, opsi, P = _cpxrbm_case(10, 20, 0, visible=True)
    H, Ho = tfim_pair((10,), -1.0, -0.7)
    es, eso = jv.sampler.ExactSampler(psi, 10), osamp.ExactSampler(opsi, 10)
    kw = dict(snrTol=2, pinvTol=1e-8, rhsPrefactor=1.0, makeReal="real")
    eq, eqo = jv.util.TDVP(es, **kw), osolv.TDVP(eso, **kw)
    y, _ = jv.util.stepper.Euler(timeStep=1e-2).step(0, eq, psi.get_parameters(), hamiltonian=H, psi=psi)
    yo, _ = osolv.Euler(timeStep=1e-2).step(0, eqo, opsi.get_parameters(), hamiltonian=Ho, psi=opsi)
    assert rel(to_np(y) - P, yo - P) < 1e-7
    assert abs(eq.get_energy_mean() - eqo.get_energy_mean()) < 1e-11 * abs(eqo.get_energy_mean())
    assert eq.get_metadata()["pinv_cutoff"] == eqo.metaData["pinv_cutoff"]
    assert abs(eq.get_metadata()["tdvp_residual"] - eqo.metaData["tdvp_residual"]) < 1e-6 * eqo.metaData[
        "tdvp_residual"] + 1e-12


def test_tdvp_heun_rhs_l32_cpxrbm_16384_samples():
    """One Heun right-hand side of the L = 32, 64-hidden CpxRBM (Pc = 2112,
    Pr = 4224) real-time TDVP with 16384 Metropolis samples from 1024 chains
    (25 thermalisation sweeps of 32): E, Var E, F, S, the spectrum, the SNR
    and theta_dot against the oracle's TDVP.solve on the same samples.  'imag'
    S = i Im S0 has eigenvalues +-lambda(G^H G), non-degenerate here, so the
    SNR is basis independent and compared mode by mode where the eigenvalue is
    kept (|ev / ev_max| > 1e-14)."""
    psi, opsi, P = _cpxrbm_case(32, 64, 1)
    H, Ho = tfim_pair((32,), -1.0, -1.0)
    ms = _Recording(jv.sampler.MCSampler(psi, (32,), 3, updateProposer=jv.sampler.propose_spin_flip,
                                         numChains=1024, numSamples=16384, thermalizationSweeps=25,
                                         sweepSteps=32))
    eq = jv.util.TDVP(ms, rhsPrefactor=1.j, makeReal="imag")
    upd = to_np(eq(psi.get_parameters(), 0.0, psi=psi, hamiltonian=H, intStep=0))
    s, lp, p = ms.last
    cs, pw = to_np(s)[0], to_np(p)[0]
    mo = osamp.MCSampler(opsi, (32,), 3, numChains=1024)
    mo.globNumSamples = ms.globNumSamples
    eqo = osolv.TDVP(mo, rhsPrefactor=1.j, makeReal="imag")
    Elo = OObs(Ho.get_O_loc(cs, opsi, opsi(cs), 0.0), pw)
    Go = OObs(opsi.gradients(cs), pw)
    ref, res, cut = eqo.solve(Elo, Go)
    assert abs(complex(eq.ElocMean) - Elo.mean()[0]) < 1e-11 * abs(Elo.mean()[0])
    assert abs(eq.get_energy_variance() - Elo.var()[0]) < 1e-9 * Elo.var()[0]
    assert rel(to_np(eq.F), osolv.imagFun(eqo.F0)) < 1e-10
    assert rel(to_np(eq.get_S()), eqo.S) < 1e-10
    ev, evo = to_np(eq.ev), eqo.ev
    assert rel(ev, evo) < 1e-10
    kept = np.abs(evo / evo[-1]) > 1e-14
    assert rel(to_np(eq.snr)[kept], eqo.snr[kept]) < 1e-7
    assert eq.get_metadata()["pinv_cutoff"] == cut
    assert rel(upd, ref) < 1e-7
