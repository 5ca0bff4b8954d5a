"""Pins the FP64 oracle against known answers, closed forms and brute force.

Each check is chosen so that a plausible mistake in the restatement (dropped
term, wrong sign or index, transposed operand) fails it.
"""
import numpy as np
import pytest
import scipy.linalg

from oracle import nets as onets
from oracle import operator as oop
from oracle import sampler as osamp
from oracle import solvers as osolv
from oracle.philox import philox4x32_10, proposal_stream
from oracle.stats import SampledObs

# 16-weight CpxRBM(2 hidden, L=4) vector of tdvp_test.py:65-69.
WEIGHTS16 = np.array([0.23898957, 0.12614753, 0.19479055, 0.17325271, 0.14619853, 0.21392751,
                      0.19648707, 0.17103704, -0.15457255, 0.10954413, 0.13228065, -0.14935214,
                      -0.09963073, 0.17610707, 0.13386381, -0.14836467])
# tdvp_test.py:122-126
REF_ZZ = np.array([0.882762129306284, 0.8936168721790617, 0.9257753299594491, 0.9779836185039352,
                   1.0482156449061142, 1.1337654450614298, 1.231369697427413, 1.337354107391303,
                   1.447796176316155, 1.558696104640795, 1.666147269524912, 1.7664978782554912,
                   1.8564960156892512, 1.9334113379450693, 1.9951280521882777, 2.0402054805651546,
                   2.067904337137255, 2.078178742959828, 2.071635856483114, 2.049466698269522])


def tfim(L, J, h, lattice=None):
    H = oop.BranchFreeOperator()
    for l in range(L):
        H.add(oop.scal_opstr(J, (oop.Sz(l), oop.Sz((l + 1) % L))))
        H.add(oop.scal_opstr(h, (oop.Sx(l),)))
    return H


def kron_tfim(L, J, h):
    """Dense TFIM from Kronecker products; site i is bit i (LSB = site 0)."""
    X = np.array([[0, 1], [1, 0]], float)
    Z = np.diag([-1.0, 1.0])  # s=0 -> -1, s=1 -> +1 (branch_free.py:84-87)
    I = np.eye(2)

    def op_at(ops):
        M = np.array([[1.0]])
        for site in reversed(range(L)):  # kron order: most significant bit first
            M = np.kron(M, ops.get(site, I))
        return M
    H = np.zeros((2 ** L, 2 ** L))
    for l in range(L):
        H += J * op_at({l: Z, (l + 1) % L: Z})
        H += h * op_at({l: X})
    return H


# ---------------------------------------------------------------- philox
def test_philox_known_answers():
    # Random123 kat_vectors for philox4x32_10.
    assert [int(x) for x in philox4x32_10([0, 0, 0, 0], [0, 0])] == \
        [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]
    assert [int(x) for x in philox4x32_10([0xffffffff] * 4, [0xffffffff] * 2)] == \
        [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]
    assert [int(x) for x in philox4x32_10([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344],
                                          [0xa4093822, 0x299f31d0])] == \
        [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]


def test_proposal_stream_distribution():
    site, u = proposal_stream(7, 0, 3, 20000, 10)
    assert site.min() == 0 and site.max() == 9
    assert abs(np.bincount(site, minlength=10) / 20000 - 0.1).max() < 0.01
    assert 0 <= u.min() and u.max() < 1 and abs(u.mean() - 0.5) < 0.01


# ---------------------------------------------------------------- nets
def test_log_cosh_closed_form():
    x = np.array([0.3, -2.0, 1.5 + 0.7j, -0.2 - 3.1j, 40.0 + 1j, -40.0 + 0.5j])
    np.testing.assert_allclose(onets.log_cosh(x), np.log(np.cosh(x)), rtol=1e-13, atol=1e-13)


def _fd_check(net, P, s, holo, h=1e-6, tol=1e-7):
    psi = onets.NQS(net, P)
    G = psi.gradients(s)
    f0 = psi(s)
    for j in range(len(P)):
        e = np.zeros(len(P)); e[j] = h
        psi.set_parameters(P + e); fp = psi(s)
        psi.set_parameters(P - e); fm = psi(s)
        psi.set_parameters(P)
        fd = (fp - fm) / (2 * h)
        np.testing.assert_allclose(fd, G[:, j], atol=tol, rtol=tol, err_msg="param %d" % j)
    return f0, G


def test_cpxrbm_gradients_fd():
    rng = np.random.default_rng(1)
    net = onets.CpxRBM((5,), numHidden=3, bias=True, visibleBias=True)
    P = 0.3 * rng.normal(size=net.num_real_params())
    s = rng.integers(0, 2, size=(6, 5))
    _fd_check(net, P, s, True)


def test_rbm_gradients_fd():
    rng = np.random.default_rng(2)
    net = onets.RBM((5,), numHidden=3, bias=True)
    P = 0.3 * rng.normal(size=net.num_real_params())
    _fd_check(net, P, rng.integers(0, 2, size=(6, 5)), False)


def test_cnn_gradients_fd_2d():
    rng = np.random.default_rng(3)
    net = onets.CNN((4, 4), F=(3, 3), channels=(2, 3, 2), strides=(1, 1))
    P = 0.4 * rng.normal(size=net.num_real_params())
    _fd_check(net, P, rng.integers(0, 2, size=(3, 4, 4)), False)


def test_cnn_gradients_fd_1d():
    rng = np.random.default_rng(4)
    net = onets.CNN((6,), F=(3,), channels=(3, 2))
    P = 0.4 * rng.normal(size=net.num_real_params())
    _fd_check(net, P, rng.integers(0, 2, size=(3, 6)), False)


def test_cnn_param_layout_c3():
    net = onets.CNN((10, 10), F=(3, 3), channels=(8, 8, 8), strides=(1, 1))
    assert [l.name for l in net.leaves] == ["Conv_0/kernel", "Conv_1/bias", "Conv_1/kernel",
                                            "Conv_2/bias", "Conv_2/kernel"]
    assert net.num_real_params() == 1240 and np.isclose(net.nrm, np.sqrt(800))


def test_cnn_translation_invariance():
    """nets_test.py:29-46: all translates of a configuration give the same log psi."""
    rng = np.random.default_rng(5)
    net = onets.CNN((4, 4), F=(3, 3), channels=(3, 2, 5), strides=(1, 1))
    psi = onets.NQS(net, rng.normal(size=net.num_real_params()))
    S0 = np.array([[1, 0, 1, 1], [0, 1, 1, 1], [0, 0, 1, 0], [1, 0, 0, 1]])
    S = np.array([np.roll(np.roll(S0, -i, 0), -j, 1) for i in range(4) for j in range(4)])
    v = psi(S)
    assert np.max(np.abs(v - v[0])) < 1e-12


def test_cnn_conv_matches_direct_sum():
    """Direct 4-fold loop restatement of the circular VALID correlation."""
    rng = np.random.default_rng(6)
    net = onets.CNN((3, 4), F=(2, 3), channels=(2,), strides=(1, 1), firstLayerBias=True)
    P = rng.normal(size=net.num_real_params())
    leaves = net.unflatten(P)
    K = net.kernel(leaves, 0); b = net.biasv(leaves, 0)
    s = rng.integers(0, 2, size=(3, 4))
    x = 2 * s - 1
    out = 0.0
    for i in range(3):
        for j in range(4):
            for c in range(2):
                z = b[c] + sum(x[(i + a) % 3, (j + d) % 4] * K[a, d, 0, c] for a in range(2) for d in range(3))
                out += onets.log_cosh(z)
    assert np.isclose(net.forward(leaves, s[None])[0], out / np.sqrt(24), rtol=1e-13)


def test_holomorphic_layout():
    net = onets.CpxRBM((3,), numHidden=2, bias=True)
    psi = onets.NQS(net, np.arange(net.num_real_params()) * 0.01)
    G = psi.gradients(np.array([[0, 1, 1]]))
    # layout [g_b, i g_b, g_W, i g_W] (vqs.py:66-69)
    np.testing.assert_allclose(G[0, 2:4], 1j * G[0, 0:2])
    np.testing.assert_allclose(G[0, 10:16], 1j * G[0, 4:10])


# ---------------------------------------------------------------- operator
def test_tfim_dense_matches_kron():
    L = 6
    H = tfim(L, -1.0, -0.7)
    np.testing.assert_allclose(H.dense_matrix(L), kron_tfim(L, -1.0, -0.7), atol=1e-14)


def test_sy_sz_string_dense():
    """Sy(i) Sz(i+1) string: compare with kron (Sy in the s-basis: <s|Sy|s'>)."""
    L = 4
    H = oop.BranchFreeOperator()
    for i in range(L):
        H.add(oop.scal_opstr(2., (oop.Sx(i),)))
        H.add(oop.scal_opstr(2., (oop.Sy(i), oop.Sz((i + 1) % L))))
    D = H.dense_matrix(L)
    # branch_free.py:66-69: s=0 -> s'=1 with +i, s=1 -> s'=0 with -i; rows s, cols s'
    Y = np.array([[0, 1j], [-1j, 0]])
    X = np.array([[0, 1], [1, 0]]); Z = np.diag([-1.0, 1.0]); I = np.eye(2)

    def op_at(ops):
        M = np.array([[1.0 + 0j]])
        for site in reversed(range(L)):
            M = np.kron(M, ops.get(site, I))
        return M
    ref = sum(2 * op_at({i: X}) + 2 * op_at({i: Y}) @ op_at({(i + 1) % L: Z}) for i in range(L))
    np.testing.assert_allclose(D, ref, atol=1e-14)


def test_sp_nonzero_counts():
    """operator_test.py:48-68."""
    rng = np.random.default_rng(3)
    s = rng.integers(0, 2, size=(24, 4)).astype(np.int32)
    h = oop.BranchFreeOperator()
    h += 2. * oop.Sp(0)
    h += 2. * oop.Sp(1)
    h += 2. * oop.Sp(2)
    sp, m = h.get_s_primes(s)
    tmp = h.get_O_loc_unbatched(np.ones(24), np.ones(sp.shape[0]))
    assert np.sum(np.abs(tmp - 2. * np.sum(-(s[:, :3] - 1), axis=-1))) < 1e-7


def test_time_dependent_prefactor():
    """operator_test.py:70-93."""
    rng = np.random.default_rng(3)
    s = rng.integers(0, 2, size=(24, 4)).astype(np.int32)
    h = oop.BranchFreeOperator()
    f = lambda t: 2.0 * t
    for i in range(3):
        h.add(oop.scal_opstr(f, (oop.Sp(i),)))
    for t in [0.5, 2, 13.9]:
        sp, m = h.get_s_primes(s, t)
        tmp = h.get_O_loc_unbatched(np.ones(24), np.ones(sp.shape[0]))
        assert np.sum(np.abs(tmp - f(t) * np.sum(-(s[:, :3] - 1), axis=-1))) < 1e-7


def test_compaction_order_and_padding():
    """base.py:91-160: nonzero strings kept in order, padded with the last
    string's s' and m = 0."""
    h = oop.BranchFreeOperator()
    h += 1. * oop.Sp(0)
    h += 1. * oop.Sz(1)
    h += 1. * oop.Sp(2)
    s = np.array([[1, 0, 0], [0, 0, 1]], np.int32)
    sp, m = h.get_s_primes(s)
    # sample 0: Sp(0) zero (s0=1), Sz(1) -> -1, Sp(2) -> 1 ; sample 1: Sp(0) ok, Sz(1) -1, Sp(2) zero
    np.testing.assert_allclose(m, [[-1, 1], [1, -1]])
    np.testing.assert_array_equal(sp.reshape(2, 2, 3), [[[1, 0, 0], [1, 0, 1]], [[1, 0, 1], [0, 0, 1]]])


def test_eloc_matches_dense():
    rng = np.random.default_rng(8)
    L = 6
    H = tfim(L, -1.0, -0.7)
    net = onets.CpxRBM((L,), numHidden=4, bias=True)
    psi = onets.NQS(net, 0.2 * rng.normal(size=net.num_real_params()))
    es = osamp.ExactSampler(psi, L)
    basis, logpsi, p = es.sample()
    eloc = H.get_O_loc(basis, psi, logpsi)
    v = np.exp(logpsi)
    ref = (kron_tfim(L, -1.0, -0.7) @ v) / v
    np.testing.assert_allclose(eloc, ref, rtol=1e-12, atol=1e-12)


# ---------------------------------------------------------------- stats
def test_sampled_obs_known_answers():
    """stats_test.py:15-38."""
    p = np.ones(3) / 3
    o1 = SampledObs(np.array([1, 2, 3.]), p)
    o2 = SampledObs(np.array([[1, 4], [2, 5], [3, 7.]]), p)
    assert np.isclose(o1.mean()[0], 2.0) and np.isclose(o1.var()[0], 2. / 3)
    np.testing.assert_allclose(o2.covar(), [[2. / 3, 1], [1., 14. / 9]])
    np.testing.assert_allclose(o2.mean(), [2, 16. / 3])
    np.testing.assert_allclose(o1.covar(o2), [[2. / 3, 1.]])
    np.testing.assert_allclose(o1.covar(o2).ravel(), o1.covar_data(o2).mean())
    np.testing.assert_allclose(o1.covar_var(o2).ravel(), o1.covar_data(o2).var())
    O = o2._data
    np.testing.assert_allclose(o2.tangent_kernel(), O @ O.conj().T)


def test_subset():
    """stats_test.py:133-154."""
    N = 10
    obs = np.arange(N, dtype=float)[:, None]
    p = np.random.default_rng(123).uniform(size=N); p /= p.sum()
    o1 = SampledObs(obs, p)
    o2 = o1.subset(0, N // 2)
    assert np.isclose(o2.mean()[0], np.sum(obs[:N // 2, 0] * p[:N // 2]) / p[:N // 2].sum())
    o3 = SampledObs(obs[:N // 2], p[:N // 2] / p[:N // 2].sum())
    np.testing.assert_allclose(o3.covar(), o2.covar())


# ---------------------------------------------------------------- samplers
def test_exact_sampler_closed_form():
    net = onets.CpxRBM((4,), numHidden=2)
    psi = onets.NQS(net, WEIGHTS16)
    es = osamp.ExactSampler(psi, 4)
    b, lp, p = es.sample()
    assert b[5].tolist() == [1, 0, 1, 0]  # bit i of k is site i
    w = np.abs(np.exp(lp)) ** 2
    np.testing.assert_allclose(p, w / w.sum(), rtol=1e-13)
    _, _, p2 = es.sample()  # running lastNorm leaves p unchanged
    np.testing.assert_allclose(p2, p, rtol=1e-13)


def test_mcmc_histogram_matches_exact():
    """sampler_test.py:32-82 method (fewer samples, tolerance scaled)."""
    net = onets.CpxRBM((4,), numHidden=2)
    psi = onets.NQS(net, WEIGHTS16)
    pex = osamp.ExactSampler(psi, 4).sample()[2]
    mc = osamp.MCSampler(psi, (4,), seed=0, numChains=777)
    s, lp, p = mc.sample(numSamples=200000)
    assert s.shape[0] >= 200000
    k = (s * (1 << np.arange(4))).sum(1)
    pmc = np.bincount(k, weights=p, minlength=16)
    # 2e-3 at 500k samples in the reference; scaled by sqrt(500k / 200k)
    assert np.max(np.abs(pmc - pex)) < 3.2e-3
    np.testing.assert_allclose(lp, psi(s))


def test_distribute_sampling():
    assert osamp.distribute_sampling(1000, 1, 300) == (4, 1200)
    assert osamp.distribute_sampling(8192, 1, 512) == (16, 8192)


# ---------------------------------------------------------------- solvers / steppers
def test_adaptive_heun_vs_expm():
    """stepper_test.py:14-39."""
    N = 4
    np.random.seed(123)
    st = osolv.AdaptiveHeun()
    mat = np.random.rand(N, N) + 1.j * np.random.rand(N, N)
    y0 = np.random.rand(N)
    y = y0.copy()
    t = 0
    norm = lambda v: np.real(np.conj(v).dot(v))
    diffs = []
    for _ in range(100):
        y, dt = st.step(t, lambda yy, tt, **a: a['mat'].dot(yy), y, normFunction=norm, mat=mat)
        t += dt
        diffs.append(norm(y - scipy.linalg.expm(t * mat).dot(y0)) / N)
    assert max(diffs) < 1e-5


def test_pinv_hermitian_matches_numpy():
    rng = np.random.default_rng(9)
    A = rng.normal(size=(6, 3)); T = A @ A.T
    np.testing.assert_allclose(osolv.pinv_hermitian(T, 1e-10), np.linalg.pinv(T, rcond=1e-10, hermitian=True),
                               atol=1e-10)


def test_time_evolution_zz_golden():
    """tdvp_test.py:59-128: exact-sampler real-time TDVP, ZZ(t) curve < 1e-3."""
    from scipy.interpolate import interp1d
    L = 4
    psi = onets.NQS(onets.CpxRBM((L,), numHidden=2), WEIGHTS16)
    H = tfim(L, -1.0, -0.3)
    ZZ = oop.BranchFreeOperator()
    for l in range(L):
        ZZ.add((oop.Sz(l), oop.Sz((l + 1) % L)))
    es = osamp.ExactSampler(psi, L)
    stepper = osolv.AdaptiveHeun(timeStep=1e-3, tol=1e-5)
    eq = osolv.TDVP(es, snrTol=1, pinvTol=0.0, pinvCutoff=1e-8, rhsPrefactor=1.j, diagonalShift=0., makeReal='imag')
    t = 0
    times = [0]
    m = osolv.measure({'E': H, 'ZZ': ZZ}, psi, es)
    obs = [[m['E']['mean'][0], m['ZZ']['mean'][0]]]
    while t < 0.5:
        dp, dt = stepper.step(0, eq, psi.get_parameters(), hamiltonian=H, psi=psi, numSamples=0)
        psi.set_parameters(dp)
        t += dt
        times.append(t)
        m = osolv.measure({'E': [(H, t)], 'ZZ': ZZ}, psi, es)
        obs.append([m['E']['mean'][0], m['ZZ']['mean'][0]])
    obs = np.array(obs)
    assert np.max(np.abs((obs[:, 0] - obs[0, 0]) / obs[0, 0])) < 1e-3
    zz = interp1d(np.array(times), obs[:, 1])(np.arange(0, 0.5, 0.05))
    assert np.max(np.abs(zz - REF_ZZ[:len(zz)])) < 1e-3


def _gs_init(net, seed):
    rng = np.random.default_rng(seed)
    return 0.01 * rng.uniform(size=net.num_real_params())


@pytest.mark.parametrize("hx,exE", [(-1.3, -6.10160339), (-0.3, -4.09296160)])
def test_gs_search_tdvp(hx, exE):
    """tdvp_test.py:26-56."""
    L = 4
    net = onets.CpxRBM((L,), numHidden=6)
    psi = onets.NQS(net, _gs_init(net, 0))
    H = tfim(L, -1.0, hx)
    es = osamp.ExactSampler(psi, L)
    eq = osolv.TDVP(es, snrTol=1, pinvTol=0.0, pinvCutoff=1e-8, rhsPrefactor=1., diagonalShift=2, makeReal='real')
    osolv.ground_state_search(psi, H, eq, es, numSteps=100, stepSize=5e-2)
    E = osolv.measure({"energy": H}, psi, es)["energy"]["mean"][0]
    assert abs((E - exE) / exE) < 1e-3
    assert np.isclose(np.linalg.eigvalsh(kron_tfim(L, -1.0, hx))[0], exE, atol=1e-7)


@pytest.mark.parametrize("hx,exE", [(-1.3, -6.10160339), (-0.3, -4.09296160)])
def test_gs_search_minsr(hx, exE):
    """minsr_test.py:14-43."""
    L = 4
    net = onets.CpxRBM((L,), numHidden=8)
    psi = onets.NQS(net, _gs_init(net, 1))
    H = tfim(L, -1.0, hx)
    es = osamp.ExactSampler(psi, L)
    eq = osolv.MinSR(es, pinvTol=1e-6, diagonalShift=0.)
    osolv.ground_state_search(psi, H, eq, es, numSteps=200, stepSize=1e-2)
    E = osolv.measure({"energy": H}, psi, es)["energy"]["mean"][0]
    assert abs((E - exE) / exE) < 1e-3


def test_snr_identity():
    """tdvp_test.py:201-263: the covar_data/transform route equals the direct
    centred product for the SNR variance."""
    L = 4
    psi = onets.NQS(onets.CpxRBM((L,), numHidden=2), WEIGHTS16)
    H = tfim(L, -1.0, -0.3)
    mc = osamp.MCSampler(psi, (L,), seed=0, numChains=500)
    s, lp, p = mc.sample(numSamples=1000)
    El = H.get_O_loc(s, psi, lp, 0.0)
    G = psi.gradients(s)
    Eo, Go = SampledObs(El, p), SampledObs(G, p)
    S = osolv.imagFun(Go.covar())
    ev, V = np.linalg.eigh(S)
    Ec = El - p @ El
    Gc = G - p @ G
    EO = -(-1.) * (p * Ec)[:, None] * np.conj(Gc) * mc.globNumSamples
    EO = osolv.imagFun(EO) @ np.conj(V)
    rho_old = np.einsum("ij,ij,i->j", np.conj(EO - p @ EO), EO - p @ EO, np.ones(len(p)) / mc.globNumSamples)
    rho_new = Go.covar_data(Eo).transform(linearFun=np.conj(V).T, nonLinearFun=osolv.imagFun).var()
    np.testing.assert_allclose(rho_old, rho_new, rtol=1e-8, atol=1e-12)


def test_time_evolution_mc_zz_golden():
    """tdvp_test.py:131-198: real-time TDVP under Metropolis sampling (mu = 1,
    500 chains, 5000 samples per RHS, snrTol = 1, pinvTol = 1e-8, AdaptiveHeun
    tol 1e-4, crossValidation on); energy drift < 1e-1 and ZZ(t) within 2e-2 of
    the exact-sampler golden curve up to t = 0.2.  The bound is statistical:
    the reference fixes it for JAX's PRNGKey(0); with the Philox stream seed 0
    ends 0.0155 from the curve.  Runs with the SNR factor dropped (0.024) or
    inverted (snr / snrTol: 0.096) fail the bound, so this pins the presence and
    direction of the SNR soft cutoff (tdvp.py:198-205); its exponent is pinned
    by test_tdvp_regulariser_closed_form."""
    from scipy.interpolate import interp1d
    L = 4
    psi = onets.NQS(onets.CpxRBM((L,), numHidden=2), WEIGHTS16)
    H = tfim(L, -1.0, -0.3)
    ZZ = oop.BranchFreeOperator()
    for l in range(L):
        ZZ.add((oop.Sz(l), oop.Sz((l + 1) % L)))
    mc = osamp.MCSampler(psi, (L,), seed=0, numSamples=50000, numChains=500, mu=1)
    stepper = osolv.AdaptiveHeun(timeStep=1e-3, tol=1e-4)
    eq = osolv.TDVP(mc, snrTol=1, pinvTol=1e-8, rhsPrefactor=1.j, diagonalShift=0., makeReal='imag',
                    crossValidation=True)
    t = 0
    times = [0]
    m = osolv.measure({'E': H, 'ZZ': ZZ}, psi, mc)
    obs = [[m['E']['mean'][0], m['ZZ']['mean'][0]]]
    with np.errstate(divide="ignore"):
        while t < 0.2:
            dp, dt = stepper.step(0, eq, psi.get_parameters(), hamiltonian=H, psi=psi, numSamples=5000)
            psi.set_parameters(dp)
            t += dt
            times.append(t)
            m = osolv.measure({'E': [(H, t)], 'ZZ': ZZ}, psi, mc)
            obs.append([m['E']['mean'][0], m['ZZ']['mean'][0]])
    obs = np.array(obs)
    assert np.max(np.abs((obs[:, 0] - obs[0, 0]) / obs[0, 0])) < 1e-1
    zz = interp1d(np.array(times), obs[:, 1])(np.arange(0, 0.2, 0.05))
    assert np.max(np.abs(zz - REF_ZZ[:len(zz)])) < 2e-2
    assert np.isfinite(eq.metaData["tdvp_residual_cross_validation_ratio"])
    assert np.isfinite(eq.metaData["tdvp_error_cross_validation_ratio"])


def _mc_tdvp_case(seed=3, n=2000):
    L = 4
    psi = onets.NQS(onets.CpxRBM((L,), numHidden=2), WEIGHTS16)
    H = tfim(L, -1.0, -0.3)
    mc = osamp.MCSampler(psi, (L,), seed=seed, numChains=500)
    s, lp, p = mc.sample(numSamples=n)
    return mc, H.get_O_loc(s, psi, lp, 0.0), psi.gradients(s), p


def test_tdvp_regulariser_closed_form():
    """The regularised inverse of the TDVP docstring (tdvp.py:43-46, arXiv
    1912.08828): sigma_k^-1 / ((1 + (eps_SVD / (sigma_k / sigma_max))^6)
    (1 + (eps_SNR / SNR(rho_k))^6)), with S, F, the eigenbasis and the SNR built
    here directly from centred samples (SNR as in test_snr_identity), and the
    cutoff the oracle's loop stopped at; snrTol is set to the median SNR so that
    half the modes are suppressed.  A different exponent or an inverted ratio
    changes the update by far more than the tolerance."""
    mc, El, G, p = _mc_tdvp_case()
    Ec = El - p @ El
    Gc = G - p @ G
    S0 = np.einsum("n,ni,nj->ij", p, np.conj(Gc), Gc)
    F0 = -(1.j) * np.einsum("n,ni,n->i", p, np.conj(Gc), Ec)
    S, F = osolv.imagFun(S0), osolv.imagFun(F0)
    # S has degenerate +-lambda pairs: take the oracle's eigenbasis (the basis
    # inside a degenerate pair is not fixed by S) after checking S and the pairs
    eq = osolv.TDVP(mc, snrTol=1.0, pinvTol=1e-10, pinvCutoff=1e-8, rhsPrefactor=1.j, makeReal='imag')
    eq.solve(SampledObs(El, p), SampledObs(G, p))
    np.testing.assert_allclose(eq.S, S, atol=1e-13 * np.abs(S).max())
    ev, V = eq.ev, eq.V
    np.testing.assert_allclose(S @ V, V * ev[None, :], atol=1e-12 * np.abs(ev).max())
    VtF = np.conj(V).T @ F
    # per-sample projection rho_n = V^dagger q(-x conj(O_n - <O>) (E_n - <E>)) N; its weighted variance
    EO = osolv.imagFun(-(1.j) * np.conj(Gc) * Ec[:, None]) @ np.conj(V)
    rho_var = np.einsum("n,nj->j", p, np.abs(EO - p @ EO) ** 2)
    snr = np.sqrt(np.abs(mc.globNumSamples * np.abs(VtF) ** 2 / rho_var))
    rel_ev = np.abs(ev / ev[-1])
    kept = rel_ev > 1e-14  # null modes: invEv = 0, their SNR is a ratio of round-off
    snrTol = float(np.median(snr[kept]))
    eq.snrTol = snrTol
    upd, res, cut = eq.solve(SampledObs(El, p), SampledObs(G, p))
    np.testing.assert_allclose(eq.snr[kept], snr[kept], rtol=1e-8)
    snr = eq.snr
    inv = np.where(rel_ev > 1e-14, 1.0 / ev, 0.0)

    def expected(c, k_snr=6, k_ev=6):
        r = 1.0 / (1.0 + (c / rel_ev) ** k_ev) / (1.0 + (snrTol / snr) ** k_snr)
        return np.real(V @ (inv * r * VtF))
    ref = expected(cut)
    assert np.linalg.norm(upd - ref) <= 1e-9 * np.linalg.norm(ref)
    assert np.linalg.norm(expected(cut, k_snr=2) - ref) > 1e-3 * np.linalg.norm(ref)
    # a cutoff floor above the smallest kept eigenvalue (2.4e-4 of the largest) pins the other exponent
    eq.pinvTol, eq.pinvCutoff = 0.0, 1e-3
    upd, res, cut = eq.solve(SampledObs(El, p), SampledObs(G, p))
    assert cut == 1e-3
    ref = expected(cut)
    assert np.linalg.norm(upd - ref) <= 1e-9 * np.linalg.norm(ref)
    assert np.linalg.norm(expected(cut, k_ev=2) - ref) > 1e-3 * np.linalg.norm(ref)


def test_cross_validation_duplicated_samples():
    """tdvp.py:306-324 with every sample present twice in a row (equal
    weights): the even and odd halves both equal the full sample, so the
    half-sample update is the full one and both cross-validation ratios are 1."""
    mc, El, G, p = _mc_tdvp_case(n=1000)
    El2, G2, p2 = np.repeat(El, 2), np.repeat(G, 2, axis=0), np.repeat(p, 2) / 2
    eq = osolv.TDVP(mc, snrTol=1, pinvTol=1e-8, rhsPrefactor=1.j, makeReal='imag', crossValidation=True)
    Eo, Go = SampledObs(El2, p2), SampledObs(G2, p2)
    u_full, res, _ = eq.solve(Eo, Go)
    eq.ElocVar0 = eq.ElocVar
    eq.metaData = {"tdvp_error": eq._get_tdvp_error(u_full)}
    u_cv, res_cv = eq.cross_validate(Eo, Go)
    np.testing.assert_allclose(u_cv, u_full, rtol=1e-12, atol=0)
    assert abs(eq.metaData["tdvp_residual_cross_validation_ratio"] - 1.0) < 1e-9
    # tdvp_error = |1 + q / Var E| with q close to -Var E: the cancellation costs
    # about log10(1 / tdvp_error) digits
    assert abs(eq.metaData["tdvp_error_cross_validation_ratio"] - 1.0) < 1e-6


def test_cross_validation_halves_direct():
    """The cross-validation ratios from halves built directly from the raw
    observations (SampledObs of the even / odd samples with renormalised
    weights) instead of SampledObs.subset."""
    mc, El, G, p = _mc_tdvp_case(n=1000)
    eq = osolv.TDVP(mc, snrTol=1, pinvTol=1e-8, rhsPrefactor=1.j, makeReal='imag', crossValidation=True)
    u_full, res, _ = eq.solve(SampledObs(El, p), SampledObs(G, p))
    evar0 = eq.ElocVar
    S0, F0 = eq.S0, eq.F0
    terr0 = np.abs(1. + np.real(u_full @ S0 @ u_full - 2. * np.real(u_full @ F0)) / evar0)
    eq.ElocVar0 = evar0
    eq.metaData = {"tdvp_error": eq._get_tdvp_error(u_full)}
    eq.cross_validate(SampledObs(El, p), SampledObs(G, p))
    ref = osolv.TDVP(mc, snrTol=1, pinvTol=1e-8, rhsPrefactor=1.j, makeReal='imag')
    pe, po = p[0::2] / p[0::2].sum(), p[1::2] / p[1::2].sum()
    u1, _, _ = ref.solve(SampledObs(El[0::2], pe), SampledObs(G[1 - 1::2], pe))
    S2, F2 = ref.get_tdvp_equation(SampledObs(El[1::2], po), SampledObs(G[1::2], po))
    r_ratio = np.linalg.norm(S2 @ u1 - F2) / np.linalg.norm(F2) / res
    t_ratio = np.abs(1. + np.real(u1 @ ref.S0 @ u1 - 2. * np.real(u1 @ ref.F0)) / evar0) / terr0
    assert abs(eq.metaData["tdvp_residual_cross_validation_ratio"] / r_ratio - 1) < 1e-8
    assert abs(eq.metaData["tdvp_error_cross_validation_ratio"] / t_ratio - 1) < 1e-8
