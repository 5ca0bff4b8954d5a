"""Mixed-precision form: the compensated BF16 Gram against FP64 products.

Bound: each entry of S = D^dagger D within 3e-5 of sum_n |D_ni| |D_nj| (the
split x = bf16(x) + bf16(x - bf16(x)) leaves 2^-18 |x| per factor and the
dropped lo*lo product 2^-18 |x y|; FP32 accumulation adds ~sqrt(slice) 2^-24
relative), and much closer in the norm.
"""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

_lib = pytest.importorskip("vmc_jax_amd._lib")


def _gram_mixed(D, f32, cpx):
    n = D.shape[0]
    p = D.shape[1] // (2 if cpx else 1)
    t = torch.as_tensor(D.astype(np.float32 if f32 else np.float64), device="cuda").contiguous()
    S = torch.empty((p, p * (2 if cpx else 1)), dtype=torch.float64, device="cuda")
    _lib.call("jv_gram_pp_mixed", _lib.ptr(t), int(f32), n, p, int(cpx), _lib.ptr(S), _lib.stream_ptr())
    torch.cuda.synchronize()
    S = S.cpu().numpy()
    return S[:, 0::2] + 1j * S[:, 1::2] if cpx else S


@pytest.mark.parametrize("n,p,f32,cpx", [
    (8192, 1240, True, False),   # 10x10 CNN dual Gram (FP32 data, vector loads)
    (1000, 333, True, False),    # ragged: p % 4 != 0 (element loads), n % 64 != 0
    (4096, 250, False, False),   # FP64 data
    (16384, 211, False, True),   # complex holomorphic Gram (interleaved, 2p % 4 != 0)
    (3000, 130, False, True),
    (1, 5, True, False),         # a single sample
])
def test_gram_pp_mixed_bound(n, p, f32, cpx):
    rng = np.random.default_rng(n + p)
    if cpx:
        Dc = rng.normal(size=(n, p)) + 1j * rng.normal(size=(n, p))
        Dc *= rng.uniform(0.01, 10.0, size=p)  # columns of different scales
        D = np.empty((n, 2 * p))
        D[:, 0::2], D[:, 1::2] = Dc.real, Dc.imag
        ref = Dc.conj().T @ Dc
        A = np.abs(Dc)
    else:
        D = rng.normal(size=(n, p)) * rng.uniform(0.01, 10.0, size=p)
        if f32:
            D = D.astype(np.float32).astype(np.float64)  # the values the kernel sees
        ref = D.T @ D
        A = np.abs(D)
    S = _gram_mixed(D, f32, cpx)
    bound = 3e-5 * (A.T @ A)
    assert np.all(np.abs(S - ref) <= bound + 1e-300)
    assert np.linalg.norm(S - ref) <= 1e-5 * np.linalg.norm(ref)
    # exactly Hermitian (the lower triangle is the mirror)
    assert np.array_equal(S, S.conj().T)


# ---------------------------------------------------------------- mixed network kernels (10x10 CNN workload)
def _c3_pair(precision, seed=3):
    import vmc_jax_amd as jv
    from oracle import nets as onets
    rng = np.random.default_rng(seed)
    shape = (10, 10)
    net = jv.nets.CNN(F=(3, 3), channels=(8, 8, 8), strides=(1, 1))
    psi = jv.NQS(net, precision=precision)
    psi(torch.zeros((1, 1) + shape, dtype=torch.int32))
    P = net.he_normal(psi.desc, rng)
    psi.set_parameters(P)
    opsi = onets.NQS(onets.CNN(shape, F=(3, 3), channels=(8, 8, 8), strides=(1, 1)), P)
    return jv, shape, psi, opsi, P


def test_mixed_eloc_and_gradients_tfim10x10():
    """FP32 window E_loc and FP32 O of the 10x10 CNN against the FP64 oracle on
    the same configurations: per-sample |dE_loc| <= 1e-5 |E_loc| + 1e-5, the
    mean to |dE/E| <= 1e-6 (the metric's accuracy condition), O to 1e-5
    relative (Frobenius)."""
    import sys, os
    sys.path.insert(0, os.path.dirname(__file__))
    from helpers import tfim_pair
    jv, shape, psi, opsi, P = _c3_pair("mixed")
    H, Ho = tfim_pair(shape, -1.0, -3.0)
    rng = np.random.default_rng(5)
    n = 512
    cs = rng.integers(0, 2, size=(n,) + shape).astype(np.int32)
    s = torch.as_tensor(cs[None], device="cuda")
    lp = psi(s)
    El = H.get_O_loc(s, psi, lp, 0.0).cpu().numpy().reshape(-1)
    olp = opsi(cs)
    Eo = Ho.get_O_loc(cs, opsi, olp, 0.0)
    assert np.all(np.abs(El - Eo) <= 1e-5 * np.abs(Eo) + 1e-5)
    assert abs(El.mean() - Eo.mean()) <= 1e-6 * abs(Eo.mean())
    G = psi.gradients(s, layout=0)
    assert G.dtype == torch.float32
    Go = np.real(opsi.gradients(cs))  # compact real layout of a real network
    g = G.cpu().numpy().reshape(n, -1).astype(np.float64)
    assert np.linalg.norm(g - Go) <= 1e-5 * np.linalg.norm(Go)


@pytest.mark.parametrize("shape,ch", [((6, 6), (8, 8, 8)), ((12, 12), (8, 8)), ((7, 9), (4, 8, 4))])
def test_mixed_eloc_incremental_windows(shape, ch):
    """The mixed form's incremental window update on other lattices and
    depths: 36 and 63 flips (a last chunk shorter than the others), two
    layers, uneven channel counts (padded to CM = 8) and a non-square lattice;
    per-sample |dE_loc| <= 1e-5 |E_loc| + 1e-5 against the FP64 oracle."""
    import sys, os
    sys.path.insert(0, os.path.dirname(__file__))
    import vmc_jax_amd as jv
    from oracle import nets as onets
    from helpers import tfim_pair
    rng = np.random.default_rng(sum(shape) + len(ch))
    net = jv.nets.CNN(F=(3, 3), channels=ch, strides=(1, 1))
    psi = jv.NQS(net, precision="mixed")
    psi(torch.zeros((1, 1) + shape, dtype=torch.int32))
    P = net.he_normal(psi.desc, rng)
    psi.set_parameters(P)
    opsi = onets.NQS(onets.CNN(shape, F=(3, 3), channels=ch, strides=(1, 1)), P)
    H, Ho = tfim_pair(shape, -1.0, -3.0)
    cs = rng.integers(0, 2, size=(96,) + shape).astype(np.int32)
    s = torch.as_tensor(cs[None], device="cuda")
    El = H.get_O_loc(s, psi, psi(s), 0.0).cpu().numpy().reshape(-1)
    Eo = Ho.get_O_loc(cs, opsi, opsi(cs), 0.0)
    assert np.all(np.abs(El - Eo) <= 1e-5 * np.abs(Eo) + 1e-5)


def test_mixed_metropolis_runs_tfim10x10():
    """The FP32 Metropolis kernel: emitted configurations are valid spins, log psi
    of the emitted samples equals the FP64 forward, and the acceptance rate is
    within 0.03 of the FP64 kernel's on the same network and Philox stream."""
    jv, shape, psi_m, _, P = _c3_pair("mixed")
    _, _, psi_d, _, _ = _c3_pair("fp64")
    rates = []
    for psi in (psi_m, psi_d):
        ms = jv.sampler.MCSampler(psi, shape, 17, updateProposer=jv.sampler.propose_spin_flip, numChains=128,
                                  numSamples=2048)
        s, lp, p = ms.sample()
        torch.cuda.synchronize()
        cs = s.cpu().numpy()
        assert set(np.unique(cs)) <= {0, 1}
        assert torch.allclose(lp, psi_d(s), rtol=0, atol=1e-12)
        rates.append(ms.acceptance_ratio())
    assert abs(float(rates[0]) - float(rates[1])) <= 0.03


def test_mixed_sample_logpsi_on_side_stream():
    """MCSampler.sample(logpsi_stream=...) of the mixed form: the same chains
    (configurations), the same FP64 log psi (computed on the other stream) and
    the same weights as a plain call from a sampler with the same key."""
    jv, shape, psi, _, P = _c3_pair("mixed")
    a = jv.sampler.MCSampler(psi, shape, 23, updateProposer=jv.sampler.propose_spin_flip, numChains=128,
                             numSamples=1024)
    b = jv.sampler.MCSampler(psi, shape, 23, updateProposer=jv.sampler.propose_spin_flip, numChains=128,
                             numSamples=1024)
    side = torch.cuda.Stream()
    for _ in range(2):  # the second call continues the chains
        s1, l1, p1 = a.sample()
        s2, l2, p2 = b.sample(logpsi_stream=side)
        torch.cuda.current_stream().wait_stream(side)
        torch.cuda.synchronize()
        assert torch.equal(s1, s2) and torch.equal(l1, l2) and torch.equal(p1, p2)


def test_mixed_minsr_step_tfim10x10():
    """One MinSR right-hand side of the mixed form (FP32 O, compensated-BF16
    Gram, FP64 solve) against the oracle's FP64 T-form MinSR on the same 2048
    samples: the energy to 1e-6 relative; the update is a least-squares
    solution, so its quality is judged by the residual ||D u + e|| of the FP64
    data, within 1e-4 of the FP64 solution's; the update itself moves by the
    Gram error times the kept spectrum's conditioning (pinvTol 1e-4 bounds the
    latter by 1e4)."""
    import sys, os
    sys.path.insert(0, os.path.dirname(__file__))
    from helpers import tfim_pair
    from oracle import sampler as osamp, solvers as osolv
    jv, shape, psi, opsi, P = _c3_pair("mixed")
    H, Ho = tfim_pair(shape, -1.0, -3.0)

    class Rec:
        def __init__(self, s):
            self.s, self.last = s, None

        def sample(self, *a, **k):
            self.last = self.s.sample(*a, **k)
            return self.last

    ms = Rec(jv.sampler.MCSampler(psi, shape, 21, updateProposer=jv.sampler.propose_spin_flip, numChains=128,
                                  numSamples=2048))
    eq = jv.util.MinSR(ms, pinvTol=1e-4)
    upd = eq(psi.get_parameters(), 0.0, psi=psi, hamiltonian=H, intStep=0).cpu().numpy()
    s, lp, p = ms.last
    cs = s[0].cpu().numpy()
    pw = p[0].cpu().numpy()
    from oracle.stats import SampledObs as OObs
    Eo = OObs(Ho.get_O_loc(cs, opsi, opsi(cs), 0.0), pw)
    Go = OObs(opsi.gradients(cs), pw)
    ref = osolv.MinSR(None, pinvTol=1e-4).solve(Eo, Go, False)
    e_o = complex(Eo.mean()[0])
    assert abs(eq.get_energy_mean() - e_o.real) <= 1e-6 * abs(e_o)
    D, e = np.real(Go._data), np.real(Eo._data).reshape(-1)
    res, res_ref = np.linalg.norm(D @ upd + e), np.linalg.norm(D @ ref + e)
    rel_upd = np.linalg.norm(upd - ref) / np.linalg.norm(ref)
    print("mixed MinSR: residual %.6e vs %.6e, update rel err %.3e" % (res, res_ref, rel_upd))
    assert abs(res - res_ref) <= 1e-4 * res_ref
    assert rel_upd <= 5e-2


@pytest.mark.parametrize("n,p,cpx", [(8192, 1240, False), (1000, 333, False), (333, 7, True), (16384, 211, True),
                                     (5, 300, False), (1, 1, False), (4097, 130, True)])
def test_gram_pp_fp64(n, p, cpx):
    """jv_gram_pp (FP64 matrix cores, 128 x 128 upper tiles, sample slices
    summed in FP64) against numpy: 1e-13 relative in norm, exactly Hermitian."""
    rng = np.random.default_rng(3 * n + p)
    if cpx:
        Dc = (rng.normal(size=(n, p)) + 1j * rng.normal(size=(n, p))) * rng.uniform(0.1, 3.0, size=p)
        D = np.empty((n, 2 * p))
        D[:, 0::2], D[:, 1::2] = Dc.real, Dc.imag
        ref = Dc.conj().T @ Dc
    else:
        D = rng.normal(size=(n, p)) * rng.uniform(0.1, 3.0, size=p)
        ref = D.T @ D
    t = torch.as_tensor(D, device="cuda").contiguous()
    S = torch.empty((p, p * (2 if cpx else 1)), dtype=torch.float64, device="cuda")
    _lib.call("jv_gram_pp", _lib.ptr(t), n, p, int(cpx), _lib.ptr(S), _lib.stream_ptr())
    torch.cuda.synchronize()
    S = S.cpu().numpy()
    S = S[:, 0::2] + 1j * S[:, 1::2] if cpx else S
    assert np.linalg.norm(S - ref) <= 1e-13 * np.linalg.norm(ref)
    assert np.array_equal(S, S.conj().T)
