"""Batched field sweep (vmc_jax_amd.batch): every instance's samples, log psi,
O_loc and O_k from the one-launch-per-phase batch equal its own MCSampler /
get_O_loc / gradients calls bit for bit (same Philox key and call index, same
device routines), on the first pass and on a replay of the captured graph
after a parameter update; two Euler steps of the batch equal two steps of the
instances run one by one."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

jv = pytest.importorskip("vmc_jax_amd")
from vmc_jax_amd import batch as jb  # noqa: E402

from helpers import tfim_pair, xxz_pair  # noqa: E402


def _spec(kind, L, alpha, seed):
    rng = np.random.default_rng(seed)
    if kind == "tdvp":
        psi = jv.NQS(jv.nets.CpxRBM(numHidden=alpha * L, bias=True), seed=seed)
    else:
        psi = jv.NQS(jv.nets.RBM(numHidden=alpha * L, bias=True), seed=seed)
    psi(torch.zeros((1, 1, L), dtype=torch.int32))
    psi.set_parameters(0.05 * rng.normal(size=psi.numRealParameters))
    H = tfim_pair((L,), -1.0, -0.8)[0] if kind == "tdvp" else xxz_pair(L, 0.7)[0]
    return psi, H


def _make(kind, psi, H, sampler=None):
    if kind == "tdvp":
        return jv.util.TDVP(sampler, snrTol=2, pinvTol=1e-8, rhsPrefactor=1., makeReal='real')
    return jv.util.MinSR(sampler, pinvTol=1e-8)


KINDS = [("tdvp", 8, 2, 11), ("minsr", 10, 1, 12), ("tdvp", 12, 1, 13)]
CHAINS, SAMPLES = 64, 512


def _batched(tmp_path, monkeypatch, graph=True):
    monkeypatch.setenv("VMC_JAX_AMD_TUNE_CACHE", str(tmp_path / "tune.json"))
    inst = []
    for kind, L, a, seed in KINDS:
        psi, H = _spec(kind, L, a, seed)
        inst.append(jb.Instance(psi, H, _make(kind, psi, H), 1e-2, key=seed, numChains=CHAINS, numSamples=SAMPLES))
    return jb.BatchedSweep(inst, workers=2, tune=True, graph=graph)


def _alone():
    out = []
    for kind, L, a, seed in KINDS:
        psi, H = _spec(kind, L, a, seed)
        ms = jv.sampler.MCSampler(psi, (L,), seed, updateProposer=jv.sampler.propose_spin_flip, numChains=CHAINS,
                                  numSamples=SAMPLES)
        out.append((kind, psi, H, ms))
    return out


def _check_pass(b, alone):
    for ins, (kind, psi, H, ms) in zip(b.inst, alone):
        s, lp, p = ms.sample()
        E = H.get_O_loc(s, psi, lp, 0.0)
        G = psi.gradients(s, layout=ins.layout)
        torch.cuda.synchronize()
        assert torch.equal(ins.configs, s.reshape(ins.configs.shape))
        assert torch.equal(ins.logpsi, psi.logpsi_complex(s.reshape(ins.configs.shape)))
        assert torch.equal(ins.eloc, E.reshape(-1).to(torch.complex128))
        assert torch.equal(ins.grad, G.reshape(ins.grad.shape))
        assert torch.equal(ins.counters, ms.counters)
        assert torch.equal(ins.walkers, ms.states)


def test_batch_pass_equals_alone(tmp_path, monkeypatch):
    b = _batched(tmp_path, monkeypatch)
    alone = _alone()
    try:
        assert b.tuned_from == "trials"
        torch.cuda.current_stream().wait_event(b.sample_phase())
        _check_pass(b, alone)
        # a parameter update in place, then a replay of the captured graph
        for ins, (kind, psi, H, ms) in zip(b.inst, alone):
            d = 0.01 * torch.arange(psi.numRealParameters, dtype=torch.float64, device="cuda") / psi.numRealParameters
            ins.psi.params.add_(d)
            psi.set_parameters(psi.get_parameters() + d)
        torch.cuda.current_stream().wait_event(b.sample_phase())
        _check_pass(b, alone)
    finally:
        b.close()
    # the tuned shapes come from the cache the second time
    b2 = _batched(tmp_path, monkeypatch, graph=False)
    try:
        assert b2.tuned_from == "cache"
    finally:
        b2.close()


def test_batch_steps_equal_alone(tmp_path, monkeypatch):
    """Two Euler steps of the batch (graph + threaded solves) against the
    instances stepped one at a time through TDVP / MinSR.__call__."""
    b = _batched(tmp_path, monkeypatch)
    alone = _alone()
    try:
        for _ in range(2):
            b.step()
            for kind, psi, H, ms in alone:
                eq = _make(kind, psi, H, ms)
                st = jv.util.stepper.Euler(timeStep=1e-2)
                if kind == "minsr":
                    eq.overlap = False
                y, _ = st.step(0.0, eq, psi.get_parameters(), hamiltonian=H, psi=psi)
                psi.set_parameters(y)
        torch.cuda.synchronize()
        for ins, (kind, psi, H, ms) in zip(b.inst, alone):
            a, r = ins.psi.params.cpu().numpy(), psi.params.cpu().numpy()
            assert np.linalg.norm(a - r) <= 1e-12 * np.linalg.norm(r), kind
    finally:
        b.close()
