"""GPU serving: jobs on concurrent streams give the same results alone and in
the mix (bit-identical energies and parameters), and the server reports
latency percentiles."""
import numpy as np
import pytest
import torch

from helpers import tfim_pair

pytestmark = pytest.mark.gpu

jv = pytest.importorskip("vmc_jax_amd")
from vmc_jax_amd import serving  # noqa: E402


def rbm_job(job_id, L, seed, steps):
    def make():
        psi = jv.NQS(jv.nets.CpxRBM(numHidden=L, bias=True), seed=seed)
        psi(torch.zeros((1, 1, L), dtype=torch.int32))
        H, _ = tfim_pair((L,), -1.0, -0.8)
        ms = jv.sampler.MCSampler(psi, (L,), seed, updateProposer=jv.sampler.propose_spin_flip, numChains=64,
                                  numSamples=1024)
        eq = jv.util.TDVP(ms, rhsPrefactor=1.0, makeReal='real', diagonalShift=0.1)
        return psi, H, eq, jv.util.stepper.Euler(timeStep=0.02)
    return serving.Job(job_id, make, steps, serving.footprint_bytes(1024, 4 * L * L, L, True, "tdvp", 64))


def cnn_job(job_id, side, seed, steps):
    def make():
        net = jv.nets.CNN(F=(3, 3), channels=(4, 4), strides=(1, 1))
        psi = jv.NQS(net, seed=seed)
        psi(torch.zeros((1, 1, side, side), dtype=torch.int32))
        H, _ = tfim_pair((side, side), -1.0, -3.0)
        ms = jv.sampler.MCSampler(psi, (side, side), seed, updateProposer=jv.sampler.propose_spin_flip,
                                  numChains=128, numSamples=1024)
        eq = jv.util.MinSR(ms, pinvTol=1e-10, solver="auto")
        return psi, H, eq, jv.util.stepper.Euler(timeStep=5e-3)
    return serving.Job(job_id, make, steps, serving.footprint_bytes(1024, 200, side * side, False, "minsr", 128))


def run(jobs, slots):
    srv = serving.Server(max_slots=slots)
    for j in jobs:
        assert srv.submit(j)
    return srv.run(), jobs


def test_same_result_alone_and_in_the_mix():
    _, (a_alone,) = run([rbm_job(0, 8, 3, 4)], 1)
    _, (c_alone,) = run([cnn_job(1, 6, 5, 4)], 1)
    rep, mix = run([rbm_job(0, 8, 3, 4), cnn_job(1, 6, 5, 4), rbm_job(2, 10, 7, 4), cnn_job(3, 8, 9, 4)], 4)
    assert rep["jobs_done"] == 4, [j.error for j in mix]
    assert mix[0].energies == a_alone.energies
    assert torch.equal(mix[0].params, a_alone.params)
    assert mix[1].energies == c_alone.energies
    assert torch.equal(mix[1].params, c_alone.params)
    assert rep["p99_step_ms"] >= rep["p50_step_ms"] > 0


def test_admission_queues_when_memory_is_short():
    jobs = [cnn_job(i, 6, 11 + i, 2) for i in range(3)]
    srv = serving.Server(max_slots=3, hbm_budget=jobs[0].footprint + 1)  # room for one job at a time
    for j in jobs:
        srv.submit(j)
    rep = srv.run()
    assert rep["jobs_done"] == 3
    starts = sorted(j.admitted_at for j in jobs)
    ends = sorted(j.finished_at for j in jobs)
    assert starts[1] >= ends[0] and starts[2] >= ends[1]  # strictly one at a time


def rbm_heun_job(job_id, L, seed, steps):
    def make():
        psi = jv.NQS(jv.nets.CpxRBM(numHidden=L, bias=True), seed=seed)
        psi(torch.zeros((1, 1, L), dtype=torch.int32))
        H, _ = tfim_pair((L,), -1.0, -0.5)
        ms = jv.sampler.MCSampler(psi, (L,), seed, updateProposer=jv.sampler.propose_spin_flip, numChains=64,
                                  numSamples=1024)
        eq = jv.util.TDVP(ms, rhsPrefactor=1.j, makeReal='imag')
        return psi, H, eq, jv.util.stepper.Heun(timeStep=1e-3)
    return serving.Job(job_id, make, steps, serving.footprint_bytes(1024, 4 * L * L, L, True, "tdvp", 64))


def test_shared_batch_server_equals_threads():
    """RBM jobs joined into the shared batch (one launch per phase for all of
    them, rebuilt as jobs arrive and finish) give bit-identical energies and
    parameters to the same jobs on their own threads; a CNN job runs beside
    the batch on its own thread."""
    def jobs():
        return [rbm_job(0, 8, 3, 4), rbm_heun_job(1, 10, 4, 3), cnn_job(2, 6, 5, 3), rbm_job(3, 12, 6, 2)]
    srv = serving.Server(max_slots=4)
    ref = jobs()
    for j in ref:
        srv.submit(j)
    rep_t = srv.run()
    srv = serving.Server(max_slots=4, batch_rbm=True, workers=2)
    got = jobs()
    for j in got:
        srv.submit(j)
    rep_b = srv.run()
    assert rep_t["jobs_done"] == rep_b["jobs_done"] == 4, ([j.error for j in ref], [j.error for j in got])
    for a, b in zip(ref, got):
        assert a.energies == b.energies, a.job_id
        assert torch.equal(a.params, b.params), a.job_id
    assert rep_b["p99_step_ms"] >= rep_b["p50_step_ms"] > 0


def test_served_workload_sizes_alone_and_in_the_mix():
    """Jobs at the served workload's upper sizes: an L = 64 complex RBM
    ground-state job (MinSR, 16384 samples from 1024 chains; Pc = 4160, so the
    complex dual Gram's eigendecomposition) and a 12 x 12 TFIM job with the
    3-layer, 8-channel CNN (MinSR, 2048 samples): energies and parameters
    bit-identical alone and in the mix."""
    def big_rbm(job_id):
        L = 64

        def make():
            psi = jv.NQS(jv.nets.CpxRBM(numHidden=L, bias=True), seed=21)
            psi(torch.zeros((1, 1, L), dtype=torch.int32))
            H, _ = tfim_pair((L,), -1.0, -1.0)
            ms = jv.sampler.MCSampler(psi, (L,), 21, updateProposer=jv.sampler.propose_spin_flip, numChains=1024,
                                      numSamples=16384)
            return psi, H, jv.util.MinSR(ms, pinvTol=1e-8), jv.util.stepper.Euler(timeStep=1e-2)
        pr = 2 * (L * L + L)
        return serving.Job(job_id, make, 2, serving.footprint_bytes(16384, pr, L, True, "minsr", 1024))

    def big_cnn(job_id):
        side = 12

        def make():
            net = jv.nets.CNN(F=(3, 3), channels=(8, 8, 8), strides=(1, 1))
            psi = jv.NQS(net, seed=22)
            psi(torch.zeros((1, 1, side, side), dtype=torch.int32))
            psi.set_parameters(net.he_normal(psi.desc, np.random.default_rng(22)))  # the workload's init
            H, _ = tfim_pair((side, side), -1.0, -3.0)
            ms = jv.sampler.MCSampler(psi, (side, side), 22, updateProposer=jv.sampler.propose_spin_flip,
                                      numChains=256, numSamples=2048)
            return psi, H, jv.util.MinSR(ms, pinvTol=1e-10), jv.util.stepper.Euler(timeStep=5e-3)
        return serving.Job(job_id, make, 2, serving.footprint_bytes(2048, 1240, side * side, False, "minsr", 256))

    print("served sizes: alone", flush=True)
    _, (r_alone,) = run([big_rbm(0)], 1)
    _, (c_alone,) = run([big_cnn(1)], 1)
    assert r_alone.state == c_alone.state == "done", (r_alone.error, c_alone.error)
    print("served sizes: mix", flush=True)
    rep, mix = run([big_rbm(0), big_cnn(1)], 2)
    assert rep["jobs_done"] == 2, [j.error for j in mix]
    assert mix[0].energies == r_alone.energies and len(r_alone.energies) == 2
    assert torch.equal(mix[0].params, r_alone.params)
    assert mix[1].energies == c_alone.energies
    assert torch.equal(mix[1].params, c_alone.params)
