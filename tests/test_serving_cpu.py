"""Host-side tests of the serving loop: admission by footprint, FIFO order,
refusal, slot limits and failure isolation (no GPU: jobs are host objects)."""
import threading
import time

import pytest

from vmc_jax_amd import serving


class FakeEq:
    def __init__(self, e, fail_at=None):
        self.e = e
        self.calls = 0
        self.fail_at = fail_at

    def __call__(self, y, t, **kw):
        self.calls += 1
        if self.fail_at is not None and self.calls > self.fail_at:
            raise RuntimeError("injected kernel failure")
        time.sleep(0.002)
        return -0.1 * y

    def get_energy_mean(self):
        return self.e


class FakePsi:
    def __init__(self):
        self.p = 1.0

    def get_parameters(self):
        return self.p

    def set_parameters(self, p):
        self.p = p


class Euler:
    def step(self, t, f, y, **kw):
        return y + 0.1 * f(y, t, **kw), 0.1


def job(i, steps, fp, arrival=0.0, e=-1.0, fail_at=None, log=None):
    def make():
        if log is not None:
            log.append(i)
        return FakePsi(), None, FakeEq(e, fail_at), Euler()
    return serving.Job(i, make, steps, fp, arrival)


def test_admission_respects_budget_and_slots():
    log = []
    srv = serving.Server(max_slots=2, hbm_budget=100, stream_factory=lambda: None)
    jobs = [job(i, 5, 60 if i == 0 else 30, arrival=0.001 * i, log=log) for i in range(5)]
    for j in jobs:
        assert srv.submit(j)
    peak = {"slots": 0, "mem": 0}
    orig = srv._admit

    def watch(now):
        r = orig(now)
        peak["slots"] = max(peak["slots"], len(srv.active))
        peak["mem"] = max(peak["mem"], srv.used)
        return r
    srv._admit = watch
    rep = srv.run()
    assert rep["jobs_done"] == 5 and rep["steps"] == 25
    assert peak["slots"] <= 2 and peak["mem"] <= 100
    assert log == sorted(log)  # FIFO admission by arrival
    assert all(j.energies == [-1.0] * 5 for j in jobs)


def test_refusal_of_oversized_job():
    srv = serving.Server(max_slots=2, hbm_budget=100, stream_factory=lambda: None)
    big = job(0, 3, 1000)
    assert not srv.submit(big)
    assert big.state == "refused" and "exceeds" in big.error
    assert srv.run()["jobs_refused"] == 1


def test_failure_is_isolated():
    srv = serving.Server(max_slots=3, hbm_budget=100, stream_factory=lambda: None)
    a = job(0, 6, 10)
    b = job(1, 6, 10, fail_at=2)
    c = job(2, 6, 10)
    for j in (a, b, c):
        srv.submit(j)
    rep = srv.run()
    assert b.state == "failed" and "injected" in b.error and b.steps_done == 2
    assert a.state == "done" and c.state == "done" and a.steps_done == c.steps_done == 6
    assert rep["jobs_failed"] == 1 and rep["jobs_done"] == 2
    assert srv.used == 0


def test_nan_energy_marks_job_failed():
    srv = serving.Server(max_slots=1, hbm_budget=100, stream_factory=lambda: None)
    j = job(0, 4, 10, e=float("nan"))
    srv.submit(j)
    srv.run()
    assert j.state == "failed" and "non-finite" in j.error


def test_latency_percentiles_reported():
    srv = serving.Server(max_slots=4, hbm_budget=1000, stream_factory=lambda: None)
    for i in range(6):
        srv.submit(job(i, 4, 10, arrival=0.002 * i))
    rep = srv.run()
    assert rep["p50_step_ms"] > 0 and rep["p99_step_ms"] >= rep["p50_step_ms"]
    assert rep["steps_per_s"] > 0


def test_footprint_grows_with_problem_size():
    small = serving.footprint_bytes(2048, 1240, 36, False, "minsr", 256)
    large = serving.footprint_bytes(16384, 1240, 144, False, "minsr", 1024)
    tdvp = serving.footprint_bytes(16384, 4224, 32, True, "tdvp", 1024)
    assert small < large < tdvp
