"""The serving layer's HBM estimate (serving.footprint_bytes) against the
device memory one job really takes: each job is built and stepped once on a
fresh host thread (so the library's per-thread workspaces start empty), and the
drop in free device memory (torch's reservations, the library's workspaces and
solver handles) must stay within the estimate.  Device code of the libraries is
loaded once per process beforehand (it belongs to no job).  Cases cover every solve route:
MinSR's dual form (real: Cholesky route; complex: eigendecomposition), MinSR's
T form (real, and complex through the real 2N x 2N Cholesky form), the
holomorphic TDVP solve, and the mixed-precision dual form."""
import ctypes
import threading

import pytest
import torch

from helpers import tfim_pair

pytestmark = pytest.mark.gpu

jv = pytest.importorskip("vmc_jax_amd")
from vmc_jax_amd import _lib, serving  # noqa: E402


def _cnn(side, n, chains, precision="fp64"):
    net = jv.nets.CNN(F=(3, 3), channels=(4, 4), strides=(1, 1))
    psi = jv.NQS(net, seed=5, precision=precision)
    psi(torch.zeros((1, 1, side, side), dtype=torch.int32))
    H, _ = tfim_pair((side, side), -1.0, -3.0)
    ms = jv.sampler.MCSampler(psi, (side, side), 5, updateProposer=jv.sampler.propose_spin_flip, numChains=chains,
                              numSamples=n)
    return psi, H, jv.util.MinSR(ms, pinvTol=1e-10), jv.util.stepper.Euler(timeStep=5e-3), side * side


def _rbm(L, hidden, n, eq_kind):
    psi = jv.NQS(jv.nets.CpxRBM(numHidden=hidden, bias=True), seed=3)
    psi(torch.zeros((1, 1, L), dtype=torch.int32))
    H, _ = tfim_pair((L,), -1.0, -0.8)
    ms = jv.sampler.MCSampler(psi, (L,), 3, updateProposer=jv.sampler.propose_spin_flip, numChains=64,
                              numSamples=n)
    if eq_kind == "tdvp":
        eq = jv.util.TDVP(ms, rhsPrefactor=1.j, makeReal='imag')
    else:
        eq = jv.util.MinSR(ms, pinvTol=1e-10)
    return psi, H, eq, jv.util.stepper.Euler(timeStep=1e-3), L


CASES = {
    "minsr_dual_real": (lambda: _cnn(6, 1024, 128), "minsr", 1024, 128, "fp64"),
    "minsr_T_real": (lambda: _cnn(6, 128, 128), "minsr", 128, 128, "fp64"),
    "minsr_dual_mixed": (lambda: _cnn(6, 1024, 128, "mixed"), "minsr", 1024, 128, "mixed"),
    "minsr_dual_complex": (lambda: _rbm(8, 16, 1024, "minsr"), "minsr", 1024, 64, "fp64"),
    "minsr_T_complex": (lambda: _rbm(16, 32, 512, "minsr"), "minsr", 512, 64, "fp64"),
    "tdvp_holomorphic": (lambda: _rbm(16, 32, 1024, "tdvp"), "tdvp", 1024, 64, "fp64"),
}


@pytest.fixture(scope="module")
def warm():
    """Loads the solver libraries' and this library's kernels once (process-wide
    device code, not part of any job) by stepping every case on this thread."""
    _lib.call("jv_init", 0)
    for make, *_ in CASES.values():
        psi, H, eq, stepper, _ = make()
        stepper.step(0.0, eq, psi.get_parameters(), hamiltonian=H, psi=psi, numSamples=None)
    torch.cuda.synchronize()


@pytest.mark.parametrize("case", sorted(CASES))
def test_footprint_bounds_one_step(case, warm):
    make, solver, n, chains, precision = CASES[case]
    print("footprint case", case, flush=True)
    out = {}

    def body():
        try:
            torch.cuda.synchronize()
            torch.cuda.empty_cache()
            free0 = torch.cuda.mem_get_info()[0]
            res0 = torch.cuda.memory_reserved()
            psi, H, eq, stepper, sites = make()
            dp, _ = stepper.step(0.0, eq, psi.get_parameters(), hamiltonian=H, psi=psi, numSamples=None)
            psi.set_parameters(dp)
            torch.cuda.synchronize()
            out["used"] = free0 - torch.cuda.mem_get_info()[0]
            out["torch_reserved"] = torch.cuda.memory_reserved() - res0
            ws = ctypes.c_int64(0)
            _lib.call("jv_workspace_bytes", ctypes.byref(ws))
            out["workspace"] = ws.value
            out["estimate"] = serving.footprint_bytes(n, psi.numRealParameters, sites, not psi.realParams, solver,
                                                      chains, precision)
            out["energy"] = eq.get_energy_mean()
        except Exception as exc:  # reported on the main thread
            out["error"] = repr(exc)
        finally:
            # release this thread's workspaces and handles now: Python's join()
            # can return before the thread's C++ thread_local destructor has
            # freed them, which would then land in the next case's measurement
            _lib.call("jv_shutdown")

    t = threading.Thread(target=body)
    t.start()
    t.join()
    assert "error" not in out, out.get("error")
    print(case, {k: v for k, v in out.items()}, flush=True)
    assert out["energy"] == out["energy"]
    assert out["workspace"] <= out["used"] + (8 << 20)
    assert out["used"] <= out["estimate"], out
