"""Benchmark: TDVP (MinSR) steps/s on the 2D TFIM 10x10 / CNN workload.

Workload (one "step" = one Euler step of the MinSR equation): H = -sum ZZ -
3 sum X on a periodic 10x10 lattice; real CNN with 3 conv layers of 8 channels,
3x3 kernels, circular padding, log-cosh, sum / sqrt(800) (1240 parameters,
He-normal init); 8192 Metropolis samples from 512 chains per step
(thermalizationSweeps = sweepSteps = 10, the sampler defaults); MinSR with
pinvTol = 1e-10 solved through the 8192 x 8192 T = O O^T in FP64; Euler
dt = 5e-3.  Everything runs in FP64 on the GPU; inputs are resident on the
device when the timed region starts.

Prints one JSON line (rank 0) with steps/s, samples/s, the roofline of the
dominant hand-written kernel (from HIP events recorded live on its stream) and
the CPU baseline (the FP64 oracle, timed on a bounded sample of the same step).
"""
import argparse
import ctypes
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

SHAPE = (10, 10)
CHANNELS = (8, 8, 8)
CHAINS = 512
SAMPLES = 8192
DT = 5e-3
PINV_TOL = 1e-10
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6


def tfim_strings(op, shape, J=-1.0, h=-3.0):
    H = op.BranchFreeOperator()
    Lx, Ly = shape
    for i in range(Lx):
        for j in range(Ly):
            s = i * Ly + j
            for nb in (((i + 1) % Lx) * Ly + j, i * Ly + (j + 1) % Ly):
                H.add(op.scal_opstr(J, (op.Sz(s), op.Sz(nb))))
            H.add(op.scal_opstr(h, (op.Sx(s),)))
    return H


def build_gpu(mode):
    import torch
    import vmc_jax_amd as jv
    net = jv.nets.CNN(F=(3, 3), channels=CHANNELS, strides=(1, 1))
    psi = jv.NQS(net)
    psi(torch.zeros((1, 1) + SHAPE, dtype=torch.int32))
    psi.set_parameters(net.he_normal(psi.desc, np.random.default_rng(0)))
    H = tfim_strings(jv.operator, SHAPE)
    sampler = jv.sampler.MCSampler(psi, SHAPE, 1234, updateProposer=jv.sampler.propose_spin_flip,
                                   numChains=CHAINS, numSamples=SAMPLES)
    eq = jv.util.MinSR(sampler, pinvTol=PINV_TOL, solver=mode)
    stepper = jv.util.stepper.Euler(timeStep=DT)
    return jv, psi, H, sampler, eq, stepper


def gpu_step(psi, H, eq, stepper):
    dp, _ = stepper.step(0.0, eq, psi.get_parameters(), hamiltonian=H, psi=psi, numSamples=None)
    psi.set_parameters(dp)


def timer(jv, name):
    tot = ctypes.c_double(0.0)
    cnt = ctypes.c_int64(0)
    jv._lib.call("jv_timer_read", name.encode(), ctypes.byref(tot), ctypes.byref(cnt))
    return tot.value, cnt.value


def run_gpu(steps, warmup, mode):
    import torch
    jv, psi, H, sampler, eq, stepper = build_gpu(mode)
    for _ in range(warmup):
        gpu_step(psi, H, eq, stepper)
    torch.cuda.synchronize()
    jv._lib.call("jv_timer_reset")
    jv._lib.call("jv_timer_enable", 1)
    energies = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        gpu_step(psi, H, eq, stepper)
        energies.append(eq.get_energy_mean())
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    jv._lib.call("jv_timer_enable", 0)
    kt = {k: timer(jv, k) for k in ("gram", "eigh", "eloc", "mc_sweep", "net_grad", "net_forward", "gemm")}
    return elapsed, kt, energies, eq


def cpu_baseline(n_sub=1024, chains=512):
    """FP64 oracle timed on a bounded sample of the same step: the oracle runs
    sampling (full-network re-evaluation per proposal), local energies with
    materialised s', gradients and the T-form MinSR solve on n_sub samples from
    `chains` chains; each phase is scaled to 8192 samples by its complexity
    (sampling and sweeps by samples per chain, E_loc / gradients linearly, the
    T Gram quadratically, the eigendecomposition cubically)."""
    from oracle import nets as onets, operator as oop, sampler as osamp, solvers as osolv
    from oracle.stats import SampledObs
    onet = onets.CNN(SHAPE, F=(3, 3), channels=CHANNELS, strides=(1, 1))
    P = np.zeros(onet.num_real_params())
    rng = np.random.default_rng(0)
    for leaf_i, leaf in enumerate(onet.leaves):
        pass
    # He-normal on kernels (same distribution as the GPU run, independent draw)
    off = 0
    for leaf in onet.leaves:
        if leaf.name.endswith("kernel"):
            fan_in = leaf.shape[0] * leaf.shape[1] * leaf.shape[2]
            P[off:off + leaf.size] = rng.normal(size=leaf.size) * math.sqrt(2.0 / fan_in)
        off += leaf.size
    psi = onets.NQS(onet, P, batchSize=4096)
    H = tfim_strings(oop, SHAPE)
    spc_full = SAMPLES // CHAINS
    spc_sub = max(1, n_sub // chains)
    ms = osamp.MCSampler(psi, SHAPE, 0, numChains=chains, thermalizationSweeps=10, sweepSteps=10)
    t = {}
    t0 = time.perf_counter(); s, lp, p = ms.sample(numSamples=spc_sub * chains); t["sampling"] = time.perf_counter() - t0
    t0 = time.perf_counter(); El = H.get_O_loc(s, psi, lp, 0.0); t["eloc"] = time.perf_counter() - t0
    t0 = time.perf_counter(); G = psi.gradients(s); t["grad"] = time.perf_counter() - t0
    Eo, Go = SampledObs(El, p), SampledObs(G, p)
    D = np.real(Go._data)
    t0 = time.perf_counter(); T = D @ D.T; t["gram"] = time.perf_counter() - t0
    t0 = time.perf_counter(); ev, V = np.linalg.eigh(T); t["eigh"] = time.perf_counter() - t0
    n = len(p)
    scale_n = SAMPLES / n
    sweeps_sub = (10 + spc_sub) * 10
    sweeps_full = (10 + spc_full) * 10
    est = (t["sampling"] * sweeps_full / sweeps_sub * (CHAINS / chains) + (t["eloc"] + t["grad"]) * scale_n
           + t["gram"] * scale_n ** 2 + t["eigh"] * scale_n ** 3)
    cores = len(os.sched_getaffinity(0))
    return {"value": 1.0 / est, "unit": "steps/s", "cores": cores, "kind": "port",
            "sample": ("FP64 numpy oracle, one MinSR step on %d samples from %d chains (10x10 TFIM, same CNN); "
                       "phases scaled to 8192 samples (sampling x%.2f, E_loc+gradients x%.0f, T Gram x%.0f, "
                       "eigh x%.0f); measured %s" % (n, chains, sweeps_full / sweeps_sub * CHAINS / chains,
                                                      scale_n, scale_n ** 2, scale_n ** 3,
                                                      {k: round(v, 3) for k, v in t.items()}))}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--mode", default="dual", choices=["T", "dual"])
    ap.add_argument("--t-steps", type=int, default=3, help="steps of the reference T-form solve reported beside")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (every box has one MI355X); "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    elapsed, kt, energies, eq = run_gpu(args.steps, args.warmup, args.mode)
    other = None
    if args.t_steps > 0:
        om = "T" if args.mode == "dual" else "dual"
        e2, kt2, _, _ = run_gpu(args.t_steps, 1, om)
        other = {"solver_form": om, "value": args.t_steps / e2, "unit": "steps/s",
                 "ms_per_step": 1000.0 * e2 / args.t_steps,
                 "kernel_ms_per_step": {k: v[0] / args.t_steps for k, v in kt2.items() if v[1] > 0}}
    ms_step = 1000.0 * elapsed / args.steps
    steps_s = args.steps / elapsed
    # dominant hand-written kernel
    cands = {k: v for k, v in kt.items() if k in ("gram", "eloc", "mc_sweep", "net_grad") and v[1] > 0}
    dom = max(cands, key=lambda k: cands[k][0])
    dom_ms = cands[dom][0] / cands[dom][1]
    if dom == "gram":
        n, p = SAMPLES, 1240
        flops = float(n) * (n + 64) * p  # triangle-only tiles: N(N+64)P multiply-adds x 2 / 2
        roof = {"bound": "mfma", "achieved": flops / (dom_ms * 1e-3) / 1e12, "peak": FP64_PEAK_TFLOPS,
                "unit": "TFLOP/s", "traffic": None}
    else:
        flops_per = {"eloc": 2.03e11, "mc_sweep": 3.3e10, "net_grad": 3 * 2 * 122400.0 * SAMPLES}[dom]
        roof = {"bound": "mfma", "achieved": flops_per / (dom_ms * 1e-3) / 1e12, "peak": FP64_PEAK_TFLOPS,
                "unit": "TFLOP/s", "traffic": None}
    roof["frac"] = roof["achieved"] / roof["peak"]
    roof["kernel"] = dom
    out = {
        "metric": "TDVP steps/s + samples/s (TFIM 10x10 CNN) at |dE/E|<=1e-6; served p99 latency",
        "value": steps_s,
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (He-normal CNN parameters, Metropolis samples of the network itself)",
        "config": {"workload": "TFIM 10x10 periodic, h=3; CNN 3x(8ch,3x3) log-cosh, P=1240; 512 chains, "
                               "8192 samples/step; MinSR pinvTol=1e-10 (%s form); Euler dt=5e-3" % (
            "dual form: eigendecomposition of the 1240x1240 D^T D, the same update as the 8192x8192 T solve"
            if args.mode == "dual" else "T"),
                   "samples_per_step": SAMPLES, "solver_form": args.mode},
        "samples_per_s": steps_s * SAMPLES,
        "kernel_ms_per_step": {k: v[0] / args.steps for k, v in kt.items() if v[1] > 0},
        "energy_per_site_last": energies[-1] / 100.0 if energies else None,
        "roofline": roof,
    }
    if other is not None:
        out["other_solver_form"] = other
    if not args.no_cpu_baseline:
        try:
            out["cpu_baseline"] = cpu_baseline()
        except Exception as e:  # report, never hide the GPU number
            out["cpu_baseline"] = {"error": repr(e)}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
