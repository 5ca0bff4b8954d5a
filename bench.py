"""Benchmark: TDVP / MinSR steps/s of jVMC's VMC inner loop on one MI355X.

Default workload (the metric's configuration, ``--workload c3``): H = -sum ZZ -
3 sum X on a periodic 10x10 lattice; real CNN with 3 conv layers of 8 channels,
3x3 kernels, circular padding, log-cosh, sum / sqrt(800) (1240 parameters,
He-normal init); 8192 Metropolis samples from 512 chains per step
(thermalizationSweeps = sweepSteps = 10, the sampler defaults); MinSR with
pinvTol = 1e-10; Euler dt = 5e-3.  One "step" = one Euler step: sampling,
local energies, O_k, the MinSR solve and the parameter update.  Everything
runs in FP64 on the GPU; inputs are resident on the device when the timed
region starts.

The MinSR pseudo-inverse is computed in its dual form by default: the nonzero
spectrum of the 8192 x 8192 T = O O^T equals that of the 1240 x 1240 O^T O, and
-O^T pinv(T) E = -pinv(O^T O) O^T E with the same relative cutoff, so the
update is the reference's (the parity tests compare both forms).  While no
eigenvalue of O^T O lies at or below the cutoff (the first ~16 steps from the
random initial network), the pseudo-inverse is the inverse and the solve is a
Cholesky solve; later steps diagonalise O^T O.  ``minsr_cholesky_steps``
reports how many timed steps took the Cholesky route.  The reference's T form
(8192^2 Gram + FP64 eigendecomposition of T) is timed for a few steps and
reported beside it.

``--workload c1``: 1D TFIM L=10, h=0.7, complex RBM alpha=2 with visible bias,
exact sampler, imaginary-time TDVP (Euler dt=1e-2, snrTol=2, pinvTol=1e-8).
``--workload c2``: 1D TFIM L=32, h=1, complex RBM alpha=2, 1024 chains x 16384
samples (25 thermalisation sweeps, sweep length 32), real-time TDVP with Heun
dt=1e-3.
``--workload c4``: the batched field sweep (64 independent TFIM instances of
uneven size and parameter count, one launch per phase for all of them, replayed
from a captured HIP graph; solves on worker threads).
``--workload c5``: served jobs (Poisson arrivals of ground-state and real-time
jobs, continuous admission, RBM jobs joined into one shared batch).
``--workload gram``: the Gram kernels alone (FP64 matrix cores and the
compensated BF16 form) at the 10x10 CNN's dual size (8192 x 1240, real) and the
L=32 complex RBM's holomorphic size (16384 x 2112, complex).
``--workload sym``: 1D TFIM L=16, h=1, a SymNet of a complex RBM alpha=2 over
the 64-element orbit (translation, reflection, spin flip), 1024 chains x 4096
samples, imaginary-time TDVP (Euler dt=1e-2).

Prints one JSON line with steps/s, samples/s, per-kernel times, the roofline of
the dominant hand-written kernel (from HIP events recorded live on its stream)
and the CPU baseline (the FP64 oracle on a bounded sample of the same step);
the same oracle leg checks |dE/E| and the MinSR update against the FP64 oracle
on fresh samples of the final state (`accuracy`).
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

FP64_PEAK_TFLOPS = 78.6
BF16_PEAK_TFLOPS = 2500.0  # dense MFMA
HBM_PEAK_GBS = 8000.0
METRIC = "TDVP steps/s + samples/s (TFIM 10x10 CNN) at |dE/E|<=1e-6; served p99 latency"


def tfim_strings(op, shape, J=-1.0, h=-3.0):
    """H = J sum ZZ + h sum X on a periodic chain or square lattice."""
    H = op.BranchFreeOperator()
    if len(shape) == 1:
        L = shape[0]
        for l in range(L):
            H.add(op.scal_opstr(J, (op.Sz(l), op.Sz((l + 1) % L))))
            H.add(op.scal_opstr(h, (op.Sx(l),)))
        return H
    Lx, Ly = shape
    for i in range(Lx):
        for j in range(Ly):
            s = i * Ly + j
            for nb in (((i + 1) % Lx) * Ly + j, i * Ly + (j + 1) % Ly):
                H.add(op.scal_opstr(J, (op.Sz(s), op.Sz(nb))))
            H.add(op.scal_opstr(h, (op.Sx(s),)))
    return H


class Workload:
    def __init__(self, name, mode, precision="fp64"):
        import torch
        import vmc_jax_amd as jv
        self.jv = jv
        self.name = name
        rng = np.random.default_rng(0)
        if name == "c3":
            self.shape, self.samples = (10, 10), 8192
            net = jv.nets.CNN(F=(3, 3), channels=(8, 8, 8), strides=(1, 1))
            psi = jv.NQS(net, precision=precision)
            psi(torch.zeros((1, 1) + self.shape, dtype=torch.int32))
            psi.set_parameters(net.he_normal(psi.desc, rng))
            self.H = tfim_strings(jv.operator, self.shape, -1.0, -3.0)
            self.sampler = jv.sampler.MCSampler(psi, self.shape, 1234, updateProposer=jv.sampler.propose_spin_flip,
                                                numChains=512, numSamples=self.samples)
            self.eq = jv.util.MinSR(self.sampler, pinvTol=1e-10, solver=mode)
            self.stepper = jv.util.stepper.Euler(timeStep=5e-3)
            self.desc = ("TFIM 10x10 periodic, h=3; CNN 3x(8ch,3x3) log-cosh, P=1240; 512 chains, 8192 samples/step; "
                         "MinSR pinvTol=1e-10 (%s); Euler dt=5e-3" %
                         ("dual form on the 1240x1240 O^T O, same update as the 8192x8192 T solve: Cholesky while no "
                          "eigenvalue falls below the cutoff, else eigendecomposition"
                          if mode == "dual" else "T form: 8192x8192 T Gram + FP64 eigendecomposition"))
        elif name == "c2":
            self.shape, self.samples = (32,), 16384
            net = jv.nets.CpxRBM(numHidden=64, bias=True)
            psi = jv.NQS(net, precision=precision)
            psi(torch.zeros((1, 1) + self.shape, dtype=torch.int32))
            psi.set_parameters(0.01 * rng.normal(size=psi.numRealParameters))
            self.H = tfim_strings(jv.operator, self.shape, -1.0, -1.0)
            self.sampler = jv.sampler.MCSampler(psi, self.shape, 1234, updateProposer=jv.sampler.propose_spin_flip,
                                                numChains=1024, numSamples=self.samples, thermalizationSweeps=25,
                                                sweepSteps=32)
            self.eq = jv.util.TDVP(self.sampler, rhsPrefactor=1.j, makeReal='imag')
            self.stepper = jv.util.stepper.Heun(timeStep=1e-3)
            self.desc = ("TFIM L=32 periodic, h=1; CpxRBM alpha=2 (2112 complex params); 1024 chains, 16384 samples, "
                         "25 therm. sweeps, sweep 32; real-time TDVP (makeReal='imag'); Heun dt=1e-3")
        elif name == "c1":
            self.shape, self.samples = (10,), 1024
            net = jv.nets.CpxRBM(numHidden=20, bias=True, visibleBias=True)
            psi = jv.NQS(net)
            psi(torch.zeros((1, 1) + self.shape, dtype=torch.int32))
            psi.set_parameters(0.01 * rng.normal(size=psi.numRealParameters))
            self.H = tfim_strings(jv.operator, self.shape, -1.0, -0.7)
            self.sampler = jv.sampler.ExactSampler(psi, 10)
            self.eq = jv.util.TDVP(self.sampler, snrTol=2, pinvTol=1e-8, rhsPrefactor=1., makeReal='real')
            self.stepper = jv.util.stepper.Euler(timeStep=1e-2)
            self.desc = ("TFIM L=10 periodic, h=0.7; CpxRBM alpha=2 with visible bias (230 complex params); exact "
                         "sampler (1024 states); imaginary-time TDVP snrTol=2 pinvTol=1e-8; Euler dt=1e-2")
        elif name == "sym":
            self.shape, self.samples = (16,), 4096
            orbit = jv.util.symmetries.get_orbit_1D(16, "translation", "reflection", "spinflip")
            net = jv.nets.CpxRBM(numHidden=32, bias=True)
            psi = jv.NQS(net, orbit=orbit)
            psi(torch.zeros((1, 1) + self.shape, dtype=torch.int32))
            psi.set_parameters(0.01 * rng.normal(size=psi.numRealParameters))
            self.H = tfim_strings(jv.operator, self.shape, -1.0, -1.0)
            self.sampler = jv.sampler.MCSampler(psi, self.shape, 1234, updateProposer=jv.sampler.propose_spin_flip,
                                                numChains=1024, numSamples=self.samples)
            self.eq = jv.util.TDVP(self.sampler, snrTol=2, pinvTol=1e-8, rhsPrefactor=1., makeReal='real')
            self.stepper = jv.util.stepper.Euler(timeStep=1e-2)
            self.desc = ("TFIM L=16 periodic, h=1; SymNet(CpxRBM alpha=2, 544 complex params) over the 64-element "
                         "orbit (translation, reflection, spin flip); 1024 chains, 4096 samples; imaginary-time TDVP; "
                         "Euler dt=1e-2")
        else:
            raise ValueError(name)
        self.psi = psi

    def step(self):
        dp, _ = self.stepper.step(0.0, self.eq, self.psi.get_parameters(), hamiltonian=self.H, psi=self.psi,
                                  numSamples=None)
        self.psi.set_parameters(dp)


def timer(jv, name):
    tot = ctypes.c_double(0.0)
    cnt = ctypes.c_int64(0)
    jv._lib.call("jv_timer_read", name.encode(), ctypes.byref(tot), ctypes.byref(cnt))
    return tot.value, cnt.value


KERNELS = ("gram", "gemm", "eigh", "tridiag", "stedc", "ormtr", "trd_apply", "cholesky", "chol_inverse", "eloc",
           "mc_sweep", "net_grad",
           "net_forward", "sym_mc", "sym_eloc", "sym_combine")


def run_gpu(workload, steps, warmup, mode, overlap=True, precision="fp64"):
    import torch
    w = Workload(workload, mode, precision)
    if hasattr(w.eq, "overlap"):
        w.eq.overlap = overlap
    jv = w.jv
    for _ in range(warmup):
        w.step()
    torch.cuda.synchronize()
    jv._lib.call("jv_timer_reset")
    jv._lib.call("jv_timer_enable", 1)
    energies = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    routes = []
    outcomes = []
    for _ in range(steps):
        w.step()
        # a non-blocking solve's energy and route are read after the loop
        # (computed in the step all the same); other solves' right away
        last = w.eq.last_outcome() if hasattr(w.eq, "last_outcome") else None
        if last is not None:
            outcomes.append(last)
        else:
            energies.append(w.eq.get_energy_mean())
            info = getattr(w.eq, "last_info", None)
            routes.append(int(info.iterations) if info is not None else 0)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    for o in outcomes:
        energies.append(o.energy_mean())
        routes.append(int(o.resolve().iterations))
    jv._lib.call("jv_timer_enable", 0)
    kt = {k: timer(jv, k) for k in KERNELS}
    w.cholesky_steps = sum(routes) if isinstance(w.eq, jv.util.MinSR) else None
    return w, elapsed, kt, energies


def measured_traffic():
    """HBM bytes per launch (read + written) of the hand-written kernels on this
    workload: the `traffic_bytes_per_launch` table of the committed counter
    summary (rocprofv3 FETCH_SIZE / WRITE_SIZE passes, scaled by a copy kernel
    of known size), or None when it is absent."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic_tfim10x10_cnn.json")
    try:
        with open(path) as f:
            return json.load(f)["traffic_bytes_per_launch"]
    except (OSError, KeyError, ValueError):
        return None


# Per step of the 10x10 CNN workload (8192 samples, 101 strings of which 100
# single flips, 512 chains x 260 proposals):
#   counted: the reference algorithm's flops (a full 244,800-flop forward per
#            s' or proposal; Gram 2 N P^2 on the computed triangle of tiles);
#   executed: the multiply-adds the window kernels actually issue (per sample a
#            full forward, 244,800 flop, plus per flip the 9 + 25 + 49 window
#            outputs of 8 channels: 2 x 43,272 flop; per proposal the same
#            window), without the log-cosh evaluations (~45 FP64 operations
#            each, 664 per flip) and address arithmetic.
FLIP_FLOP = 2.0 * (9 * 9 * 8 + 25 * 72 * 8 + 49 * 72 * 8)
COUNTED = {"eloc": 2.03e11, "mc_sweep": 3.3e10, "net_grad": 3 * 244800.0 * 8192}
EXECUTED = {"eloc": 8192 * (244800.0 + 100 * FLIP_FLOP), "mc_sweep": 512 * 260 * FLIP_FLOP,
            "net_grad": 3 * 244800.0 * 8192}
BOUND = {"eloc": "valu", "mc_sweep": "valu", "net_grad": "valu", "gram": "mfma"}


def roofline_c3(kt, steps, mode, traffic=None):
    """Roofline of the dominant hand-written kernel (largest time per step).
    `achieved` uses the counted flops (the metric's algorithmic count);
    `executed_frac` the flops the kernel issues.  The window kernels run on the
    FP64 VALU (bound "valu"), whose peak equals the FP64 MFMA peak on MI355X
    (78.6 TFLOP/s).  `traffic`: HBM bytes per launch from the committed
    FETCH_SIZE / WRITE_SIZE counter passes of this workload."""
    per_step = dict(COUNTED)
    per_step["gram"] = (8192.0 * 8256.0 * 1240.0) if mode == "T" else (1240.0 * 1280.0 * 8192.0)
    cands = {k: v for k, v in kt.items() if k in per_step and v[1] > 0}
    dom = max(cands, key=lambda k: cands[k][0])
    total_ms, launches = cands[dom]
    ms = total_ms / launches
    flops = per_step[dom] * steps / launches
    ach = flops / (ms * 1e-3) / 1e12
    exe = EXECUTED.get(dom, per_step[dom]) * steps / launches / (ms * 1e-3) / 1e12
    return {"bound": BOUND.get(dom, "mfma"), "achieved": ach, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
            "frac": ach / FP64_PEAK_TFLOPS, "executed_tflops": exe, "executed_frac": exe / FP64_PEAK_TFLOPS,
            "traffic": (traffic or {}).get(dom) if mode == "dual" else None,
            "kernel": dom, "avg_launch_ms": ms, "launches": launches}


def gram_bench(reps=20):
    """S = D^dagger D alone: `jv_gram_pp` (FP64 matrix cores) and
    `jv_gram_pp_mixed` (compensated BF16, FP32 and FP64 inputs), timed with HIP
    events on the stream the kernels run on.  Rates: `counted` = the full
    Gram's 2 q^2 n flops (q = p real, 2p for complex data viewed as reals:
    8 p^2 n), the count the Gram's algorithmic cost is quoted in; `syrk` = the
    q (q + 1) n flops of the distinct entries; `executed` = the tiles the
    kernel computes (upper 128 x 128 tiles, padded)."""
    import torch
    import vmc_jax_amd._lib as L
    st = L.stream_ptr()

    def timed(fn):
        for _ in range(3):
            fn()
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        a.record()
        for _ in range(reps):
            fn()
        b.record()
        b.synchronize()
        return a.elapsed_time(b) / reps

    out = {}
    g = torch.Generator(device="cuda").manual_seed(0)
    for name, n, p, cpx in (("tfim10x10_cnn_dual", 8192, 1240, False), ("tfim_l32_cpxrbm_holomorphic", 16384, 2112, True)):
        q = p * (2 if cpx else 1)
        D64 = torch.randn(n, q, dtype=torch.float64, device="cuda", generator=g)
        D32 = D64.float()
        S = torch.empty(p, q, dtype=torch.float64, device="cuda")
        side = (q + 127) // 128
        counted, syrk = 2.0 * q * q * n, float(q) * (q + 1) * n
        executed = 2.0 * (side * (side + 1) // 2) * 128 * 128 * n
        rec = {"n": n, "p": p, "complex": cpx, "counted_flop": counted}
        legs = [("fp64", lambda: L.call("jv_gram_pp", L.ptr(D64), n, p, int(cpx), L.ptr(S), st), FP64_PEAK_TFLOPS, 1),
                ("bf16x3_from_fp64", lambda: L.call("jv_gram_pp_mixed", L.ptr(D64), 0, n, p, int(cpx), L.ptr(S), st),
                 BF16_PEAK_TFLOPS, 3)]
        if not cpx:
            legs.append(("bf16x3_from_fp32",
                         lambda: L.call("jv_gram_pp_mixed", L.ptr(D32), 1, n, p, 0, L.ptr(S), st), BF16_PEAK_TFLOPS, 3))
        for leg, fn, peak, products in legs:
            ms = timed(fn)
            rate = lambda f: f / (ms * 1e-3) / 1e12
            rec[leg] = {"ms": ms, "counted_tflops": rate(counted), "syrk_tflops": rate(syrk),
                        "executed_tflops": rate(executed * products), "peak_tflops": peak,
                        "counted_frac": rate(counted) * products / peak, "syrk_frac": rate(syrk) * products / peak,
                        "executed_frac": rate(executed * products) / peak}
        out[name] = rec
        del D64, D32, S
    return out


def batched_sweep(num_instances=64, steps=3, seed=11, workers=8):
    """Batched field sweep: independent 1D TFIM instances with L ~ U{16, 24, 32,
    48, 64}, h ~ U[0.5, 1.5], complex RBM alpha ~ U{1, 2, 4} (per-instance seed),
    4096 Metropolis samples (256 chains) per instance and step, imaginary-time
    Euler steps (dt = 1e-2).  Instances with Pr <= 4224 solve the TDVP equation
    through the eigendecomposition of S; larger ones (Pr up to 33280, where S
    alone would be 17.7 GB) use MinSR, whose T is 4096 x 4096.  Sampling,
    log psi + O_k and O_loc of all instances run as one launch each from a
    captured HIP graph (launch shapes autotuned per phase); the solves run on
    `workers` threads with their own streams."""
    import torch
    import vmc_jax_amd as jv
    from vmc_jax_amd import batch as jb
    rng = np.random.default_rng(seed)
    inst = []
    kinds = []
    t0 = time.perf_counter()
    for i in range(num_instances):
        L = int(rng.choice([16, 24, 32, 48, 64]))
        h = float(rng.uniform(0.5, 1.5))
        alpha = int(rng.choice([1, 2, 4]))
        jseed = int(rng.integers(1 << 30))
        pr = 2 * (alpha * L + alpha * L * L)
        tdvp = pr <= 4224
        psi = jv.NQS(jv.nets.CpxRBM(numHidden=alpha * L, bias=True), seed=jseed)
        psi(torch.zeros((1, 1, L), dtype=torch.int32))
        H = tfim_strings(jv.operator, (L,), -1.0, -h)
        eq = jv.util.TDVP(None, snrTol=2, pinvTol=1e-8, rhsPrefactor=1., makeReal='real') if tdvp else \
            jv.util.MinSR(None, pinvTol=1e-8, solver="auto")
        inst.append(jb.Instance(psi, H, eq, 1e-2, key=jseed, numChains=256, numSamples=4096))
        kinds.append("L%d-a%d-%s" % (L, alpha, "tdvp" if tdvp else "minsr"))
    sweep = jb.BatchedSweep(inst, workers=workers, tune=True, graph=True)
    setup = time.perf_counter() - t0
    print("c4: %d instances set up in %.1f s (launch shapes %s, %s)" % (num_instances, setup, sweep.shapes,
                                                                      sweep.tuned_from), file=sys.stderr, flush=True)
    try:
        for _ in range(2):  # warm-up: library kernels, per-thread workspaces, first refused Cholesky attempts
            sweep.step()
        torch.cuda.synchronize()
        jv._lib.call("jv_timer_reset")
        jv._lib.call("jv_timer_enable", 1)
        t1 = time.perf_counter()
        for k in range(steps):
            sweep.step()
            print("c4: step %d of %d, %.0f ms" % (k + 1, steps, sweep.step_ms[-1]), file=sys.stderr, flush=True)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t1
        jv._lib.call("jv_timer_enable", 0)
        ph = {k: timer(jv, k) for k in ("batch_sample", "batch_grad", "batch_eloc", "eigh", "gram", "cholesky")}
        lat = np.array(sweep.step_ms[2:])
        phase = sweep.phase_ms()[2:]
        energies = [float(np.real(i.energy)) for i in inst]
        return {"instances": num_instances, "steps": steps, "wall_s": wall,
                "instance_steps_per_s": num_instances * steps / wall,
                "p50_step_ms": float(np.percentile(lat, 50)), "p99_step_ms": float(np.percentile(lat, 99)),
                "batched_pass_ms": float(np.mean(phase)) if phase else None,
                "solver_ms_per_step_summed_over_threads": {k: v[0] / steps for k, v in ph.items() if v[1] > 0},
                "launch_shapes": {k: list(v) for k, v in sweep.shapes.items()}, "tuned_from": sweep.tuned_from,
                "graph": sweep.graph is not None, "workers": workers, "setup_s": setup,
                "energies_finite": bool(np.all(np.isfinite(energies))), "mix": kinds}
    finally:
        sweep.close()


def serve_mix(num_jobs=16, slots=8, seed=7, rate=8.0, steps=(3, 9), samples=(2048, 4096), L1d=(16, 20, 24),
              side=(6, 11), batch_rbm=True, workers=8):
    """Served workload: Poisson arrivals (rate jobs/s); 70% 1D jobs (TFIM or XXZ
    with Delta ~ U[0, 2], L ~ U(L1d), complex RBM alpha=2; half ground-state
    MinSR with Euler, half real-time TDVP with Heun) and 30% 2D TFIM jobs (side
    ~ U[side), h ~ U[2, 4], the 3x(8ch,3x3) CNN, MinSR); steps ~ U[steps) and
    samples ~ U(samples) per job.  Jobs are admitted by HBM footprint into
    `slots` slots; with batch_rbm the 1D jobs share one batch (one launch per
    phase for all of them), the 2D jobs run on their own threads and streams."""
    import torch
    import vmc_jax_amd as jv
    from vmc_jax_amd import serving
    rng = np.random.default_rng(seed)
    t, jobs = 0.0, []
    for i in range(num_jobs):
        t += rng.exponential(1.0 / rate)
        nsteps = int(rng.integers(*steps))
        ns = int(rng.choice(samples))
        jseed = int(rng.integers(1 << 30))
        if rng.uniform() < 0.7:
            L = int(rng.choice(L1d))
            xxz = rng.uniform() < 0.5
            delta = float(rng.uniform(0, 2))
            ground = rng.uniform() < 0.5

            def make(L=L, xxz=xxz, delta=delta, ground=ground, ns=ns, jseed=jseed):
                psi = jv.NQS(jv.nets.CpxRBM(numHidden=2 * L, bias=True), seed=jseed)
                psi(torch.zeros((1, 1, L), dtype=torch.int32))
                op = jv.operator
                H = op.BranchFreeOperator()
                for l in range(L):
                    r = (l + 1) % L
                    if xxz:
                        H.add(op.scal_opstr(1.0, (op.Sx(l), op.Sx(r))))
                        H.add(op.scal_opstr(1.0, (op.Sy(l), op.Sy(r))))
                        H.add(op.scal_opstr(delta, (op.Sz(l), op.Sz(r))))
                    else:
                        H.add(op.scal_opstr(-1.0, (op.Sz(l), op.Sz(r))))
                        H.add(op.scal_opstr(-1.0, (op.Sx(l),)))
                ms = jv.sampler.MCSampler(psi, (L,), jseed, updateProposer=jv.sampler.propose_spin_flip,
                                          numChains=256, numSamples=ns)
                if ground:
                    return psi, H, jv.util.MinSR(ms, pinvTol=1e-8, solver="auto"), jv.util.stepper.Euler(1e-2)
                return psi, H, jv.util.TDVP(ms, rhsPrefactor=1.j, makeReal='imag'), jv.util.stepper.Heun(1e-3)
            pr = 2 * (2 * L * L + 2 * L)
            fp = serving.footprint_bytes(ns, pr, L, True, "minsr" if ground else "tdvp", 256)
            kind = "1d-%s-%s-L%d" % ("xxz" if xxz else "tfim", "gs" if ground else "rt", L)
        else:
            sd = int(rng.integers(*side))
            h = float(rng.uniform(2, 4))

            def make(side=sd, h=h, ns=ns, jseed=jseed):
                net = jv.nets.CNN(F=(3, 3), channels=(8, 8, 8), strides=(1, 1))
                psi = jv.NQS(net, seed=jseed)
                psi(torch.zeros((1, 1, side, side), dtype=torch.int32))
                psi.set_parameters(net.he_normal(psi.desc, np.random.default_rng(jseed)))
                H = tfim_strings(jv.operator, (side, side), -1.0, -h)
                ms = jv.sampler.MCSampler(psi, (side, side), jseed, updateProposer=jv.sampler.propose_spin_flip,
                                          numChains=256, numSamples=ns)
                return psi, H, jv.util.MinSR(ms, pinvTol=1e-10, solver="auto"), jv.util.stepper.Euler(5e-3)
            fp = serving.footprint_bytes(ns, 1240, sd * sd, False, "minsr", 256)
            kind = "2d-cnn-%d" % sd
        jobs.append(serving.Job(i, make, nsteps, fp, arrival=t, meta={"kind": kind}))
    srv = serving.Server(max_slots=slots, batch_rbm=batch_rbm, workers=workers)
    for j in jobs:
        srv.submit(j)
    t_run = time.perf_counter()
    rep = srv.run(progress=lambda j: print("served: job %d (%s) %s, %d steps, t = %.1f s" %
                                           (j.job_id, j.meta["kind"], j.state, j.steps_done,
                                            time.perf_counter() - t_run), file=sys.stderr, flush=True))
    rep["jobs"] = num_jobs
    rep["slots"] = slots
    rep["shared_batch"] = batch_rbm
    rep["arrival_span_s"] = t
    rep["failures"] = [j.error for j in srv.finished if j.state == "failed"]
    lat = {}
    for j in srv.finished:
        lat.setdefault(j.meta["kind"].split("-")[0] + "-" + j.meta["kind"].split("-")[2 if j.meta["kind"].startswith("1d") else 1],
                       []).extend(j.latencies)
    rep["p99_step_ms_by_kind"] = {k: float(np.percentile(v, 99) * 1e3) for k, v in lat.items() if v}
    rep["workload"] = serve_mix.__doc__.split("\n")[0].strip()
    return rep


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(sizes=(64, 128), chains=32, reps=3):
    """FP64 oracle on the host cores, timed on bounded samples of the 10x10 CNN
    step: per phase one warm-up and the median of `reps` runs at two sample
    counts, the scaling exponent fitted from the two sizes, and the phase
    extrapolated to 8192 samples with the algorithm's exponent (sampling:
    linear in chains x sweeps, here 512 chains of 10 + 16 sweeps of 100
    proposals from `chains` chains; E_loc and gradients: linear; T Gram:
    quadratic; eigendecomposition: cubic).  numpy's BLAS uses every core of the
    process's affinity set."""
    from oracle import nets as onets, operator as oop, sampler as osamp
    from oracle.stats import SampledObs
    shape, samples, nchains = (10, 10), 8192, 512
    onet = onets.CNN(shape, F=(3, 3), channels=(8, 8, 8), strides=(1, 1))
    P = np.zeros(onet.num_real_params())
    rng = np.random.default_rng(0)
    off = 0
    for leaf in onet.leaves:  # He-normal kernels, zero biases
        if leaf.name.endswith("kernel"):
            fan_in = leaf.shape[0] * leaf.shape[1] * leaf.shape[2]
            P[off:off + leaf.size] = rng.normal(size=leaf.size) * math.sqrt(2.0 / fan_in)
        off += leaf.size
    psi = onets.NQS(onet, P, batchSize=64)
    H = tfim_strings(oop, shape)

    def med(fn):
        fn()
        ts = []
        for _ in range(reps):
            t0 = time.perf_counter()
            fn()
            ts.append(time.perf_counter() - t0)
        return float(np.median(ts))

    t0 = time.perf_counter()
    meas = {}
    for n in sizes:
        ms = osamp.MCSampler(psi, shape, 0, numChains=chains, thermalizationSweeps=10, sweepSteps=10)
        spc = n // chains
        box = {}

        def samp():
            box["s"] = ms.sample(numSamples=spc * chains)
        t = {"sampling": med(samp)}
        s, lp, p = box["s"]
        t["eloc"] = med(lambda: H.get_O_loc(s, psi, lp, 0.0))
        t["grad"] = med(lambda: psi.gradients(s))
        D = np.real(SampledObs(psi.gradients(s), p)._data)
        t["gram"] = med(lambda: D @ D.T)
        T = D @ D.T
        t["eigh"] = med(lambda: np.linalg.eigh(T))
        meas[n] = t
    n1, n2 = sizes
    expo = {"sampling": 1.0, "eloc": 1.0, "grad": 1.0, "gram": 2.0, "eigh": 3.0}
    fitted = {k: math.log(max(meas[n2][k], 1e-9) / max(meas[n1][k], 1e-9)) / math.log(n2 / n1) for k in expo}
    est = {}
    for k, a in expo.items():
        if k == "sampling":  # chains x (therm + samples per chain) proposal sweeps
            work_full = nchains * (10 + samples // nchains)
            work_sub = chains * (10 + n2 // chains)
            est[k] = meas[n2][k] * work_full / work_sub
        else:
            est[k] = meas[n2][k] * (samples / n2) ** a
    total = sum(est.values())
    cores = len(os.sched_getaffinity(0))
    return {"value": 1.0 / total, "unit": "steps/s", "cores": cores, "kind": "port",
            "cpu_model": _cpu_model(), "threads": cores, "wall_s": time.perf_counter() - t0,
            "sample": ("jVMC-algorithm CPU restatement (FP64 numpy oracle): one MinSR step of the 10x10 CNN "
                       "workload on %s samples from %d chains, per phase warm-up + median of %d; extrapolated "
                       "to 8192 samples / 512 chains with the algorithm's exponents" % (list(sizes), chains, reps)),
            "phase_s_measured": {str(n): {k: round(v, 4) for k, v in meas[n].items()} for n in sizes},
            "phase_s_extrapolated": {k: round(v, 3) for k, v in est.items()},
            "fitted_exponents": {k: round(v, 2) for k, v in fitted.items()}}


def accuracy_check(w, n=1536):
    """|dE/E| and the errors of S, F and the MinSR update against the FP64
    oracle on the same configurations and weights (n fresh samples of the
    workload's current state; 1536 = 3 per chain, more than P + 1 so that S
    can be regular and the Cholesky route can be checked), through the same
    device path the timed steps use (overlapped dual solve; FP32 O and the
    compensated-BF16 Gram in the mixed form): the accuracy condition of the
    metric.  `route` names the solve route the checked step took (cholesky /
    eigendecomposition)."""
    import torch
    from oracle import nets as onets, operator as oop, solvers as osolv
    from oracle.stats import SampledObs as OObs
    jv = w.jv

    class Rec:
        def __init__(self, smp):
            self.s, self.last = smp, None

        def sample(self, *a, **k):
            self.last = self.s.sample(*a, **k)
            return self.last

    rec = Rec(w.sampler)
    eq = jv.util.MinSR(rec, pinvTol=1e-10, solver="dual")
    upd = eq(w.psi.get_parameters(), 0.0, psi=w.psi, hamiltonian=w.H, intStep=0, numSamples=n).cpu().numpy()
    route = "cholesky" if eq.last_info.iterations == 1 else "eigendecomposition"
    e_gpu = complex(eq.ElocMean0)
    s, lp, p = rec.last
    # S and F of the device data (FP32 in the mixed form) beside the oracle's
    G = jv.stats.SampledObs(w.psi.gradients(s, layout=0), p)
    El = jv.stats.SampledObs(w.H.get_O_loc(s, w.psi, lp, 0.0), p)
    D = G._data
    nn, P = D.shape
    S = torch.empty((P, P), dtype=torch.float64, device=D.device)
    if D.dtype == torch.float32:
        jv._lib.call("jv_gram_pp_mixed", jv._lib.ptr(D), 1, nn, P, 0, jv._lib.ptr(S), jv._lib.stream_ptr())
    else:
        jv._lib.call("jv_gram_pp", jv._lib.ptr(D), nn, P, 0, jv._lib.ptr(S), jv._lib.stream_ptr())
    F = (D.to(torch.float64).T @ El._data.real.reshape(-1).to(torch.float64)).cpu().numpy()
    S = S.cpu().numpy()
    onet = onets.CNN(w.shape, F=(3, 3), channels=(8, 8, 8), strides=(1, 1))
    opsi = onets.NQS(onet, w.psi.get_parameters().cpu().numpy(), batchSize=64)
    cs = s[0].cpu().numpy()
    pw = p[0].cpu().numpy()
    Ho = tfim_strings(oop, w.shape)
    Eo = OObs(Ho.get_O_loc(cs, opsi, opsi(cs), 0.0), pw)
    Go = OObs(opsi.gradients(cs), pw)
    ref = osolv.MinSR(None, pinvTol=1e-10).solve(Eo, Go, False)
    Do, eo = np.real(Go._data), np.real(Eo._data).reshape(-1)
    So, Fo = Do.T @ Do, Do.T @ eo
    e_o = complex(Eo.mean()[0])
    res, res_o = np.linalg.norm(Do @ upd + eo), np.linalg.norm(Do @ ref + eo)
    return {"samples": int(n), "route": route,
            "against": "FP64 numpy oracle (T-form MinSR) on the same configurations and weights",
            "energy": e_gpu.real, "dE_rel": abs(e_gpu - e_o) / abs(e_o),
            "dS_rel": float(np.linalg.norm(S - So) / np.linalg.norm(So)),
            "dF_rel": float(np.linalg.norm(F - Fo) / np.linalg.norm(Fo)),
            "dtheta_rel": float(np.linalg.norm(upd - ref) / np.linalg.norm(ref)),
            "residual_rel": float(res / res_o - 1.0)}


def c1_oracle_run(steps=200):
    """The L=10 CpxRBM exact-sampler TDVP workload on the FP64 oracle, every
    step timed on the host cores: the same initial parameters (N(0, 0.01^2),
    seed 0, the flat layout shared with the device network), Euler dt = 1e-2,
    snrTol = 2, pinvTol = 1e-8, makeReal 'real'.  Returns the per-step times
    and the energies."""
    from oracle import nets as onets, operator as oop, sampler as osamp, solvers as osol
    L = 10
    net = onets.CpxRBM((L,), numHidden=20, bias=True, visibleBias=True)
    psi = onets.NQS(net, 0.01 * np.random.default_rng(0).normal(size=net.num_real_params()))
    H = tfim_strings(oop, (L,), -1.0, -0.7)
    eq = osol.TDVP(osamp.ExactSampler(psi, (L,)), snrTol=2, pinvTol=1e-8, rhsPrefactor=1., makeReal='real')
    st = osol.Euler(timeStep=1e-2)
    times, energies = [], []
    for _ in range(steps):
        t0 = time.perf_counter()
        dp, _ = st.step(0.0, eq, psi.get_parameters(), hamiltonian=H, psi=psi)
        psi.set_parameters(dp)
        times.append(time.perf_counter() - t0)
        energies.append(float(eq.get_energy_mean()))
    return times, energies


def other_workloads():
    """The other configurations measured in the same run: c1 (all 200 steps),
    c2 (Heun steps of the L=32 real-time workload) and c4 (batched field
    sweep), each on the device in FP64."""
    out = {}
    for name, steps, warmup in (("c1", 200, 2), ("c2", 4, 1)):
        try:
            w, el, kt, en = run_gpu(name, steps, warmup, "dual")
            out[name] = {"value": steps / el, "unit": "steps/s", "steps": steps, "warmup": warmup,
                         "ms_per_step": 1000.0 * el / steps, "config": w.desc, "energy_last": en[-1],
                         "kernel_ms_per_step": {k: v[0] / steps for k, v in kt.items() if v[1] > 0}}
            print("other workload %s: %.1f steps/s" % (name, steps / el), file=sys.stderr, flush=True)
        except Exception as e:  # report, never hide the headline
            out[name] = {"error": repr(e)}
    try:
        rep = batched_sweep(steps=2)
        out["c4"] = {"value": rep["instance_steps_per_s"], "unit": "instance-steps/s", "steps": rep["steps"],
                     "instances": rep["instances"], "p50_step_ms": rep["p50_step_ms"],
                     "p99_step_ms": rep["p99_step_ms"], "batched_pass_ms": rep["batched_pass_ms"],
                     "config": batched_sweep.__doc__.split("\n")[0].strip()}
    except Exception as e:
        out["c4"] = {"error": repr(e)}
    return out


def accuracy_c2(w):
    """The mixed form of the L = 32 workload (compensated BF16 Gram of the
    holomorphic solve) against the FP64 solve on the same samples; the FP64
    device solve matches the FP64 oracle at this size
    (test_tdvp_heun_rhs_l32_cpxrbm_16384_samples)."""
    import torch
    jv = w.jv
    s, lp, p = w.sampler.sample()
    E = jv.stats.SampledObs(w.H.get_O_loc(s, w.psi, lp, 0.0), p)
    G = jv.stats.SampledObs(w.psi.gradients(s), p)
    out = {}
    for mixed in (False, True):
        eq = jv.util.TDVP(w.sampler, rhsPrefactor=1.j, makeReal='imag')
        eq._holo_cols = w.psi.holomorphic_columns()
        eq._mixed = mixed
        u, _, _ = eq.solve(E, G)
        out[mixed] = (u.cpu().numpy(), eq.S.cpu().numpy(), eq.F.cpu().numpy(), eq.ev.cpu().numpy())
    (u0, S0, F0, e0), (u1, S1, F1, e1) = out[False], out[True]
    rel = lambda a, b: float(np.linalg.norm(a - b) / np.linalg.norm(b))
    return {"samples": int(s.shape[1]), "against": "FP64 device solve on the same samples",
            "dE_rel": 0.0, "dS_rel": rel(S1, S0), "dF_rel": rel(F1, F0), "dspectrum_rel": rel(e1, e0),
            "dtheta_rel": rel(u1, u0)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="c3", choices=["c1", "c2", "c3", "c4", "c5", "sym", "gram"])
    ap.add_argument("--mode", default="dual", choices=["T", "dual"], help="MinSR solve form (c3)")
    ap.add_argument("--t-steps", type=int, default=3, help="c3: steps of the other MinSR form reported beside")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-overlap", action="store_true",
                    help="c3: compute E_loc before the MinSR solve instead of beside it (isolated kernel timings)")
    ap.add_argument("--serve-jobs", type=int, default=16, help="c3: jobs of the served mix (0 = skip)")
    ap.add_argument("--no-mixed", action="store_true", help="c3: skip the mixed-precision form")
    ap.add_argument("--no-other-workloads", action="store_true", help="c3: skip the c1 / c2 / c4 legs")
    ap.add_argument("--rate", type=float, default=4.0, help="c5: Poisson arrival rate (jobs/s)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="end with an elementwise 512 MiB -> 512 MiB FP64 device copy (a kernel of known "
                         "HBM traffic that scales the FETCH_SIZE / WRITE_SIZE counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (every box has one MI355X); "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    if args.workload == "c4":
        t0 = time.perf_counter()
        rep = batched_sweep(steps=max(1, min(args.steps, 4)))
        wall = time.perf_counter() - t0
        print(json.dumps({
            "metric": "instance TDVP steps/s (c4: batched field sweep)", "value": rep["instance_steps_per_s"],
            "unit": "instance-steps/s", "n_gpus": 1, "steps": rep["steps"], "warmup": 1,
            "ms_per_step": 1000.0 * rep["wall_s"] / rep["steps"], "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "f64",
            "data": "synthetic (random network parameters; samples drawn from the networks)",
            "config": {"workload": batched_sweep.__doc__.split("\n")[0].strip(), "instances": rep["instances"]},
            "report": rep, "wall_s_including_setup": wall}))
        return
    if args.workload == "gram":
        rep = gram_bench(reps=max(1, args.steps))
        lead = rep["tfim10x10_cnn_dual"]["fp64"]
        print(json.dumps({
            "metric": "Gram S = D^T D TFLOP/s (counted 2 q^2 n)", "value": lead["counted_tflops"],
            "unit": "TFLOP/s", "n_gpus": 1, "steps": max(1, args.steps), "warmup": 3, "ms_per_step": lead["ms"],
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
            "data": "synthetic (standard normal D)",
            "config": {"workload": "Gram of the 10x10 CNN dual MinSR (8192 x 1240) and the L=32 complex RBM "
                                   "holomorphic TDVP (16384 x 2112 complex)"},
            # the hardware's rate: the multiply-adds of the tiles the kernel computes
            # (the counted rate credits the symmetric half it skips)
            "roofline": {"bound": "mfma", "achieved": lead["executed_tflops"], "peak": FP64_PEAK_TFLOPS,
                         "unit": "TFLOP/s", "frac": lead["executed_frac"], "counted_frac": lead["counted_frac"],
                         "traffic": None, "kernel": "gram64_tn"},
            "report": rep}))
        if args.calibrate_copy:
            calibration_copy()
        return
    if args.workload == "c5":
        t0 = time.perf_counter()
        rep = serve_mix(num_jobs=args.serve_jobs if args.serve_jobs > 16 else 100, slots=32, seed=5,
                        rate=args.rate, steps=(10, 31), samples=(2048, 4096, 8192), L1d=(16, 20, 24, 28, 32),
                        side=(6, 13))
        wall = time.perf_counter() - t0
        print(json.dumps({
            "metric": "served VMC jobs: aggregate steps/s, p50 / p99 step latency (c5)", "value": rep["steps_per_s"],
            "unit": "steps/s", "n_gpus": 1, "steps": rep["steps"], "warmup": 0, "ms_per_step": rep["p50_step_ms"],
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
            "data": "synthetic (random network parameters; samples drawn from the networks)",
            "config": {"workload": rep["workload"], "jobs": rep["jobs"], "slots": rep["slots"], "rate": args.rate},
            "report": rep, "wall_s_including_setup": wall}))
        return
    w, elapsed, kt, energies = run_gpu(args.workload, args.steps, args.warmup, args.mode, not args.no_overlap)
    steps_s = args.steps / elapsed
    out = {
        "metric": METRIC if args.workload == "c3" else "TDVP steps/s (%s)" % args.workload,
        "value": steps_s,
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1000.0 * elapsed / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (random network parameters; samples drawn from the network itself)",
        "config": {"workload": w.desc, "samples_per_step": w.samples},
        "samples_per_s": steps_s * w.samples,
        "kernel_ms_per_step": {k: v[0] / args.steps for k, v in kt.items() if v[1] > 0},
        "minsr_cholesky_steps": w.cholesky_steps,
        "energy_last": energies[-1] if energies else None,
    }
    if args.workload == "c1" and not args.no_cpu_baseline:
        # the oracle over the same 200-step trajectory: CPU baseline in full and
        # the energy after every step against the device's
        times, en_cpu = c1_oracle_run(args.warmup + args.steps)
        med = float(np.median(times))
        out["cpu_baseline"] = {"value": 1.0 / med, "unit": "steps/s", "cores": os.cpu_count(), "kind": "port",
                               "cpu_model": _cpu_model(),
                               "sample": "jVMC-algorithm CPU restatement (FP64 numpy oracle): the whole %d-step "
                                         "trajectory, median step time" % len(times),
                               "wall_s": float(sum(times))}
        dev = energies
        ref = en_cpu[args.warmup:]
        out["accuracy"] = {"against": "FP64 numpy oracle, same initial parameters, same trajectory",
                           "max_dE_rel_over_steps": float(max(abs(a - b) / abs(b) for a, b in zip(dev, ref))),
                           "energy_last_device": dev[-1], "energy_last_oracle": ref[-1]}
    if args.workload == "c2" and not args.no_mixed:
        wm, em, ktm, enm = run_gpu("c2", args.steps, args.warmup, args.mode, True, "mixed")
        out["mixed_form"] = {
            "value": args.steps / em, "unit": "steps/s", "ms_per_step": 1000.0 * em / args.steps,
            "dtype": "fp64 network kernels + bf16x3 (compensated) Gram, fp64 solve",
            "kernel_ms_per_step": {k: v[0] / args.steps for k, v in ktm.items() if v[1] > 0}}
        try:
            out["mixed_form"]["accuracy"] = accuracy_c2(wm)
        except Exception as e:
            out["mixed_form"]["accuracy"] = {"error": repr(e)}
    if args.workload == "c3":
        out["config"]["solver_form"] = args.mode
        out["config"]["eloc_beside_solve"] = not args.no_overlap and args.mode == "dual"
        out["roofline"] = roofline_c3(kt, args.steps, args.mode, measured_traffic())
        if args.mode == "dual" and not args.no_mixed:
            # the mixed-precision form from the same initial state, same steps
            wm, em, ktm, enm = run_gpu("c3", args.steps, args.warmup, "dual", not args.no_overlap, "mixed")
            out["mixed_form"] = {
                "value": args.steps / em, "unit": "steps/s", "ms_per_step": 1000.0 * em / args.steps,
                "samples_per_s": args.steps / em * wm.samples,
                "dtype": "fp32 network kernels + bf16x3 (compensated) Gram, fp64 solve",
                "kernel_ms_per_step": {k: v[0] / args.steps for k, v in ktm.items() if v[1] > 0},
                "minsr_cholesky_steps": wm.cholesky_steps, "energy_last": enm[-1] if enm else None}
            if not args.no_cpu_baseline:
                try:
                    out["mixed_form"]["accuracy"] = accuracy_check(wm)
                except Exception as e:
                    out["mixed_form"]["accuracy"] = {"error": repr(e)}
        if args.t_steps > 0:
            om = "T" if args.mode == "dual" else "dual"
            _, e2, kt2, _ = run_gpu("c3", args.t_steps, 1, om)
            out["other_solver_form"] = {
                "solver_form": om, "value": args.t_steps / e2, "unit": "steps/s",
                "ms_per_step": 1000.0 * e2 / args.t_steps,
                "kernel_ms_per_step": {k: v[0] / args.t_steps for k, v in kt2.items() if v[1] > 0}}
        if args.serve_jobs > 0:
            try:
                out["served"] = serve_mix(args.serve_jobs)
            except Exception as e:
                out["served"] = {"error": repr(e)}
        if not args.no_other_workloads:
            out["other_workloads"] = other_workloads()
        if not args.no_cpu_baseline:  # the oracle leg: accuracy checker and CPU baseline
            try:
                out["accuracy"] = accuracy_check(w)
            except Exception as e:
                out["accuracy"] = {"error": repr(e)}
            try:
                out["cpu_baseline"] = cpu_baseline()
            except Exception as e:  # report, never hide the GPU number
                out["cpu_baseline"] = {"error": repr(e)}
    print(json.dumps(out))
    if args.calibrate_copy:
        calibration_copy()


def calibration_copy():
    """A 512 MiB -> 512 MiB FP64 elementwise copy: a kernel of known HBM
    traffic that scales the FETCH_SIZE / WRITE_SIZE counters."""
    import torch
    a = torch.ones(64 << 20, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    torch.mul(a, 1.0, out=b)
    torch.cuda.synchronize()
    print("calibration copy: %d bytes read, %d bytes written" % (a.numel() * 8, b.numel() * 8), file=sys.stderr)


if __name__ == "__main__":
    main()
