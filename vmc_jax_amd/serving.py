"""Continuous-batching request loop for many independent VMC jobs on one GPU.

A job is one variational state with its Hamiltonian, sampler, TDVP / MinSR
equation and stepper, plus a step budget.  The server

* admits queued jobs whenever a slot is free and the job's HBM footprint fits
  the remaining budget (jobs that can never fit are refused),
* runs every admitted job on its own HIP stream (default priority) from its own
  host thread, so the kernels of different jobs overlap on the device and a
  job's latency-bound phases (the FP64 eigendecomposition, host reads of the
  solver status) hide behind other jobs' work,
* keeps each job's parameters, walkers and Philox counters on the device
  between steps,
* records the wall-clock latency of every step and reports p50 / p99 per-step
  latency with the aggregate steps/s,
* isolates failures: a job whose step raises (or produces a non-finite energy)
  is marked failed and its slot and memory are released; other jobs continue.

With ``batch_rbm=True`` the RBM jobs (CpxRBM / RBM with a Metropolis sampler,
a time-independent operator and an Euler or Heun stepper) do not get threads:
they join one shared batch (``vmc_jax_amd.batch``), so the sampling, O_k and
O_loc of all of them run as one launch per phase, and the batch is rebuilt
whenever jobs join or leave; their solves run on a pool of worker threads.
Every other job keeps its own thread and stream beside the batch.

Results do not depend on the mix: every job draws from its own Philox stream,
all reductions have a fixed order, and the library keeps solver workspaces per
host thread, so a job's trajectory is bit-identical alone and in the mix.

jVMC has no serving layer; the per-step work is the reference's
``stepper.step(t, tdvp, params, hamiltonian=..., psi=...)`` (util.py:188-212).
"""
import collections
import heapq
import math
import threading
import time

import numpy as np
import torch


class Job:
    """One VMC job.  ``make`` builds (psi, hamiltonian, equation, stepper) on
    the device when the job is admitted; ``footprint`` is its HBM estimate in
    bytes."""

    def __init__(self, job_id, make, num_steps, footprint, arrival=0.0, meta=None):
        self.job_id = job_id
        self.make = make
        self.num_steps = int(num_steps)
        self.footprint = int(footprint)
        self.arrival = float(arrival)
        self.meta = meta or {}
        self.state = "queued"
        self.error = None
        self.steps_done = 0
        self.latencies = []
        self.energies = []
        self.t = 0.0
        self.admitted_at = None
        self.finished_at = None
        self.params = None

    def run(self, stream, stop_event, built=None):
        """Body of the job's worker thread (built: the objects of make() when the
        server has already built them on this stream)."""
        try:
            ctx = torch.cuda.stream(stream) if stream is not None else _NullCtx()
            with ctx:
                psi, H, eq, stepper = built if built is not None else self.make()
                while self.steps_done < self.num_steps and not stop_event.is_set():
                    t0 = time.perf_counter()
                    dp, dt = stepper.step(self.t, eq, psi.get_parameters(), hamiltonian=H, psi=psi,
                                          numSamples=None)
                    psi.set_parameters(dp)
                    e = eq.get_energy_mean()
                    if stream is not None:
                        stream.synchronize()
                    self.latencies.append(time.perf_counter() - t0)
                    if not math.isfinite(e):
                        raise FloatingPointError("non-finite energy at step %d" % self.steps_done)
                    self.energies.append(e)
                    self.t += dt
                    self.steps_done += 1
                p = psi.get_parameters()
                self.params = p.cpu() if torch.is_tensor(p) else p
            self.state = "done"
        except Exception as exc:  # isolate the failure to this job
            self.state = "failed"
            self.error = repr(exc)
        finally:
            self.finished_at = time.perf_counter()


class _BatchedJob:
    """A job advanced inside the shared batch: its Instance plus the stepper
    stages (Euler: one right-hand side per step; Heun: two, stepper.py:45-93)."""

    def __init__(self, job, psi, H, eq, stepper):
        from .batch import Instance
        ms = eq.sampler
        self.job = job
        self.inst = Instance(psi, H, eq, stepper.dt, key=ms.key, numChains=ms.numChains, numSamples=ms.numSamples,
                             thermalizationSweeps=ms.thermalizationSweeps, sweepSteps=ms.sweepSteps)
        self.inst.calls = ms.callIndex
        self.inst.walkers.copy_(ms.states)
        self.heun = type(stepper).__name__ == "Heun"
        self.y = psi.params.clone() if self.heun else None
        self.stage = 0
        self.k0 = None
        self.energy0 = None
        self.t_start = None
        self.completed = False

    def on_rhs(self):
        """Runs on a worker thread after the batched pass: the solve, then the
        stepper stage; sets `completed` when a step has finished."""
        k = self.inst.rhs()
        dt = self.inst.dt
        P = self.inst.psi.params
        if not self.heun:
            self.energy0 = self.inst.energy
            P.add_(dt * k)
            self.completed = True
            return
        if self.stage == 0:
            self.k0 = k
            self.energy0 = self.inst.energy
            P.copy_(self.y + dt * k)
            self.stage = 1
            self.completed = False
            return
        self.y = self.y + 0.5 * dt * (self.k0 + k)
        P.copy_(self.y)
        self.stage = 0
        self.completed = True


def _batchable(psi, H, eq, stepper, max_dim):
    """RBM job with a Metropolis sampler (mu = 2), a time-independent operator,
    an Euler or Heun stepper, and a solve of dimension <= max_dim (larger
    solves would hold every other batched job back for their duration)."""
    from . import _lib
    from .sampler import MCSampler
    if psi._sym is not None or psi.desc.kind not in (_lib.JV_NET_CPXRBM, _lib.JV_NET_RBM):
        return False
    if not isinstance(getattr(eq, "sampler", None), MCSampler) or eq.sampler.mu != 2:
        return False
    n = eq.sampler.numSamples
    dim = min(n, psi.numParameters) if type(eq).__name__ == "MinSR" else psi.numParameters
    if dim > max_dim:
        return False
    if type(stepper).__name__ not in ("Euler", "Heun"):
        return False
    if type(eq).__name__ == "TDVP" and eq.crossValidation:
        return False
    if not H.compiled:
        H.compile()
    return bool(torch.equal(H.prefactors(0.0), H.prefactors(1.0)))  # time-independent strings only


class _NullCtx:
    def __enter__(self):
        return self

    def __exit__(self, *a):
        return False


class Server:
    """Admission and scheduling of ``Job`` objects on one GPU.
    ``stream_factory`` makes the per-slot streams (``torch.cuda.Stream`` by
    default; host-only tests pass ``lambda: None``)."""

    def __init__(self, max_slots=8, hbm_budget=None, stream_factory=None, batch_rbm=False, workers=8,
                 batch_max_dim=2112):
        self.max_slots = int(max_slots)
        self.batch_rbm = bool(batch_rbm)
        self.batch_max_dim = int(batch_max_dim)
        self.workers = int(workers)
        self.batched = {}        # job_id -> _BatchedJob
        self._batch = None
        self._batch_ids = None
        self._pool = None
        self.stream_factory = stream_factory or torch.cuda.Stream
        if hbm_budget is None:
            free, total = torch.cuda.mem_get_info()
            hbm_budget = int(0.8 * free)
        self.hbm_budget = int(hbm_budget)
        self.queue = []          # heap of (arrival, seq, job)
        self._seq = 0
        self.active = {}         # job_id -> (job, thread, stream)
        self.finished = []
        self.refused = []
        self.used = 0
        self._stop = threading.Event()
        self._streams = collections.deque()

    def submit(self, job):
        if job.footprint > self.hbm_budget:
            job.state = "refused"
            job.error = "footprint %d B exceeds the HBM budget %d B" % (job.footprint, self.hbm_budget)
            self.refused.append(job)
            return False
        heapq.heappush(self.queue, (job.arrival, self._seq, job))
        self._seq += 1
        return True

    def _stream(self):
        return self._streams.popleft() if self._streams else self.stream_factory()

    def _admit(self, now):
        admitted = []
        while self.queue and len(self.active) + len(self.batched) < self.max_slots:
            arrival, _, job = self.queue[0]
            if arrival > now or job.footprint > self.hbm_budget - self.used:
                break  # FIFO by arrival: the head waits for memory
            heapq.heappop(self.queue)
            st = self._stream()
            job.state = "running"
            job.admitted_at = time.perf_counter()
            self.used += job.footprint
            admitted.append(job)
            built = None
            if self.batch_rbm:
                try:
                    with torch.cuda.stream(st):
                        built = job.make()
                    st.synchronize()
                    if _batchable(*built, self.batch_max_dim):
                        self.batched[job.job_id] = _BatchedJob(job, *built)
                        self._streams.append(st)
                        continue
                except Exception as exc:
                    self._fail(job, exc)
                    self._streams.append(st)
                    continue
            th = threading.Thread(target=job.run, args=(st, self._stop, built), daemon=True)
            self.active[job.job_id] = (job, th, st)
            th.start()
        return admitted

    def _fail(self, job, exc):
        job.state = "failed"
        job.error = repr(exc)
        job.finished_at = time.perf_counter()
        self.used -= job.footprint
        self.finished.append(job)

    def _tick_batch(self):
        """One right-hand side of every batched job: the shared sampling / O_k /
        O_loc pass, the solves on the worker pool, the stepper stages."""
        from concurrent.futures import ThreadPoolExecutor
        from .batch import BatchedSweep
        ids = tuple(sorted(self.batched))
        if ids != self._batch_ids:
            if self._batch is not None:
                self._batch.close()
            if self._pool is None:
                self._pool = ThreadPoolExecutor(max_workers=self.workers)
            self._batch = BatchedSweep([self.batched[i].inst for i in ids], workers=self.workers, tune="cache",
                                       graph=False, pool=self._pool)
            self._batch_ids = ids
        now = time.perf_counter()
        for bj in self.batched.values():
            if bj.stage == 0:
                bj.t_start = now
        owner = {id(self.batched[i].inst): self.batched[i] for i in ids}
        errors = self._batch.evaluate(lambda inst: owner[id(inst)].on_rhs())
        done_t = time.perf_counter()
        for k, jid in enumerate(ids):
            bj = self.batched[jid]
            job = bj.job
            exc = errors.get(k)
            if exc is None and bj.completed:
                e = float(np.real(bj.energy0))
                if not math.isfinite(e):
                    exc = FloatingPointError("non-finite energy at step %d" % job.steps_done)
                else:
                    job.latencies.append(done_t - bj.t_start)
                    job.energies.append(e)
                    job.t += bj.inst.dt
                    job.steps_done += 1
                    bj.completed = False
            if exc is not None:
                del self.batched[jid]
                self._fail(job, exc)
            elif job.steps_done >= job.num_steps:
                del self.batched[jid]
                job.params = bj.inst.psi.params.cpu()
                job.state = "done"
                job.finished_at = time.perf_counter()
                self.used -= job.footprint
                self.finished.append(job)

    def _reap(self):
        for jid in list(self.active):
            job, th, st = self.active[jid]
            if not th.is_alive():
                th.join()
                self.used -= job.footprint
                self._streams.append(st)
                del self.active[jid]
                self.finished.append(job)

    def run(self, poll=0.002, time_scale=1.0, progress=None):
        """Runs until every submitted job has finished.  Arrival times are in
        seconds after the start of ``run`` (scaled by ``time_scale``)."""
        if self.stream_factory is torch.cuda.Stream:
            # solver kernels loaded once on this thread before job threads start
            from . import _lib
            _lib.call("jv_init", torch.cuda.current_device())
        t_start = time.perf_counter()
        try:
            reported = 0
            while self.queue or self.active or self.batched:
                now = (time.perf_counter() - t_start) / time_scale
                self._reap()
                if progress is not None and len(self.finished) > reported:
                    for j in self.finished[reported:]:
                        progress(j)
                    reported = len(self.finished)
                self._admit(now)
                if self.batched:
                    self._tick_batch()
                else:
                    time.sleep(poll)
        finally:
            if self._batch is not None:
                self._batch.close()
                self._batch = None
            if self._pool is not None:
                self._pool.shutdown(wait=True)
                self._pool = None
        self.wall = time.perf_counter() - t_start
        return self.report()

    def report(self):
        lat = np.array([l for j in self.finished for l in j.latencies]) if self.finished else np.zeros(0)
        steps = int(sum(j.steps_done for j in self.finished))
        return {
            "jobs_done": sum(j.state == "done" for j in self.finished),
            "jobs_failed": sum(j.state == "failed" for j in self.finished),
            "jobs_refused": len(self.refused),
            "steps": steps,
            "wall_s": getattr(self, "wall", None),
            "steps_per_s": steps / self.wall if getattr(self, "wall", 0) else None,
            "p50_step_ms": float(np.percentile(lat, 50) * 1e3) if lat.size else None,
            "p99_step_ms": float(np.percentile(lat, 99) * 1e3) if lat.size else None,
        }

    def stop(self):
        self._stop.set()


def _cdiv(a, b):
    return -(-a // b)


def _gram_slice_bytes(n, q):
    """Partial tiles of the FP64 Gram of an n x q real matrix: 128 x 128 upper
    tiles, the sample axis cut into slices of whole 32-sample stages while the
    tiles number fewer than 512 and each slice keeps >= 16 stages (the rule of
    gram_pp_f64 in csrc/gram_mixed.hip)."""
    side = _cdiv(q, 128)
    ntiles, nkb, splits = side * (side + 1) // 2, _cdiv(n, 32), 1
    while ntiles * splits < 512 and nkb // (splits * 2) >= 16:
        splits *= 2
    splits = _cdiv(n, _cdiv(nkb, splits) * 32)
    return splits * ntiles * 128 * 128 * 8


def _bf16_gram_bytes(n, q):
    """The compensated BF16 Gram's FP32 partial tiles plus its pre-split
    operand (hi and lo planes, 64-sample stages; gram_pp_bf16x3)."""
    side = _cdiv(q, 128)
    ntiles, nkb, splits = side * (side + 1) // 2, _cdiv(n, 64), 1
    while ntiles * splits < 512 and nkb // (splits * 2) >= 8:
        splits *= 2
    splits = _cdiv(nkb, _cdiv(nkb, splits))
    return splits * ntiles * 128 * 128 * 4 + nkb * side * 128 * 2 * 64 * 2


def _gemm_split_bytes(m, k, cpx):
    """Split-K partials of the symmetric FP64 GEMM that forms the m x m
    T = D D^dagger (64 x 64 tiles; K split while fewer than 1024 workgroups and
    each part keeps >= 64 of k; csrc/linalg.hip gemm)."""
    ntn = _cdiv(m, 64)
    blocks, splits = ntn * (ntn + 1) // 2, 1
    while blocks * splits < 1024 and k // (splits * 2) >= 64:
        splits *= 2
    return splits * m * m * (16 if cpx else 8) if splits > 1 else 0


def footprint_bytes(num_samples, num_real_params, num_sites, complex_params, solver, chains, precision="fp64"):
    """HBM estimate (bytes) of one job: samples, walkers and s' strings; O and
    its centred copy; every workspace buffer of the solve route the equation
    takes.  ``solver``: "tdvp", or "minsr", whose route follows MinSR's
    default: the P x P dual Gram when the samples outnumber the (compact)
    parameters, else the N x N T form (Cholesky attempt on T, or on its real
    2N x 2N form for complex data, then the eigendecomposition).  Counted per
    route: the Gram's split-sample partial tiles, the shifted copy, the second
    factor and the inverse of the Cholesky route, the eigenvectors, the
    explicit Q and its product, the divide-and-conquer buffers, the complex
    tridiagonal path's buffers and, where the library's eigendecomposition
    runs, its workspace, plus the solver library handle of the job's thread
    (rocBLAS / rocSOLVER keep ~134 MB of device memory per handle; counted
    for every route, since a real solve needs the library once its spectrum
    is cut).  ``precision="mixed"`` adds the compensated Gram's buffers.  A
    25 % margin and 64 MiB cover rounding of buffer sizes."""
    n, pr = int(num_samples), int(num_real_params)
    pc = pr // 2 if complex_params else pr
    esz = 16 if complex_params else 8
    b = n * num_sites * 4 * 6 + chains * num_sites * 4    # samples, s' slots of the operator, walkers
    b += n * 16 * 8                                       # log psi, E_loc, weights, centred E_loc copies
    if solver == "tdvp":
        b += n * pr * 16 * 4                              # O (flat layout) and centred O, SNR data Q and R
        b += n * pc * 16 * 2                              # holomorphic G and the SNR projection M
        b += pr * pr * 8 * 5                              # S0 (complex), Z, S
        b += pc * pc * 8 * 10                             # A, eigenvectors, tridiagonal Z and Zc, D&C buffers
        b += pr * pr * 16 * 3                             # library eigendecomposition workspace
        b += _gram_slice_bytes(n, 2 * pr)
        if precision == "mixed":
            b += _bf16_gram_bytes(n, 2 * pc)
    elif solver == "minsr":
        b += n * pc * esz * 2                             # O and centred O
        b += (n // 64 + 1) * pc * esz                     # O^T e partial sums (64-sample chunks)
        if n > pc:                                        # dual form
            m = pc
            b += m * m * esz * 4                          # S, eigenvectors, library eigendecomposition workspace
            b += m * m * 8 * 12                           # shifted copy, second factor, inverse, Z, Q, Q x,
            #                                               D&C buffers (3), complex tridiagonal Zc (2)
            b += _gram_slice_bytes(n, m * (2 if complex_params else 1))
            if precision == "mixed":
                b += _bf16_gram_bytes(n, m * (2 if complex_params else 1))
        else:                                             # T form
            m = 2 * n if complex_params else n            # the Cholesky route's real form
            b += n * n * esz * 4                          # T, eigenvectors, library eigendecomposition workspace
            b += m * m * 8 * 7                            # shifted copy, factor, inverse, Z, D&C buffers (3)
            b += _gemm_split_bytes(n, pc, complex_params)
    else:
        raise ValueError("solver must be 'tdvp' or 'minsr'")
    b += LIBRARY_HANDLE_BYTES
    return int(1.25 * b) + (64 << 20)


# device memory a thread's solver library handle holds (measured: 134 MB)
LIBRARY_HANDLE_BYTES = 160 << 20
