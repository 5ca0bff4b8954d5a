"""Batched field sweep: many independent RBM states advanced together.

Each instance is a complete jVMC set-up (NQS of a CpxRBM / RBM, a
BranchFreeOperator, Metropolis sampling with single spin flips, a TDVP or
MinSR equation and an Euler step); the reference runs each as its own driver
(``util.util.ground_state_search``-style loops over ``sampler.sample``,
``Operator.get_O_loc``, ``NQS.gradients`` and ``TDVP.__call__``).  Here:

* sampling, log psi + O_k and O_loc of **all** instances are one launch each
  (``jv_batch_step``: a device-resident instance table maps every workgroup to
  one instance), captured once into a HIP graph and replayed every step;
* the launch shape of each phase (waves per workgroup, samples per wave) is
  chosen by ``autotune`` from timed trials on the actual instance mix and kept
  in a JSON cache keyed by the mix;
* the per-instance solves (FP64 eigendecomposition / Cholesky) run on a pool
  of worker threads, each with its own HIP stream, after the graph's event,
  and update the parameters in place (the graph reads them at replay).

Every instance's samples, log psi, O_loc and O_k equal those of its own
``MCSampler`` / ``get_O_loc`` / ``gradients`` calls with the same key and call
index (the same device routines); the tests check this bit for bit.
"""
import ctypes
import hashlib
import json
import os
import threading
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import torch

from . import _lib
from . import global_defs
from . import mpi_wrapper as mpi
from .stats import SampledObs

PHASES = {"sample": _lib.JV_BATCH_SAMPLE, "grad": _lib.JV_BATCH_GRAD, "eloc": _lib.JV_BATCH_ELOC}


class _SamplerInfo:
    """What TDVP.solve reads from its sampler: the global sample count."""

    def __init__(self, n):
        self.globNumSamples = n


class Instance:
    """One state of the sweep: an initialised NQS of a (Cpx)RBM, its operator,
    a ``TDVP`` (gradients in jVMC's layout) or ``MinSR`` (compact layout)
    equation whose sampler argument is replaced, the Euler time step, and the
    MCSampler arguments (key, chains, samples, sweeps; mu = 2)."""

    def __init__(self, psi, H, eq, dt, key, numChains, numSamples, thermalizationSweeps=10, sweepSteps=10):
        if psi._sym is not None or psi.desc.kind not in (_lib.JV_NET_CPXRBM, _lib.JV_NET_RBM):
            raise NotImplementedError("the batched sweep runs RBM / CpxRBM networks")
        self.psi, self.H, self.eq, self.dt = psi, H, eq, float(dt)
        self.L = int(psi.desc.num_sites)
        self.key = int(key)
        self.C = int(numChains)
        self.spc, self.n = mpi.distribute_sampling(numSamples, localDevices=1, numChainsPerDevice=self.C)
        self.therm = int(thermalizationSweeps * sweepSteps)
        self.sweep = int(sweepSteps)
        dev = global_defs.device()
        self.layout = 0 if type(eq).__name__ == "MinSR" else 1
        self.walkers = torch.zeros((self.C, self.L), dtype=torch.int32, device=dev)
        self.configs = torch.empty((self.n, self.L), dtype=torch.int32, device=dev)
        self.logpsi = torch.empty(self.n, dtype=torch.complex128, device=dev)
        self.counters = torch.zeros(2, dtype=torch.int64, device=dev)
        self.eloc = torch.empty(self.n, dtype=torch.complex128, device=dev)
        cols = psi.numRealParameters if self.layout == 1 else psi.numParameters
        cpx = self.layout == 1 or psi.holomorphic
        self.grad = torch.empty((self.n, cols), dtype=torch.complex128 if cpx else torch.float64, device=dev)
        self.p = torch.full((1, self.n), 1.0 / self.n, dtype=torch.float64, device=dev)  # mu = 2
        if not H.compiled:
            H.compile()
        self.pref = H.prefactors(0.0)
        eq.sampler = _SamplerInfo(self.n)
        self.energy = None
        self.calls = 0  # sampling passes done (the device advances its own copy in step)

    def item(self):
        it = _lib.JvBatchItem()
        it.net = self.psi.desc
        it.params = self.psi.params.data_ptr()
        it.op = self.H.table
        it.prefactor = self.pref.data_ptr()
        a = it.mc
        a.num_chains, a.therm_steps, a.samples_per_chain, a.sweep_steps = self.C, self.therm, self.spc, self.sweep
        a.mu, a.seed, a.call_index, a.stream = 2.0, self.key & 0xFFFFFFFFFFFFFFFF, self.calls & 0xFFFFFFFF, 0
        it.walkers, it.configs = self.walkers.data_ptr(), self.configs.data_ptr()
        it.logpsi, it.counters = self.logpsi.data_ptr(), self.counters.data_ptr()
        it.eloc, it.grad, it.grad_layout = self.eloc.data_ptr(), self.grad.data_ptr(), self.layout
        return it

    def signature(self):
        d = self.psi.desc
        return [self.L, int(d.num_hidden), int(d.is_complex), int(self.H.table.num_ops), self.n, self.C, self.layout]

    def rhs(self):
        """The instance's right-hand side theta_dot from the batch buffers
        (tdvp.py:134-213 or minsr.py:53-80) as a flat real vector."""
        Eloc = SampledObs(self.eloc.reshape(1, -1), self.p)
        grads = SampledObs(self.grad.reshape((1,) + tuple(self.grad.shape)), self.p)
        if self.layout == 1:
            self.eq._holo_cols = self.psi.holomorphic_columns() if self.psi.holomorphic else None
            upd, _, _ = self.eq.solve(Eloc, grads)
            self.energy = self.eq.ElocMean
        else:
            upd = self.eq.solve(Eloc, grads, self.psi.holomorphic, self.psi)
            self.energy = complex(Eloc.mean()[0].item())
        return self.psi._param_unflatten(upd)

    def solve_and_update(self):
        """rhs() and the Euler update theta += dt * theta_dot, in place (the
        batch reads the parameters from this buffer)."""
        self.psi.params.add_(self.dt * self.rhs())


def tune_cache_path():
    return os.environ.get("VMC_JAX_AMD_TUNE_CACHE",
                          os.path.join(os.path.expanduser("~"), ".cache", "vmc_jax_amd", "batch_tune.json"))


def _load_cache(path):
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


class BatchedSweep:
    SHAPES = {"sample": [(w, 1) for w in (1, 2, 4, 8)],
              "grad": [(w, s) for w in (1, 2, 4, 8) for s in (1, 2, 4, 8)],
              "eloc": [(w, s) for w in (1, 2, 4, 8) for s in (1, 2, 4, 8)]}

    def __init__(self, instances, workers=8, tune=True, graph=True, pool=None):
        """tune: True runs the trials unless the mix is in the cache; "cache" uses
        cached shapes only (default shapes otherwise); False keeps the defaults.
        pool: a ThreadPoolExecutor for the solves (shared by a server that
        rebuilds the batch when its jobs change)."""
        _lib.check(_lib.lib.jv_init(torch.cuda.current_device()), "jv_init")
        self.inst = list(instances)
        self.workers = int(workers)
        self._items = (_lib.JvBatchItem * len(self.inst))(*[i.item() for i in self.inst])
        self.batch = ctypes.c_void_p()
        _lib.check(_lib.lib.jv_batch_create(self._items, len(self.inst), ctypes.byref(self.batch)), "jv_batch_create")
        self.stream = torch.cuda.Stream()
        self.shapes = {"sample": (4, 1), "grad": (4, 1), "eloc": (4, 1)}
        self.tuned_from = None
        if tune:
            self.autotune(trials=tune is True)
        self.graph = None
        if graph:
            g = ctypes.c_void_p()
            torch.cuda.synchronize()
            _lib.check(_lib.lib.jv_graph_capture(self.batch, ctypes.c_void_p(self.stream.cuda_stream), ctypes.byref(g)),
                       "jv_graph_capture")
            self.graph = g
        self._own_pool = pool is None
        self._pool = pool if pool is not None else ThreadPoolExecutor(max_workers=self.workers)
        self._tls = threading.local()
        self.step_ms = []
        self._phase_events = []

    def key(self):
        sig = json.dumps(sorted(i.signature() for i in self.inst))
        return hashlib.sha1(sig.encode()).hexdigest()

    def _set_shape(self, batch, phase, shape):
        _lib.check(_lib.lib.jv_batch_set_shape(batch, PHASES[phase], shape[0], shape[1]), "jv_batch_set_shape")

    def autotune(self, reps=3, trials=True):
        """Launch shape per phase: the fastest of the candidate shapes over `reps`
        timed launches on this instance mix, on a scratch batch that shares the
        output buffers but has its own walkers, counters and call indices (the
        real chains are not advanced).  Results are cached by the mix."""
        path = tune_cache_path()
        cache = _load_cache(path)
        k = self.key()
        if k in cache:
            self.shapes = {ph: tuple(v) for ph, v in cache[k].items()}
            self.tuned_from = "cache"
        elif not trials:
            return
        else:
            scratch = []  # (walkers, counters, item): kept alive while the scratch batch exists
            for it, ins in zip(self._items, self.inst):
                w, c = ins.walkers.clone(), torch.zeros_like(ins.counters)
                s = _lib.JvBatchItem.from_buffer_copy(it)
                s.walkers, s.counters = w.data_ptr(), c.data_ptr()
                scratch.append((w, c, s))
            arr = (_lib.JvBatchItem * len(scratch))(*[t[2] for t in scratch])
            tb = ctypes.c_void_p()
            _lib.check(_lib.lib.jv_batch_create(arr, len(scratch), ctypes.byref(tb)), "jv_batch_create")
            st = ctypes.c_void_p(self.stream.cuda_stream)
            best = {}
            try:
                for ph in ("sample", "grad", "eloc"):
                    times = {}
                    for shape in self.SHAPES[ph]:
                        try:
                            self._set_shape(tb, ph, shape)
                        except _lib.HipError:
                            continue  # e.g. more LDS than a workgroup has
                        ts = []
                        for _ in range(reps):
                            a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                            a.record(self.stream)
                            _lib.check(_lib.lib.jv_batch_phase(tb, PHASES[ph], st), "jv_batch_phase")
                            b.record(self.stream)
                            b.synchronize()
                            ts.append(a.elapsed_time(b))
                        times[shape] = min(ts)
                    best[ph] = min(times, key=times.get)
            finally:
                _lib.lib.jv_batch_destroy(tb)
            self.shapes = best
            self.tuned_from = "trials"
            cache[k] = {ph: list(v) for ph, v in best.items()}
            try:
                os.makedirs(os.path.dirname(path), exist_ok=True)
                with open(path + ".tmp", "w") as f:
                    json.dump(cache, f, indent=1)
                os.replace(path + ".tmp", path)
            except OSError:
                pass  # an unwritable cache only costs the trials next time
        for ph, shape in self.shapes.items():
            self._set_shape(self.batch, ph, shape)

    def _worker_stream(self):
        s = getattr(self._tls, "stream", None)
        if s is None:
            s = self._tls.stream = torch.cuda.Stream()
        return s

    def _solve(self, idx, ready, fn):
        s = self._worker_stream()
        errors = {}
        with torch.cuda.stream(s):
            s.wait_event(ready)
            for i in idx:
                try:
                    fn(self.inst[i])
                except Exception as exc:  # reported per instance, the others continue
                    errors[i] = exc
            done = torch.cuda.Event()
            done.record(s)
        return done, errors

    def sample_phase(self):
        """Sampling, log psi + O_k and O_loc of every instance on the batch stream."""
        st = ctypes.c_void_p(self.stream.cuda_stream)
        self.stream.wait_stream(torch.cuda.current_stream())
        t0 = torch.cuda.Event(enable_timing=True)
        t0.record(self.stream)
        if self.graph is not None:
            _lib.check(_lib.lib.jv_graph_launch(self.graph, st), "jv_graph_launch")
        else:
            _lib.check(_lib.lib.jv_batch_step(self.batch, st), "jv_batch_step")
        for i in self.inst:
            i.calls += 1
        ready = torch.cuda.Event(enable_timing=True)
        ready.record(self.stream)
        self._phase_events.append((t0, ready))
        return ready

    def phase_ms(self):
        """Device time of each batched sampling / O_k / O_loc pass so far."""
        return [a.elapsed_time(b) for a, b in self._phase_events]

    def evaluate(self, fn):
        """One right-hand-side evaluation of every instance: the batched
        sampling / O_k / O_loc pass, then fn(instance) for each instance on the
        worker threads (largest solves first, round-robin), each on its
        worker's stream.  Returns {instance index: exception} of failed ones."""
        ready = self.sample_phase()
        order = sorted(range(len(self.inst)), key=lambda i: -self.inst[i].grad.shape[1])
        parts = [order[k::self.workers] for k in range(self.workers)]
        futs = [self._pool.submit(self._solve, p, ready, fn) for p in parts if p]
        cur = torch.cuda.current_stream()
        errors = {}
        for f in futs:
            done, err = f.result()
            cur.wait_event(done)
            errors.update(err)
        cur.synchronize()
        return errors

    def step(self):
        """One Euler step of every instance."""
        t0 = time.perf_counter()
        errors = self.evaluate(lambda i: i.solve_and_update())
        if errors:
            raise next(iter(errors.values()))
        self.step_ms.append(1000.0 * (time.perf_counter() - t0))

    def close(self):
        if self._own_pool:
            self._pool.shutdown(wait=True)
        if self.graph is not None:
            _lib.lib.jv_graph_destroy(self.graph)
            self.graph = None
        if self.batch:
            _lib.lib.jv_batch_destroy(self.batch)
            self.batch = ctypes.c_void_p()
