"""MinSR right-hand side (minsr.py:16-165) with the solve on the GPU.

Real-parameter networks: T = O O^T (+ shift) is the [Re O; Im O] stacking of
minsr.py:68-78 with its identically-zero imaginary half removed, which leaves
the pseudo-inverse update unchanged.  Holomorphic networks: the compact
gradients G give T = 2 G G^dagger for the [G, iG] layout, and the update is
expanded back to jVMC's flat layout.
"""
import ctypes
import os

import numpy as np
import torch

from .. import _lib
from .tdvp import _Timer


class _Outcome:
    """The jv_solve_info of a non-blocking solve, copied by the library into
    pinned host memory in stream order; read once the event recorded after
    the solve has completed.  `hint` and `full` are the route bookkeeping's
    arguments for that solve."""

    def __init__(self, buf, event, hint, full):
        self._buf, self._event, self.hint, self.full = buf, event, hint, full
        self._info = None

    def resolve(self):
        if self._info is None:
            self._event.synchronize()
            info = _lib.JvSolveInfo.from_buffer_copy(ctypes.string_at(self._buf.data_ptr(),
                                                                      ctypes.sizeof(_lib.JvSolveInfo)))
            self._buf = self._event = None
            if info.iterations < 0:
                raise _lib.HipError("jv_minsr_solve_dual_deferred_gram_async failed: Cholesky factorisation "
                                    "failed (info %d)" % (-1 - info.iterations))
            self._info = info
        return self._info

    def energy_mean(self):
        return float(self.resolve().e_mean[0])


class MinSR:
    def __init__(self, sampler, pinvTol=1e-14, diagonalShift=0., diagonalizeOnDevice=True, solver="auto"):
        """``solver``: "T" diagonalises the N x N tangent kernel as the reference
        does; "dual" diagonalises the P x P Gram D^dagger D, which gives the same
        update (identical nonzero spectrum); "auto" (default) picks the smaller,
        so a jVMC script gets the dual form whenever N > P."""
        if solver not in ("T", "dual", "auto"):
            raise ValueError("solver must be 'T', 'dual' or 'auto'")
        self.solver = solver
        self.sampler = sampler
        self.pinvTol = pinvTol
        self.diagonalShift = diagonalShift
        self.diagonalizeOnDevice = diagonalizeOnDevice
        self.metaData = None
        self._pending = None  # _Outcome of the last overlapped solve, not yet read
        self._energy_src = None  # _Outcome whose E_loc mean and variance are ElocMean0 / ElocVar0
        self.last_info = None
        self._chol_skip = 0

    # The overlapped solve returns before the device has finished it, and its
    # info is read when first needed: by the next solve's route choice (after
    # the next step's sampling is queued), or by a read of last_info,
    # ElocMean0 or ElocVar0.  The host thus never waits at the end of a step.
    @property
    def last_info(self):
        pend = self._pending
        if pend is not None:
            self._pending = None
            self._last_info = pend.resolve()
            self._note_route(pend.hint, pend.full, self._last_info)
        return self._last_info

    @last_info.setter
    def last_info(self, info):
        if self._pending is not None:
            self.last_info  # read (and release) the outstanding outcome first
        self._last_info = info

    def last_outcome(self):
        """The last overlapped solve's outcome (resolved on first read), or None."""
        return self._pending if self._pending is not None else self._energy_src

    def _energy(self):
        src = self._energy_src
        if src is not None:
            if src is self._pending:
                self.last_info  # settles the route bookkeeping first
            info = src.resolve()
            self._energy_src = None
            self._ElocMean0 = complex(info.e_mean[0], info.e_mean[1])
            self._ElocVar0 = float(info.e_var)

    @property
    def ElocMean0(self):
        self._energy()
        return self._ElocMean0

    @ElocMean0.setter
    def ElocMean0(self, v):
        self._energy()  # an outstanding outcome is read (and released) first
        self._ElocMean0 = v

    @property
    def ElocVar0(self):
        self._energy()
        return self._ElocVar0

    @ElocVar0.setter
    def ElocVar0(self, v):
        self._energy()
        self._ElocVar0 = v

    # Solves try the Cholesky route first (exact when no eigenvalue falls below
    # the cutoff).  After a refused attempt, the attempt is skipped for the
    # next CHOLESKY_RETRY - 1 solves: the spectrum changes slowly from step to
    # step, and a refused attempt costs a factorisation (and, in the T form, a
    # host synchronisation before the eigendecomposition).  A skipped solve
    # whose eigendecomposition cut nothing the route would not make the next
    # solve try again.  The update is the same either way.
    CHOLESKY_RETRY = 8
    # The holomorphic T form (complex T = G G^dagger, N < P) tries the route
    # too: centred, uniformly weighted samples leave T one null vector, the
    # constant 1/sqrt(N), which the device deflates (the pseudo-inverse cuts
    # its eigenvalue); the attempt succeeds when nothing else is cut.  Set to
    # False to go straight to the eigendecomposition.
    try_cholesky_holomorphic = True

    def _cholesky_hint(self):
        self.last_info  # the previous solve's route first
        if self._chol_skip > 0:
            self._chol_skip -= 1
            return 0
        return 1

    def _note_route(self, tried, full, info):
        """Back-off of the dual solve; `full`: the rank of a spectrum the
        route accepts (P for the P x P Gram)."""
        if info.iterations != 0:
            return
        if tried:
            self._chol_skip = self.CHOLESKY_RETRY - 1
        elif info.rank >= full:
            self._chol_skip = 0

    def set_pinv_tol(self, tol):
        self.pinvTol = tol

    def get_metadata(self):
        return self.metaData

    def get_energy_variance(self):
        return self.ElocVar0

    def get_energy_mean(self):
        return float(np.real(self.ElocMean0))

    def solve(self, eloc, gradients, holomorphic, psi=None):
        """minsr.py:53-80 on centred data; gradients in the compact layout (FP32
        data of the mixed form are widened to FP64 here; the mixed Gram runs on
        the overlapped path of __call__)."""
        D = gradients._data.to(torch.float64).contiguous() if gradients._data.dtype == torch.float32 \
            else gradients._data.contiguous()
        e = eloc._data.reshape(-1).to(torch.complex128).contiguous()
        if not holomorphic and torch.is_complex(D):
            # real parameters with a complex output (a SymNet with complex factors):
            # the stacked real system [Re O; Im O], [Re e; Im e] of minsr.py:68-78
            D = torch.cat([D.real, D.imag], dim=0).contiguous()
            e = torch.cat([e.real, e.imag]).to(torch.complex128).contiguous()
        n, p = D.shape
        cpx = torch.is_complex(D)
        upd = torch.empty(p, dtype=torch.complex128 if cpx else torch.float64, device=D.device)
        info = _lib.JvSolveInfo()
        dual = self.solver == "dual" or (self.solver == "auto" and n > p)
        if dual:
            hint = self._cholesky_hint() if not cpx else 0
            _lib.call("jv_minsr_solve_dual", _lib.ptr(D), _lib.ptr(e), n, p, int(cpx), float(self.pinvTol),
                      float(self.diagonalShift), hint, _lib.ptr(upd), ctypes.byref(info), _lib.stream_ptr())
        else:
            # T form: a refused attempt costs one factorisation beside an
            # eigendecomposition many times its price, so every solve that can
            # pass tries (T has rank <= P: with N - 1 > P it is always cut)
            hint = int((not cpx or self.try_cholesky_holomorphic) and n - 1 <= p) and self._cholesky_hint()
            _lib.call("jv_minsr_solve", _lib.ptr(D), _lib.ptr(e), n, p, int(cpx), float(self.pinvTol),
                      float(self.diagonalShift), hint, _lib.ptr(upd), ctypes.byref(info), _lib.stream_ptr())
        self.last_info = info
        # back-off after a refused attempt; a spectrum the route accepts keeps
        # P (dual) or N - 1 (T form, the deflated centring vector) eigenvalues
        self._note_route(hint, p if dual else n - 1, info)
        if cpx:
            # T_ref = 2 G G^dagger: pinv scales by 1/2; [v, -i v] layout per leaf
            return psi.expand_compact(0.5 * upd)
        return upd

    # Real dual solves compute E_loc on a side stream while the solve builds and
    # reduces the Gram matrix (set False to run the phases one after another).
    overlap = True
    def _overlapped(self, psi, n, outp):
        """The real dual solve runs with E_loc still in flight (not when per-phase
        timing is requested: phases then overlap)."""
        if not self.overlap or outp is not None or psi.holomorphic or getattr(psi, "_cpx_weights", False):
            return False
        return self.solver == "dual" or (self.solver == "auto" and n > psi.numParameters)

    def _eloc_first(self, psi):
        """Where E_loc starts on the side stream (10x10 CNN measurements).
        FP64, when the Cholesky route is tried: right after sampling, beside
        O_k, the Gram and the factorisation (128 steps/s against 125 when it
        starts after the Gram).  FP64, when the attempt is skipped (a cut
        spectrum expected): after the Gram, beside the latency-bound
        Householder reduction (55.6 against 50.6 steps/s).  Mixed form (FP32
        O): after the Gram (202 against 195 steps/s on the Cholesky route,
        equal on cut spectra).  VMC_JAX_AMD_MINSR_ELOC_FIRST = 0 / 1
        overrides."""
        env = os.environ.get("VMC_JAX_AMD_MINSR_ELOC_FIRST")
        if env is not None:
            return env == "1"
        self.last_info  # the previous solve's route first
        return not psi.fp32_gradients and self._chol_skip == 0

    def _dual_gram(self, gradients):
        """S = D^T D (+ diagonalShift) of the centred data on the current
        stream (FP32 data of the mixed form: the compensated BF16 Gram)."""
        D = gradients._data.contiguous()
        n, P = D.shape
        S = torch.empty((P, P), dtype=torch.float64, device=D.device)
        _lib.call("jv_minsr_dual_gram", _lib.ptr(D), int(D.dtype == torch.float32), n, P, 0,
                  float(self.diagonalShift), _lib.ptr(S), _lib.stream_ptr())
        return D, S

    def _solve_deferred(self, Eraw, ready, D, S, p):
        """The non-blocking solve: its info is read later (see last_info)."""
        n, P = D.shape
        upd = torch.empty(P, dtype=torch.float64, device=D.device)
        hint = self._cholesky_hint()
        buf = torch.empty(ctypes.sizeof(_lib.JvSolveInfo), dtype=torch.uint8, pin_memory=True)
        _lib.call("jv_minsr_solve_dual_deferred_gram_async", _lib.ptr(D), int(D.dtype == torch.float32),
                  _lib.ptr(S), _lib.ptr(Eraw), _lib.ptr(p), ready._as_parameter_, n, P, 0, float(self.pinvTol),
                  float(self.diagonalShift), hint, _lib.ptr(upd), ctypes.c_void_p(buf.data_ptr()),
                  _lib.stream_ptr())
        done = torch.cuda.Event()
        done.record()
        self._pending = _Outcome(buf, done, hint, P)
        return upd

    def __call__(self, netParameters, t, *, psi, hamiltonian, **rhsArgs):
        """minsr.py:82-165."""
        from ..stats import SampledObs
        tmpParameters = psi.get_parameters()
        psi.set_parameters(netParameters)
        outp = rhsArgs.get("outp", None)
        timer = _Timer(outp)
        numSamples = rhsArgs.get("numSamples", None)
        timer.start("sampling")
        side = None
        if self.overlap and outp is None and psi.fp32_gradients and getattr(self.sampler, "defers_logpsi", False):
            # the mixed form's FP64 log psi of the samples (needed by E_loc
            # only) on the side stream, beside O_k and the Gram
            if getattr(self, "_side", None) is None:
                self._side = torch.cuda.Stream()
            side = self._side
            sampleConfigs, sampleLogPsi, p = self.sampler.sample(numSamples=numSamples, logpsi_stream=side)
        else:
            sampleConfigs, sampleLogPsi, p = self.sampler.sample(numSamples=numSamples)
        timer.stop("sampling")
        overlapped = self._overlapped(psi, sampleConfigs.shape[1], outp)
        if side is not None and not overlapped:
            torch.cuda.current_stream().wait_stream(side)
        if overlapped:
            # O_k and the Gram first, then E_loc on a side stream beside the
            # latency-bound factorisation (or reduction) of the Gram; the solve
            # waits for E_loc only where it needs it.  Starting E_loc after
            # the Gram keeps the two throughput-bound kernels from sharing the
            # CUs in an order the hardware picks.
            main = torch.cuda.current_stream()
            if getattr(self, "_side", None) is None:
                self._side = torch.cuda.Stream()
            side = self._side
            eloc_first = self._eloc_first(psi)
            if not eloc_first:
                sampleGradients = SampledObs(psi.gradients(sampleConfigs, layout=0), p)
                D, S = self._dual_gram(sampleGradients)
            side.wait_stream(main)
            with torch.cuda.stream(side):
                Eraw = hamiltonian.get_O_loc(sampleConfigs, psi, sampleLogPsi, t)
                Eraw = Eraw.reshape(-1).to(torch.complex128).contiguous()
                ready = torch.cuda.Event()
                ready.record(side)
            if eloc_first:
                sampleGradients = SampledObs(psi.gradients(sampleConfigs, layout=0), p)
                D, S = self._dual_gram(sampleGradients)
            w = p.reshape(-1).to(torch.float64).contiguous()
            update = self._solve_deferred(Eraw, ready, D, S, w)
            main.wait_stream(side)
            Eraw.record_stream(main)
            Eloc = None  # mean and variance come with the solve info
        else:
            timer.start("compute Eloc")
            Eloc = hamiltonian.get_O_loc(sampleConfigs, psi, sampleLogPsi, t)
            timer.stop("compute Eloc")
            Eloc = SampledObs(Eloc, p)
            timer.start("compute gradients")
            sampleGradients = SampledObs(psi.gradients(sampleConfigs, layout=0), p)
            timer.stop("compute gradients")
            timer.start("solve MinSR eqn.")
            update = self.solve(Eloc, sampleGradients, psi.holomorphic, psi)
            timer.stop("solve MinSR eqn.")
        psi.set_parameters(tmpParameters)
        if rhsArgs.get("intStep", None) == 0:
            if Eloc is None:
                self._energy_src = self._pending  # read on first use
            else:
                self.ElocMean0 = complex(Eloc.mean()[0].item())
                self.ElocVar0 = float(Eloc.var()[0].item())
            self.metaData = {}
        return update
