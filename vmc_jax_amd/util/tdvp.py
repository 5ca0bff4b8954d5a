"""TDVP right-hand side (tdvp.py:26-326) with the solve on the GPU.

The S / F build, FP64 eigendecomposition, SNR projection and the regulariser
loop (all cutoff candidates at once) run inside ``jv_tdvp_solve``.  For a
holomorphic network the gradient columns come in pairs (g, i g), and
``jv_tdvp_solve_holomorphic`` builds S from the Pc x Pc Gram G^H G and
diagonalises S through it (an exact rewrite, see include/jvmc_hip.h).
"""
import ctypes
import time

import numpy as np
import torch

from .. import _lib
from .. import mpi_wrapper as mpi
from ..sampler import ExactSampler


def realFun(x):
    return torch.real(x) if torch.is_tensor(x) else np.real(x)


def imagFun(x):
    return 0.5 * (x - x.conj())


class _Timer:
    def __init__(self, outp):
        self.outp = outp

    def start(self, name):
        if self.outp is not None:
            torch.cuda.synchronize()
            self.outp.start_timing(name)

    def stop(self, name):
        if self.outp is not None:
            torch.cuda.synchronize()
            self.outp.stop_timing(name)


# Smallest Pc for which each solve mode uses the holomorphic pair structure
# ('imag': S has +-eig(G^H G), always cheaper than the 2Pc eigendecomposition;
# 'real': S has every eigenvalue of G^H G twice; the complex Pc
# eigendecomposition beats the real 2Pc one once both run through the
# Householder tridiagonal paths, i.e. from Pc = 128).
HOLOMORPHIC_SOLVE = {"real": 128, "imag": 1}


class TDVP:
    def __init__(self, sampler, snrTol=2, pinvTol=1e-14, pinvCutoff=1e-8, makeReal='imag', rhsPrefactor=1.j,
                 diagonalShift=0., crossValidation=False, diagonalizeOnDevice=True):
        self.sampler = sampler
        self.snrTol = snrTol
        self.pinvTol = pinvTol
        self.pinvCutoff = pinvCutoff
        self.diagonalShift = diagonalShift
        self.rhsPrefactor = rhsPrefactor
        self.crossValidation = crossValidation
        self.diagonalizeOnDevice = diagonalizeOnDevice
        self.makeReal = imagFun if makeReal == 'imag' else realFun
        self._imag = makeReal == 'imag'
        self.metaData = None
        # mixed-precision form: set from the network (NQS(precision="mixed")) at
        # each call; holomorphic solves then take the compensated BF16 Gram
        self._mixed = False

    def set_diagonal_shift(self, delta):
        self.diagonalShift = delta

    def set_cross_validation(self, crossValidation=True):
        self.crossValidation = crossValidation

    def get_residuals(self):
        return self.metaData["tdvp_error"], self.metaData["tdvp_residual"]

    def get_snr(self):
        return self.metaData["SNR"]

    def get_spectrum(self):
        return self.metaData["spectrum"]

    def get_metadata(self):
        return self.metaData

    def get_energy_variance(self):
        return self.ElocVar0

    def get_energy_mean(self):
        return float(np.real(self.ElocMean0))

    def get_S(self):
        return self.S

    def _num_samples(self):
        """mpi.globNumSamples as the SNR uses it (tdvp.py:181).  The sampler's own
        copy is read so that jobs running concurrently in one process do not
        see each other's sample counts."""
        n = getattr(self.sampler, "globNumSamples", None)
        return mpi.globNumSamples if n is None else n

    def solve(self, Eloc, gradients):
        """tdvp.py:183-213 on centred data (SampledObs of E_loc and O)."""
        D = gradients._data
        if not torch.is_complex(D):
            D = D.to(torch.complex128)
        D = D.contiguous()
        e = Eloc._data.reshape(-1).to(torch.complex128).contiguous()
        w = gradients._weights
        n, pr = D.shape
        dev = D.device
        update = torch.empty(pr, dtype=torch.float64, device=dev)
        S = torch.empty((pr, pr), dtype=torch.complex128 if self._imag else torch.float64, device=dev)
        F = torch.empty(pr, dtype=torch.complex128, device=dev)
        ev = torch.empty(pr, dtype=torch.float64, device=dev)
        snr = torch.empty(pr, dtype=torch.float64, device=dev)
        x = (ctypes.c_double * 2)(complex(self.rhsPrefactor).real, complex(self.rhsPrefactor).imag)
        info = _lib.JvSolveInfo()
        common = (x, int(self._imag), float(self.diagonalShift), int(not isinstance(self.sampler, ExactSampler)),
                  float(self._num_samples()), float(self.snrTol), float(self.pinvTol), float(self.pinvCutoff),
                  _lib.ptr(update), _lib.ptr(S), _lib.ptr(F), _lib.ptr(ev), _lib.ptr(snr), ctypes.byref(info),
                  _lib.stream_ptr())
        cols = getattr(self, "_holo_cols", None)
        if cols is not None and cols[0].numel() * 2 == pr and \
                cols[0].numel() >= HOLOMORPHIC_SOLVE["imag" if self._imag else "real"]:
            fn = "jv_tdvp_solve_holomorphic_mixed" if self._mixed else "jv_tdvp_solve_holomorphic"
            _lib.call(fn, _lib.ptr(D), _lib.ptr(e), _lib.ptr(w), n, pr, _lib.ptr(cols[0]),
                      _lib.ptr(cols[1]), cols[0].numel(), *common)
        else:
            _lib.call("jv_tdvp_solve", _lib.ptr(D), _lib.ptr(e), _lib.ptr(w), n, pr, *common)
        self.S, self.F, self.ev, self.snr = S, F, ev, snr
        self.ElocMean = complex(Eloc.mean()[0].item())
        self.ElocVar = info.e_var
        self._tdvp_error = info.tdvp_error
        return update, info.residual, info.pinv_cutoff

    def __call__(self, netParameters, t, *, psi, hamiltonian, **rhsArgs):
        """tdvp.py:219-326."""
        from ..stats import SampledObs
        tmpParameters = psi.get_parameters()
        psi.set_parameters(netParameters)
        self._holo_cols = psi.holomorphic_columns() if psi.holomorphic else None
        self._mixed = getattr(psi, "precision", "fp64") == "mixed"
        timer = _Timer(rhsArgs.get("outp", None))
        numSamples = rhsArgs.get("numSamples", None)
        timer.start("sampling")
        sampleConfigs, sampleLogPsi, p = self.sampler.sample(numSamples=numSamples)
        timer.stop("sampling")
        timer.start("compute Eloc")
        Eloc = hamiltonian.get_O_loc(sampleConfigs, psi, sampleLogPsi, t)
        timer.stop("compute Eloc")
        Eloc = SampledObs(Eloc, p)
        timer.start("compute gradients")
        sampleGradients = psi.gradients(sampleConfigs)
        timer.stop("compute gradients")
        sampleGradients = SampledObs(sampleGradients, p)
        timer.start("solve TDVP eqn.")
        update, solverResidual, pinvCutoff = self.solve(Eloc, sampleGradients)
        timer.stop("solve TDVP eqn.")
        psi.set_parameters(tmpParameters)
        if rhsArgs.get("intStep", None) == 0:
            self.ElocMean0 = self.ElocMean
            self.ElocVar0 = self.ElocVar
            self.metaData = {"tdvp_error": self._tdvp_error, "tdvp_residual": solverResidual,
                             "pinv_cutoff": pinvCutoff, "SNR": self.snr, "spectrum": self.ev}
            if self.crossValidation:
                update = self._cross_validate(Eloc, sampleGradients)
        return update

    def _cross_validate(self, Eloc, sampleGradients):
        """tdvp.py:306-324: the update of the even samples, validated on the odd
        ones.  Its tdvp_error uses the odd samples' S0, F0 with the full sample's
        ElocVar0 (tdvp.py:102-104 after the odd-sample get_tdvp_equation); the
        residual ratio divides by the full solve's residual.  The full solve is
        repeated, as in the reference, so the solver state ends on the full
        sample."""
        Eloc1, grads1 = Eloc.subset(start=0, step=2), sampleGradients.subset(start=0, step=2)
        Eloc2, grads2 = Eloc.subset(start=1, step=2), sampleGradients.subset(start=1, step=2)
        update_1, _, _ = self.solve(Eloc1, grads1)
        D2 = grads2._data
        if not torch.is_complex(D2):
            D2 = D2.to(torch.complex128)
        D2 = D2.contiguous()
        e2 = Eloc2._data.reshape(-1).to(torch.complex128).contiguous()
        x = (ctypes.c_double * 2)(complex(self.rhsPrefactor).real, complex(self.rhsPrefactor).imag)
        out = (ctypes.c_double * 2)()
        _lib.call("jv_tdvp_cross_validate", _lib.ptr(D2), _lib.ptr(e2), D2.shape[0], D2.shape[1], x,
                  int(self._imag), float(self.diagonalShift), _lib.ptr(update_1), out, _lib.stream_ptr())
        validation_tdvpErr = abs(1.0 + out[1] / self.ElocVar0)
        update, solverResidual, _ = self.solve(Eloc, sampleGradients)
        with np.errstate(divide="ignore", invalid="ignore"):  # jnp semantics: x / 0 -> inf
            self.crossValidationFactor_residual = np.float64(out[0]) / np.float64(solverResidual)
            self.crossValidationFactor_tdvpErr = np.float64(validation_tdvpErr) / np.float64(self.metaData["tdvp_error"])
        self.metaData["tdvp_residual_cross_validation_ratio"] = self.crossValidationFactor_residual
        self.metaData["tdvp_error_cross_validation_ratio"] = self.crossValidationFactor_tdvpErr
        return update
