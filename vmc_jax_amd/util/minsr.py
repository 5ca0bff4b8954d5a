"""MinSR right-hand side (minsr.py:16-165) with the solve on the GPU.

Real-parameter networks: T = O O^T (+ shift) is the [Re O; Im O] stacking of
minsr.py:68-78 with its identically-zero imaginary half removed, which leaves
the pseudo-inverse update unchanged.  Holomorphic networks: the compact
gradients G give T = 2 G G^dagger for the [G, iG] layout, and the update is
expanded back to jVMC's flat layout.
"""
import ctypes

import numpy as np
import torch

from .. import _lib
from .tdvp import _Timer


class MinSR:
    def __init__(self, sampler, pinvTol=1e-14, diagonalShift=0., diagonalizeOnDevice=True, solver="T"):
        """``solver``: "T" diagonalises the N x N tangent kernel as the reference
        does; "dual" diagonalises the P x P Gram D^dagger D, which gives the same
        update (identical nonzero spectrum); "auto" picks the smaller."""
        if solver not in ("T", "dual", "auto"):
            raise ValueError("solver must be 'T', 'dual' or 'auto'")
        self.solver = solver
        self.sampler = sampler
        self.pinvTol = pinvTol
        self.diagonalShift = diagonalShift
        self.diagonalizeOnDevice = diagonalizeOnDevice
        self.metaData = None
        self.last_info = None

    def set_pinv_tol(self, tol):
        self.pinvTol = tol

    def get_metadata(self):
        return self.metaData

    def get_energy_variance(self):
        return self.ElocVar0

    def get_energy_mean(self):
        return float(np.real(self.ElocMean0))

    def solve(self, eloc, gradients, holomorphic, psi=None):
        """minsr.py:53-80 on centred data; gradients in the compact layout."""
        D = gradients._data.contiguous()
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
        _lib.call("jv_minsr_solve_dual" if dual else "jv_minsr_solve", _lib.ptr(D), _lib.ptr(e), n, p, int(cpx), float(self.pinvTol),
                  float(self.diagonalShift), _lib.ptr(upd), ctypes.byref(info), _lib.stream_ptr())
        self.last_info = info
        if cpx:
            # T_ref = 2 G G^dagger: pinv scales by 1/2; [v, -i v] layout per leaf
            return psi.expand_compact(0.5 * upd)
        return upd

    def __call__(self, netParameters, t, *, psi, hamiltonian, **rhsArgs):
        """minsr.py:82-165."""
        from ..stats import SampledObs
        tmpParameters = psi.get_parameters()
        psi.set_parameters(netParameters)
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
        sampleGradients = SampledObs(psi.gradients(sampleConfigs, layout=0), p)
        timer.stop("compute gradients")
        timer.start("solve MinSR eqn.")
        update = self.solve(Eloc, sampleGradients, psi.holomorphic, psi)
        timer.stop("solve MinSR eqn.")
        psi.set_parameters(tmpParameters)
        if rhsArgs.get("intStep", None) == 0:
            self.ElocMean0 = complex(Eloc.mean()[0].item())
            self.ElocVar0 = float(Eloc.var()[0].item())
            self.metaData = {}
        return update
