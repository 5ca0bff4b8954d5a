"""ODE steppers (stepper.py:6-182) on device parameter vectors."""
import torch


def _norm(v):
    return torch.linalg.norm(v)


class Euler:
    """stepper.py:6-40."""

    def __init__(self, timeStep=1e-3):
        self.dt = timeStep

    def step(self, t, f, yInitial, **rhsArgs):
        dy = f(yInitial, t, **rhsArgs, intStep=0)
        return yInitial + self.dt * dy, self.dt


class Heun:
    """stepper.py:45-93."""

    def __init__(self, timeStep=1e-3):
        self.dt = timeStep

    def set_dt(self, timeStep):
        self.dt = timeStep

    def step(self, t, f, yInitial, **rhsArgs):
        dt = rhsArgs.get("dt", self.dt)
        k0 = f(yInitial, t, **rhsArgs, intStep=0)
        y = yInitial + dt * k0
        k1 = f(y, t + dt, **rhsArgs, intStep=1)
        return yInitial + 0.5 * dt * (k0 + k1), dt


class AdaptiveHeun:
    """stepper.py:98-182."""

    def __init__(self, timeStep=1e-3, tol=1e-8, maxStep=1):
        self.dt = timeStep
        self.tolerance = tol
        self.maxStep = maxStep

    def step(self, t, f, y, normFunction=_norm, **rhsArgs):
        fe = 0.5
        dt = self.dt
        yInitial = y.clone()
        while fe < 1.:
            y = yInitial.clone()
            k0 = f(y, t, **rhsArgs, intStep=0)
            y = y + dt * k0
            k1 = f(y, t + dt, **rhsArgs, intStep=1)
            dy0 = 0.5 * dt * (k0 + k1)
            y = y - 0.5 * dt * k0
            k10 = f(y, t + 0.5 * dt, **rhsArgs, intStep=2)
            dy1 = 0.25 * dt * (k0 + k10)
            y = yInitial + dy1
            k01 = f(y, t + 0.5 * dt, **rhsArgs, intStep=3)
            y = y + 0.5 * dt * k01
            k11 = f(y, t + dt, **rhsArgs, intStep=4)
            dy1 = dy1 + 0.25 * dt * (k01 + k11)
            updateDiff = float(normFunction(dy1 - dy0))
            # jnp division: a zero difference gives inf, i.e. the x2 step growth
            fe = self.tolerance / updateDiff if updateDiff != 0.0 else float("inf")
            tmp = 0.2 if 0.2 > 0.9 * fe ** 0.33333 else 0.9 * fe ** 0.33333
            if tmp > 2.:
                tmp = 2.
            realDt = dt
            dt *= tmp
            if dt > self.maxStep:
                dt = self.maxStep
        self.dt = dt
        return yInitial + dy1, realDt
