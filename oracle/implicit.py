"""ImplicitEuler on the oracle's model engine (TEST INFRASTRUCTURE ONLY).

Restates pf.ImplicitEuler (pf/implicitEuler.go:12-208):

* the residual updateEquation (implicitEuler.go:68-94): with the fields set to x, their
  transforms x^, and R^(x) = GetRHS, D = GetDenum,
      update = orig^ exp(D dt) + integral(D, R^(x), R^prev)
      F(x)   = Re IFFT(x^ - update) / N;
* nonlinearIntegral (implicitEuler.go:151-162), the linear-interpolation exponential
  integral of the explicit terms over the step;
* Step (implicitEuler.go:165-208): orig^ and R^prev at the current fields, an explicit
  Euler step as the initial guess, then a Newton-Krylov solve of F(x) = 0.

The reference's nonlinear solver is gononlin v0.2.2's NewtonKrylov (go.mod:9) with a
gonum GMRES inner solver, neither vendored.  The step's result is the root of F, which
does not depend on the solver beyond its tolerance, so this restatement solves F = 0 with
scipy's Newton-Krylov (the Knoll-Keyes Jacobian-free method) to a residual far below
the reference's Tol = 1e-7 (implicitEuler.go:221-228).
"""
from __future__ import annotations

from typing import Dict, Sequence

import numpy as np
from scipy.optimize import newton_krylov

from . import parser as P
from .engine import OracleSolver


def nonlinear_integral(denum: np.ndarray, rhs: np.ndarray, rhs_prev: np.ndarray, dt: float) -> np.ndarray:
    """nonlinearIntegral (implicitEuler.go:151-162), elementwise on complex arrays."""
    a = rhs_prev
    b = (rhs - rhs_prev) / dt
    f = np.exp(denum * dt)
    small = np.abs(denum) < 1e-5
    safe = np.where(small, 1.0, denum)
    out = a * (f - 1.0) / safe + b * (f - safe * dt - 1.0) / (safe * safe)
    return np.where(small, 0.5 * dt * (rhs + rhs_prev * f), out)


class ImplicitEulerOracle:
    """pf.ImplicitEuler stepping a pf.Model on the C oracle's bricks."""

    def __init__(self, model: P.Model, dims: Sequence[int], fields: Dict[str, np.ndarray], dt: float,
                 f_tol: float = 1e-12, filter_order: int = 0) -> None:
        # the explicit Euler step that makes the initial guess carries the stepper's filter
        self.o = OracleSolver(model, dims, fields, dt, stepper="euler", filter_order=filter_order)
        self.model = model
        self.dt = dt
        self.n = self.o.n
        self.nf = len(model.fields)
        self.f_tol = f_tol
        self.current_step = 0

    def field(self, name: str) -> np.ndarray:
        return self.o.field(name)

    def _transform_all(self, x: np.ndarray) -> None:
        """Fields <- x; SyncDerivedFields; forward FFT of every brick (implicitEuler.go:107-115)."""
        for i in range(self.nf):
            self.o.set_brick(i, x[i * self.n:(i + 1) * self.n].astype(complex))
        self.o.sync_derived()
        for b in range(self.nf + len(self.model.derived)):
            self.o.fft_brick(b)

    def _residual(self, x: np.ndarray, orig, rhs_prev) -> np.ndarray:
        self._transform_all(x)
        out = np.empty(self.nf * self.n)
        for i in range(self.nf):
            rhs = self.o.get_rhs(i)
            den = self.o.get_denum(i)
            update = orig[i] * np.exp(den * self.dt) + nonlinear_integral(den, rhs, rhs_prev[i], self.dt)
            res = self.o.brick(i) - update
            out[i * self.n:(i + 1) * self.n] = np.real(np.fft.ifftn(res.reshape(self.o.dims)).reshape(-1))
        return out

    def step(self) -> None:
        """ImplicitEuler.Step (implicitEuler.go:165-208)."""
        x_now = np.concatenate([self.o.brick(i).real for i in range(self.nf)])
        self._transform_all(x_now)
        orig = [self.o.brick(i).copy() for i in range(self.nf)]
        rhs_prev = [self.o.get_rhs(i) for i in range(self.nf)]
        # explicit Euler step from the current fields as the initial guess
        for i in range(self.nf):
            self.o.set_brick(i, x_now[i * self.n:(i + 1) * self.n].astype(complex))
        self.o.step()
        x0 = np.concatenate([self.o.brick(i).real for i in range(self.nf)])
        sol = newton_krylov(lambda x: self._residual(x, orig, rhs_prev), x0, f_tol=self.f_tol, method="lgmres")
        for i in range(self.nf):
            self.o.set_brick(i, sol[i * self.n:(i + 1) * self.n].astype(complex))
        self.current_step += 1

    def get_time(self) -> float:
        return self.dt * self.current_step
