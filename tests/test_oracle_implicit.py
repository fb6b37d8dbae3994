"""The oracle's ImplicitEuler restatement against the reference's own known answers
(pf/implicitEuler_test.go: TestImplicitEuler, TestDissipatingHeatEquation)."""
import math

import numpy as np
import pytest

from oracle import parser as P
from oracle.implicit import ImplicitEulerOracle, nonlinear_integral

CASES = [
    (["conc"], ["dconc/dt = -conc"], [1.0], lambda t: [math.exp(-t)]),
    (["conc"], ["dconc/dt = -conc^2"], [1.0], lambda t: [1.0 / (1.0 + t)]),
    (["conc"], ["dconc/dt = conc - conc^2"], [0.5], lambda t: [math.exp(t) / (1.0 + math.exp(t))]),
    (["conc1", "conc2"], ["dconc1/dt = -conc1*conc2", "dconc2/dt = -conc2"], [1.0, 1.0],
     lambda t: [math.exp(math.exp(-t) - 1.0), math.exp(-t)]),
]


@pytest.mark.parametrize("names,eqs,init,solution", CASES, ids=["linear", "nonlinear", "logistic", "coupled"])
def test_implicit_euler_known_answers(names, eqs, init, solution):
    """TestImplicitEuler (implicitEuler_test.go:10-110): 8x8, dt 0.01, 100 steps, tol 0.005."""
    N = 8
    m = P.Model()
    for name in names:
        m.add_field(name, N * N)
    for e in eqs:
        m.add_equation(e)
    m.init()
    o = ImplicitEulerOracle(m, [N, N], {nm: np.full(N * N, v, dtype=complex) for nm, v in zip(names, init)}, 0.01)
    for _ in range(100):
        o.step()
    expect = solution(0.01 * 100)
    for j, nm in enumerate(names):
        f = o.field(nm)
        assert np.all(np.abs(f.real - expect[j]) <= 0.005) and np.all(np.abs(f.imag) <= 0.005)


def test_dissipating_heat_equation():
    """TestDissipatingHeatEquation (implicitEuler_test.go:158-229): 128x128, dt 0.005, 10 steps, tol 1e-3."""
    N, gamma, dt, nsteps = 128, 0.2, 0.005, 10
    idx = np.arange(N * N)
    ix, iy = idx // N, idx % N
    x, y = ix / N, iy / N
    field = np.sin(2.0 * x * np.pi) * np.sin(2.0 * y * np.pi)
    m = P.Model()
    m.add_field("temperature", N * N)
    m.register_polynomial("DISSIPATE", [(-gamma, {"temperature": 1.0})])
    m.add_equation("dtemperature/dt = LAP temperature + DISSIPATE")
    m.init()
    o = ImplicitEulerOracle(m, [N, N], {"temperature": field.astype(complex)}, dt)
    for _ in range(nsteps):
        o.step()
    L = float(N)
    t = nsteps * dt
    value = np.exp(-(4.0 * np.pi / (L * L) + gamma) * t) * np.sin(2.0 * np.pi * y) * np.sin(2.0 * np.pi * x)
    got = o.field("temperature")
    assert np.all(np.abs(got.real - value) <= 1e-3) and np.all(np.abs(got.imag) <= 1e-3)


def test_nonlinear_integral_limits():
    """nonlinearIntegral (implicitEuler.go:151-162): the |D| -> 0 branch is the trapezoid rule,
    and the general branch equals the exact integral of exp(D (dt - s)) (a + b s) over [0, dt]."""
    dt = 0.1
    rhs, prev = np.array([2.0 + 1j]), np.array([1.0 - 0.5j])
    small = nonlinear_integral(np.array([0.0]), rhs, prev, dt)
    assert np.allclose(small, 0.5 * dt * (rhs + prev))
    D = np.array([-3.0])
    s = np.linspace(0.0, dt, 20001)
    a, b = prev[0], (rhs[0] - prev[0]) / dt
    integrand = np.exp(D[0] * (dt - s)) * (a + b * s)
    exact = np.sum((integrand[1:] + integrand[:-1]) * 0.5 * np.diff(s))
    assert abs(nonlinear_integral(D, rhs, prev, dt)[0] - exact) <= 1e-9
