"""ImplicitEuler on the device against the reference's known answers and the oracle.

* Known answers (pf/implicitEuler_test.go): ODEs on 8x8 within 0.005, the dissipating
  heat equation on 128x128 within 1e-3.
* Oracle parity: with the Newton tolerance at 1e-12 (max norm of the residual) on both
  sides, max|gpu - oracle| / max|oracle| <= 1e-9 after the stated steps.  The step's result
  is the root of the residual equation, which does not depend on the Krylov method.
"""
import math

import numpy as np
import pytest

import gopf_amd as G
from oracle import parser as P
from oracle.implicit import ImplicitEulerOracle

pytestmark = pytest.mark.gpu

TIGHT = 1e-12


def pair_models(fields, scalars, eqs, n, polys=None):
    pm, om = G.Model(), P.Model()
    for name, data in fields.items():
        pm.AddField(name, data)
        om.add_field(name, n)
    for s, v in scalars.items():
        pm.AddScalar(s, v)
        om.add_scalar(s, v)
    for name, poly in (polys or {}).items():
        pm.RegisterPolynomial(name, poly)
        om.register_polynomial(name, poly)
    for e in eqs:
        pm.AddEquation(e)
        om.add_equation(e)
    pm.Init()
    om.init()
    return pm, om


def maxrel(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


ODE_CASES = [
    (["conc"], ["dconc/dt = -conc"], [1.0], lambda t: [math.exp(-t)]),
    (["conc"], ["dconc/dt = -conc^2"], [1.0], lambda t: [1.0 / (1.0 + t)]),
    (["conc"], ["dconc/dt = conc - conc^2"], [0.5], lambda t: [math.exp(t) / (1.0 + math.exp(t))]),
    (["conc1", "conc2"], ["dconc1/dt = -conc1*conc2", "dconc2/dt = -conc2"], [1.0, 1.0],
     lambda t: [math.exp(math.exp(-t) - 1.0), math.exp(-t)]),
]


@pytest.mark.parametrize("names,eqs,init,solution", ODE_CASES, ids=["linear", "nonlinear", "logistic", "coupled"])
def test_implicit_euler_known_answers(names, eqs, init, solution):
    """TestImplicitEuler (implicitEuler_test.go:10-110): 8x8, dt 0.01, 100 steps, tol 0.005."""
    N = 8
    m = G.Model()
    for nm, v in zip(names, init):
        m.AddField(nm, np.full(N * N, v))
    for e in eqs:
        m.AddEquation(e)
    m.Init()
    s = G.Solver(m, [N, N], 0.01, stepper="implicitEuler")
    assert not s.fused
    s.Propagate(100)
    assert abs(s.GetTime() - 1.0) < 1e-12
    st = s.NewtonStats()
    assert st["converged"] and st["residual"] < 1e-7
    expect = solution(1.0)
    for j, nm in enumerate(names):
        f = s.Field(nm)
        assert np.all(np.abs(f.real - expect[j]) <= 0.005), (nm, f.real[:4], expect[j])


def test_dissipating_heat_equation():
    """TestDissipatingHeatEquation (implicitEuler_test.go:158-229): 128x128, dt 0.005, 10 steps, tol 1e-3."""
    N, gamma, dt, nsteps = 128, 0.2, 0.005, 10
    idx = np.arange(N * N)
    x, y = (idx // N) / N, (idx % N) / N
    m = G.Model()
    m.AddField("temperature", np.sin(2.0 * x * np.pi) * np.sin(2.0 * y * np.pi))
    m.RegisterPolynomial("DISSIPATE", [(-gamma, {"temperature": 1.0})])
    m.AddEquation("dtemperature/dt = LAP temperature + DISSIPATE")
    m.Init()
    s = G.Solver(m, [N, N], dt, stepper="implicitEuler")
    s.Propagate(nsteps)
    t = nsteps * dt
    L = float(N)
    value = np.exp(-(4.0 * np.pi / (L * L) + gamma) * t) * np.sin(2.0 * np.pi * y) * np.sin(2.0 * np.pi * x)
    assert np.abs(s.Field("temperature").real - value).max() <= 1e-3


def _parity(fields, scalars, eqs, dims, dt, nsteps, polys=None, stencil=6, filter_order=0):
    n = int(np.prod(dims))
    pm, om = pair_models(fields, scalars, eqs, n, polys)
    s = G.Solver(pm, dims, dt, stepper="implicitEuler", filter_order=filter_order)
    s.SetNewtonKrylov(Tol=TIGHT, Stencil=stencil)
    o = ImplicitEulerOracle(om, dims, {k: np.asarray(v, dtype=complex) for k, v in fields.items()}, dt,
                            f_tol=TIGHT, filter_order=filter_order)
    for _ in range(nsteps):
        s.Propagate(1)
        o.step()
        st = s.NewtonStats()
        assert st["converged"], st
    worst = 0.0
    for name in fields:
        worst = max(worst, maxrel(s.Field(name).real, o.field(name).real))
    return worst, s


def test_parity_cahn_hilliard_2d():
    """Conserved dynamics with a stiff implicit LAP^2 part and an explicit LAP conc^3."""
    N = 32
    rng = np.random.default_rng(3)
    c0 = 0.1 * (rng.random(N * N) - 0.5)
    worst, _ = _parity({"conc": c0}, {"gamma": 0.5}, ["dconc/dt = LAP conc^3 - LAP conc - gamma*LAP^2 conc"],
                       [N, N], 0.1, 3)
    assert worst <= 1e-9, worst


def test_parity_cahn_hilliard_3d():
    N = 8
    rng = np.random.default_rng(4)
    c0 = 0.2 * (rng.random(N ** 3) - 0.5)
    worst, _ = _parity({"conc": c0}, {"gamma": 0.2}, ["dconc/dt = LAP conc^3 - LAP conc - gamma*LAP^2 conc"],
                       [N, N, N], 0.05, 2)
    assert worst <= 1e-9, worst


def test_parity_coupled_fields():
    """Two fields: a product term in a derived field and a cross-field linear term."""
    N = 16
    rng = np.random.default_rng(5)
    a = 1.0 + 0.2 * rng.random(N * N)
    b = 1.0 + 0.2 * rng.random(N * N)
    worst, _ = _parity({"conc1": a, "conc2": b}, {},
                       ["dconc1/dt = -conc1*conc2 + LAP conc1", "dconc2/dt = -conc2 + conc1"], [N, N], 0.05, 3)
    assert worst <= 1e-9, worst


def test_parity_allen_cahn_filtered():
    """The solverStability example's Allen-Cahn (CHEMPOT = c - c^3, gamma 0.02) at dt 1, with a
    Vandeven filter on the explicit Euler initial guess."""
    N = 32
    rng = np.random.default_rng(0)
    c0 = 2.0 * rng.random(N * N) - 1.0
    worst, _ = _parity({"conc": c0}, {"gamma": 0.02}, ["dconc/dt = CHEMPOT + gamma*LAP conc"], [N, N], 1.0, 3,
                       polys={"CHEMPOT": [(1.0, {"conc": 1.0}), (-1.0, {"conc": 3.0})]}, filter_order=4)
    assert worst <= 1e-9, worst


@pytest.mark.parametrize("stencil", [1, 2, 4, 8])
def test_stencils_reach_the_same_root(stencil):
    N = 32
    rng = np.random.default_rng(3)
    c0 = 0.1 * (rng.random(N * N) - 0.5)
    worst, _ = _parity({"conc": c0}, {"gamma": 0.5}, ["dconc/dt = LAP conc^3 - LAP conc - gamma*LAP^2 conc"],
                       [N, N], 0.1, 2, stencil=stencil)
    assert worst <= 1e-9, worst


def _allen_cahn(dt, stepper):
    N = 32
    rng = np.random.default_rng(0)
    m = G.Model()
    m.AddField("conc", 2.0 * rng.random(N * N) - 1.0)
    m.AddScalar("gamma", 0.02)
    m.RegisterPolynomial("CHEMPOT", [(1.0, {"conc": 1.0}), (-1.0, {"conc": 3.0})])
    m.AddEquation("dconc/dt = CHEMPOT + gamma*LAP conc")
    m.Init()
    return G.Solver(m, [N, N], dt, stepper=stepper)


def test_stability_large_step():
    """examples/solverStability (description.md:11-17, the stable = no NaN criterion of
    main.go:56-66): at dt = 3.4 Euler goes non-finite within 100 steps while ImplicitEuler
    stays finite; at dt = 2.1 every ImplicitEuler step converges.  (Euler's blow-up at
    dt 1.5-2.1 depends on the random initial field; at dt = 3.4 plain Newton from the
    explicit Euler guess, like the oracle's SciPy Newton-Krylov, does not reach Tol.)"""
    s = _allen_cahn(2.1, "implicitEuler")
    for _ in range(10):
        s.Propagate(1)
        assert s.NewtonStats()["converged"]
    c = s.Field("conc").real
    assert np.all(np.isfinite(c)) and np.abs(c).max() <= 1.5
    s = _allen_cahn(3.4, "implicitEuler")
    s.Propagate(2)
    assert np.all(np.isfinite(s.Field("conc").real))
    e = _allen_cahn(3.4, "euler")
    e.Propagate(100)
    assert not np.all(np.isfinite(e.Field("conc").real))


def test_fp32_is_rejected():
    m = G.Model()
    m.AddField("conc", np.zeros(64))
    m.AddEquation("dconc/dt = -conc")
    m.Init()
    with pytest.raises(G.GpfError):
        G.Solver(m, [8, 8], 0.01, stepper="implicitEuler", precision="fp32")


def test_newton_krylov_settings_are_checked():
    s = _allen_cahn(0.1, "implicitEuler")
    for bad in (dict(Stencil=3), dict(Maxiter=0), dict(Tol=0.0), dict(StepSize=-1.0), dict(Restart=65)):
        with pytest.raises(G.GpfError):
            s.SetNewtonKrylov(**bad)
    e = _allen_cahn(0.1, "euler")
    with pytest.raises(G.GpfError):
        e.SetNewtonKrylov()
    with pytest.raises(G.GpfError):
        e.NewtonStats()
