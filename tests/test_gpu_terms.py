"""SpectralViscosity and VolumeConservingLP on the device against the oracle.

Bar: FP64 max|gpu - oracle| / max|oracle| <= 1e-10 (Euler, RK4) and <= 1e-9 (ImplicitEuler,
Newton tolerance 1e-12 on both sides; the strain example's elastic term at 1e-9).
"""
import numpy as np
import pytest

import gopf_amd as G
from gopf_amd import workloads as W
from oracle import parser as P
from oracle.engine import OracleSolver
from oracle.implicit import ImplicitEulerOracle

pytestmark = pytest.mark.gpu


def maxrel(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


def sv_models(N, eq, eps=0.5, thr=0.3, power=2):
    c0 = 0.3 * (np.random.default_rng(7).random(N * N) - 0.5)
    pm, om = G.Model(), P.Model()
    pm.AddField("conc", c0)
    om.add_field("conc", N * N)
    pm.RegisterImplicitTerm("SVISC", G.SpectralViscosity(eps, thr, power))
    om.register_implicit_term("SVISC", P.KIND_SVISC, eps=eps, threshold=thr, power=power)
    pm.AddEquation(eq)
    om.add_equation(eq)
    pm.Init()
    om.init()
    return pm, om, c0


@pytest.mark.parametrize("eq", ["dconc/dt = LAP conc^3 - LAP conc - LAP^2 conc + SVISC",
                                "dconc/dt = LAP conc^3 - LAP conc - LAP^2 conc - LAP SVISC",
                                "dconc/dt = conc - conc^3 + SVISC"],
                         ids=["plain", "lap_neg", "allen_cahn"])
@pytest.mark.parametrize("stepper", ["euler", "rk4"])
def test_spectral_viscosity_parity(eq, stepper):
    N, dt, steps = 32, 0.05, 20
    pm, om, c0 = sv_models(N, eq)
    s = G.Solver(pm, [N, N], dt, stepper=stepper)
    assert not s.fused
    o = OracleSolver(om, [N, N], {"conc": c0.astype(complex)}, dt, stepper=stepper)
    s.Propagate(steps)
    o.solver_propagate(steps)
    assert maxrel(s.Field("conc").real, o.field("conc").real) <= 1e-10


def test_spectral_viscosity_fp32():
    N, dt, steps = 32, 0.05, 20
    pm, om, c0 = sv_models(N, "dconc/dt = LAP conc^3 - LAP conc - LAP^2 conc + SVISC")
    s = G.Solver(pm, [N, N], dt, precision="fp32")
    o = OracleSolver(om, [N, N], {"conc": c0.astype(complex)}, dt)
    s.Propagate(steps)
    o.solver_propagate(steps)
    a, b = s.Field("conc").real, o.field("conc").real
    assert np.linalg.norm(a - b) / np.linalg.norm(b) <= 1e-5


def test_spectral_viscosity_implicit_euler():
    N, dt = 32, 0.05
    pm, om, c0 = sv_models(N, "dconc/dt = LAP conc^3 - LAP conc - LAP^2 conc + SVISC", eps=2.0, thr=0.2, power=4)
    s = G.Solver(pm, [N, N], dt, stepper="implicitEuler")
    s.SetNewtonKrylov(Tol=1e-12)
    o = ImplicitEulerOracle(om, [N, N], {"conc": c0.astype(complex)}, dt, f_tol=1e-12)
    for _ in range(3):
        s.Propagate(1)
        o.step()
        assert s.NewtonStats()["converged"]
    assert maxrel(s.Field("conc").real, o.field("conc").real) <= 1e-9


def strain_models(N, dt, conserve=True):
    """examples/strain_single_precipitate/main.go: conc and phase, the example's registered
    functions as polynomials, the 2D Khachaturyan term, and the volume constraint."""
    n = N * N
    idx = np.arange(n)
    r, c = idx // N, idx % N
    inside = ((r > 3 * N // 8) & (r < 5 * N // 8) & (c > 3 * N // 8) & (c < 5 * N // 8)).astype(float)
    rng = np.random.default_rng(11)
    conc0 = inside + 0.01 * (rng.random(n) - 0.5)
    phase0 = inside + 0.01 * (rng.random(n) - 0.5)
    polys = dict(W.precipitate_polynomials())
    polys["SMEARING_DERIV"] = [(6.0, {"conc": 0.0, "phase": 1.0}), (-6.0, {"conc": 0.0, "phase": 2.0})]
    Cm = G.CubicMaterial(110.0, 60.0, 30.0)
    misfit = np.array([[0.05, 0, 0], [0, -0.01, 0], [0, 0, 0]])
    pm, om = G.Model(), P.Model()
    for name, data in (("conc", conc0), ("phase", phase0)):
        pm.AddField(name, data)
        om.add_field(name, n)
    pm.AddScalar("kappa", 0.1)
    om.add_scalar("kappa", 0.1)
    for name, poly in polys.items():
        pm.RegisterPolynomial(name, poly)
        om.register_polynomial(name, poly)
    pm.RegisterExplicitTerm("LIN_ELAST", G.HomogeneousModulusLinElast("phase", Cm, misfit))
    om.register_explicit_term("LIN_ELAST", P.KIND_ELAST, field="phase", C=Cm, misfit=misfit)
    eqs = ["dconc/dt = CHEMICALPOT + kappa*LAP conc",
           "dphase/dt = DERIV_PHASE_ORDER + LIN_ELAST + kappa*LAP phase" + (" + CONSERVE_PREC_VOL" if conserve else "")]
    if conserve:
        pm.RegisterExplicitTerm("CONSERVE_PREC_VOL", G.NewVolumeConservingLP("phase", "SMEARING_DERIV", dt, n))
        om.register_explicit_term("CONSERVE_PREC_VOL", P.KIND_VOLCONS, field="phase", indicator="SMEARING_DERIV",
                                  dt=dt, num_nodes=n)
    for e in eqs:
        pm.AddEquation(e)
        om.add_equation(e)
    pm.Init()
    om.init()
    return pm, om, {"conc": conc0.astype(complex), "phase": phase0.astype(complex)}


@pytest.mark.parametrize("stepper", ["euler", "rk4"])
def test_volume_conserving_strain_example(stepper):
    N, dt, steps = 64, 0.1, 30
    pm, om, init = strain_models(N, dt)
    s = G.Solver(pm, [N, N], dt, stepper=stepper)
    o = OracleSolver(om, [N, N], init, dt, stepper=stepper)
    s.Propagate(steps)
    o.solver_propagate(steps)
    for name in ("conc", "phase"):
        assert maxrel(s.Field(name).real, o.field(name).real) <= 1e-9, name


def test_volume_conserving_holds_the_volume():
    """The constraint keeps sum(phase) far closer to its value after the first steps than the
    unconstrained run does."""
    N, dt = 64, 0.1
    drift = {}
    for conserve in (True, False):
        pm, _, _ = strain_models(N, dt, conserve)
        s = G.Solver(pm, [N, N], dt)
        s.Propagate(2)
        v0 = s.Field("phase").real.sum()
        s.Propagate(100)
        drift[conserve] = abs(s.Field("phase").real.sum() - v0)
    assert drift[True] < 0.1 * drift[False], drift


def test_volume_conserving_stepper_propagate_skips_the_update():
    """Euler.Propagate does not run OnStepFinished (euler.go:55-59): the multiplier stays 0 and
    the run equals the unconstrained one."""
    N, dt = 32, 0.1
    pm, _, _ = strain_models(N, dt, True)
    pu, _, _ = strain_models(N, dt, False)
    a = G.Solver(pm, [N, N], dt)
    b = G.Solver(pu, [N, N], dt)
    a.StepperPropagate(10)
    b.StepperPropagate(10)
    assert np.array_equal(a.Field("phase").real, b.Field("phase").real)


HESS_CASES = {
    "2d_diag": ([32, 32], [1.0, 0.0, 0.0, 0.5]),
    "3d_diag": ([16, 16, 16], [1.0, 0.0, 0.0, 0.0, 0.7, 0.0, 0.0, 0.0, 0.4]),
    "2d_offdiag": ([32, 32], [1.0, 0.3, 0.3, 0.5]),
    "3d_offdiag": ([16, 16, 16], [1.0, 0.2, 0.0, 0.2, 0.7, 0.1, 0.0, 0.1, 0.4]),
}


def hess_models(dims, K):
    n = int(np.prod(dims))
    c0 = 0.3 * (np.random.default_rng(9).random(n) - 0.5)
    eq = "dconc/dt = HESSIAN + LAP conc^3 - LAP^2 conc"
    pm, om = G.Model(), P.Model()
    pm.AddField("conc", c0)
    om.add_field("conc", n)
    pm.RegisterImplicitTerm("HESSIAN", G.TensorialHessian("conc", K))
    om.register_implicit_term("HESSIAN", P.KIND_HESS, K=K)
    pm.AddEquation(eq)
    om.add_equation(eq)
    pm.Init()
    om.init()
    return pm, om, c0


@pytest.mark.parametrize("case", list(HESS_CASES))
@pytest.mark.parametrize("stepper", ["euler", "rk4"])
def test_tensorial_hessian_parity(case, stepper):
    """Anisotropic diffusion through an implicit TensorialHessian term.

    With a diagonal K the multiplier is Hermitian and the device matches the oracle to 1e-10.
    An off-diagonal K_jk multiplies by f_j f_k, which the reference's Freq (+0.5 for the
    Nyquist index of k and of -k) makes non-Hermitian on Nyquist rows: the reference's fields
    pick up imaginary parts (about 1e-6 of the field here) that feed back into the real part
    (about 1e-8 after 10 steps).  The device keeps real fields, so it is compared with the
    oracle projected onto real fields after every step (1e-10 Euler, 1e-9 RK4, whose stages
    keep the within-step coupling) and with the plain oracle at 1e-7."""
    dims, K = HESS_CASES[case]
    pm, om, c0 = hess_models(dims, K)
    dt, steps = 0.02, 10
    s = G.Solver(pm, dims, dt, stepper=stepper)
    plain = OracleSolver(om, dims, {"conc": c0.astype(complex)}, dt, stepper=stepper)
    real = OracleSolver(om, dims, {"conc": c0.astype(complex)}, dt, stepper=stepper)
    s.Propagate(steps)
    plain.solver_propagate(steps)
    for _ in range(steps):
        real.solver_propagate(1)
        real.set_field("conc", real.field("conc").real.astype(complex))
    got = s.Field("conc").real
    if "diag" in case and "off" not in case:
        assert maxrel(got, plain.field("conc").real) <= 1e-10
    else:
        assert maxrel(got, real.field("conc").real) <= (1e-10 if stepper == "euler" else 1e-9)
        assert maxrel(got, plain.field("conc").real) <= 1e-7


@pytest.mark.parametrize("case", ["2d_offdiag", "3d_offdiag"])
def test_tensorial_hessian_implicit_euler(case):
    """ImplicitEuler's unknowns are real vectors (implicitEuler.go:48-66), so its residual
    sees the Hermitian part of the multiplier: the device matches the oracle to 1e-9."""
    dims, K = HESS_CASES[case]
    pm, om, c0 = hess_models(dims, K)
    dt = 0.02
    s = G.Solver(pm, dims, dt, stepper="implicitEuler")
    s.SetNewtonKrylov(Tol=1e-12)
    o = ImplicitEulerOracle(om, dims, {"conc": c0.astype(complex)}, dt, f_tol=1e-12)
    for _ in range(3):
        s.Propagate(1)
        o.step()
    assert maxrel(s.Field("conc").real, o.field("conc").real) <= 1e-9


def test_tensorial_hessian_checks_k():
    m = G.Model()
    m.AddField("conc", np.zeros(16 ** 3))
    m.RegisterImplicitTerm("HESSIAN", G.TensorialHessian("conc", [1.0, 0.0, 0.0, 1.0]))
    m.AddEquation("dconc/dt = HESSIAN")
    m.Init()
    with pytest.raises(G.GpfError):
        G.Solver(m, [16, 16, 16], 0.1)   # a 2 x 2 K on a 3D grid
