"""The reference's stepper / term known-answer tests, restated against the oracle.

Sources: pf/euler_test.go, pf/rk4_test.go, pf/solver_test.go,
pf/squareGradientTerm_test.go, pf/homoLinElast_test.go, pf/vandeven_test.go,
pf/util_test.go (TestApplyModalFilter).
"""
import math

import numpy as np
import pytest

from oracle import parser as P


def _solver(model, dims, fields, dt, stepper="euler", filter_order=0):
    from oracle.engine import OracleSolver
    return OracleSolver(model, dims, fields, dt, stepper, filter_order)


def _decay_model(eq, n):
    m = P.Model()
    m.add_field("field", n)
    m.add_scalar("rate", -1.0)
    m.add_equation(eq)
    m.init()
    return m


def test_exponential_decay(oracle_lib):
    """TestExponentialDecay (euler_test.go:10-49)."""
    N = 8
    m = _decay_model("dfield/dt = rate*field", N * N)
    s = _solver(m, [N, N], {"field": np.ones(N * N)}, 0.001)
    s.stepper_propagate(1000)
    f = s.field("field")
    assert np.abs(f.real - math.exp(-1.0)).max() < 1e-3 and np.abs(f.imag).max() < 1e-3
    assert abs(s.get_time() - 1.0) < 1e-10


def test_euler_square_decay(oracle_lib):
    """TestEulerSquareDecary (euler_test.go:51-85)."""
    N = 8
    m = _decay_model("dfield/dt = rate*field^2", N * N)
    s = _solver(m, [N, N], {"field": np.ones(N * N)}, 0.001)
    s.stepper_propagate(1000)
    f = s.field("field")
    assert np.abs(f.real - 0.5).max() < 1e-3 and np.abs(f.imag).max() < 1e-3


def test_rk4_simple_model(oracle_lib):
    """TestSimpleModel (rk4_test.go:14-59)."""
    N = 8
    m = _decay_model("dfield/dt = rate*field^2", N * N)
    s = _solver(m, [N, N], {"field": np.ones(N * N)}, 0.1, "rk4")
    s.stepper_propagate(10)
    f = s.field("field")
    assert np.abs(f.real - 1.0 / 2.0).max() < 1e-6 and np.abs(f.imag).max() < 1e-6
    assert abs(s.get_time() - 1.0) < 1e-10


def test_rk4_with_implicit(oracle_lib):
    """TestWithImplicit (rk4_test.go:61-100): logistic growth."""
    N = 8
    m = _decay_model("dfield/dt = field + rate*field^2", N * N)
    s = _solver(m, [N, N], {"field": 0.5 * np.ones(N * N)}, 0.01, "rk4")
    s.stepper_propagate(100)
    t = 1.0
    A = 1.0 / 0.5 - 1.0
    expect = math.exp(t) / (A + math.exp(t))
    f = s.field("field")
    assert np.abs(f.real - expect).max() < 1e-3


def test_solver_diffusion(oracle_lib):
    """TestSolverDiffusion (solver_test.go:9-34)."""
    m = P.Model()
    m.add_field("conc", 256)
    m.add_equation("dconc/dt = LAP conc")
    m.init()
    c0 = np.zeros(256)
    c0[128] = 1.0
    s = _solver(m, [16, 16], {"conc": c0}, 0.1)
    for _ in range(10):
        s.solver_propagate(10)
    c = s.field("conc").real
    assert abs(c.sum() - 1.0) < 1e-4
    assert (c < 1.0).all() and (c >= 0.0).all()
    # one diffusion step is exactly the heat-kernel multiplier 1/(1 + dt (2 pi |f|)^2) per step
    k = np.fft.fftfreq(16)
    f2 = k[:, None] ** 2 + k[None, :] ** 2
    mult = (1.0 / (1.0 + 0.1 * 4 * np.pi ** 2 * f2)) ** 100
    ref = np.fft.ifft2(np.fft.fft2(c0.reshape(16, 16)) * mult).real.reshape(-1)
    assert np.abs(c - ref).max() < 1e-14


def sin_x(nx, ny):
    c = np.arange(nx * ny) % nx
    x = c / nx
    return np.sin(2 * np.pi * x), (2 * np.pi * np.cos(2 * np.pi * x) / nx) ** 2


def poly4(nx, ny):
    i = np.arange(nx * ny)
    r, c = i // nx, i % nx
    x, y = c / nx, r / ny
    vx = 16.0 * (x * x - 2 * x ** 3 + x ** 4)
    vy = 16.0 * (y * y - 2 * y ** 3 + y ** 4)
    ddx = 16.0 * (2 * x - 6 * x * x + 4 * x ** 3) * vy / nx
    ddy = vx * 16.0 * (2 * y - 6 * y * y + 4 * y ** 3) / ny
    return vx * vy, ddx * ddx + ddy * ddy


@pytest.mark.parametrize("fn,N,tol", [(sin_x, 16, 1e-10), (poly4, 32, 1e-5)])
def test_square_gradient(oracle_lib, fn, N, tol):
    """TestSquareGradient (squareGradientTerm_test.go:42-82)."""
    data, grad_sq = fn(N, N)
    m = P.Model()
    m.add_field("height", N * N)
    m.register_explicit_term("GRAD", P.KIND_SQGRAD, field="height")
    m.add_equation("dheight/dt = GRAD")
    m.init()
    s = _solver(m, [N, N], {"height": data}, 0.1)
    s.fft_brick(0)
    res = s.eval_user_term(P.KIND_SQGRAD, 0)
    res = oracle_lib.fft(res, [N, N], inverse=True) / (N * N)
    assert np.abs(res.real - grad_sq).max() < tol and np.abs(res.imag).max() < tol


def test_square_grad_with_solver(oracle_lib):
    """TestSquareGradWithSolver (squareGradientTerm_test.go:84-128)."""
    N = 16
    data, grad_sq = sin_x(N, N)
    m = P.Model()
    m.add_field("field1", N * N)
    m.add_field("field2", N * N)
    m.add_scalar("ZERO", 0.0)
    m.register_explicit_term("GRAD_SQ_f2", P.KIND_SQGRAD, field="field2")
    m.add_equation("dfield1/dt = GRAD_SQ_f2")
    m.add_equation("dfield2/dt = ZERO*field1")
    m.init()
    s = _solver(m, [N, N], {"field2": data}, 0.1)
    s.solver_propagate(10)
    f1, f2 = s.field("field1"), s.field("field2")
    assert np.abs(f2.real - data).max() < 1e-10 and np.abs(f2.imag).max() < 1e-10
    assert np.abs(f1.real - 0.1 * 10 * grad_sq).max() < 1e-10 and np.abs(f1.imag).max() < 1e-10


def indicator(x):
    return 3.0 * x * x - 2.0 * x ** 3


def indicator_deriv(x):
    return 6.0 * x - 6.0 * x * x


def test_indicator_deriv():
    """TestIndicatorDeriv (homoLinElast_test.go:12-25)."""
    dx = 0.01
    for i in range(100):
        x = dx * i
        d = (indicator(x + dx / 2) - indicator(x - dx / 2)) / dx
        assert abs(d - indicator_deriv(x)) < 1e-4


def _elastic_model(N, dims, Cijkl, misfit):
    m = P.Model()
    m.add_field("x", N)
    m.register_explicit_term("ELAST", P.KIND_ELAST, field="x", C=Cijkl, misfit=misfit)
    m.add_equation("dx/dt = ELAST")
    m.init()
    return m


def test_homogeneous_rhs(oracle_lib):
    """TestHomogeneousRHS (homoLinElast_test.go:27-88): injected displacement u_x = (x(1-x))^2."""
    N = 16
    Cijkl = oracle_lib.isotropic(60.0, 0.3)
    eps = 0.01
    misfit = np.zeros((3, 3))
    misfit[0, 0] = eps
    m = _elastic_model(N * N, [N, N], Cijkl, misfit)
    s = _solver(m, [N, N], {}, 0.1)
    row = np.arange(N * N) // N
    x = row / N
    ux = oracle_lib.fft((x * (1 - x)) ** 2, [N, N])
    disp = np.zeros((3, N * N), complex)
    disp[0] = ux
    s.inject_displacement(0, disp)
    eta = 0.5
    s.set_lag(0, np.full(N * N, eta))
    res = s.eval_user_term(P.KIND_ELAST, 0)
    res = oracle_lib.fft(res, [N, N], inverse=True) / (N * N)
    exx = 2 * x * (1 - x) * (1 - 2 * x) / N
    C0 = Cijkl[0]
    expect = C0 * eps * (exx - eps * indicator(eta)) * indicator_deriv(eta)
    assert np.abs(res.real - expect).max() < 1e-4 and np.abs(res.imag).max() < 1e-4


@pytest.mark.parametrize("dims", [[16, 16], [8, 8, 8]])
def test_elastic_isotropic_dilatational_closed_form(oracle_lib, dims):
    """Full Khachaturyan term (force, 3x3 solve, strains) against the isotropic closed form.

    For C isotropic (bulk B, shear mu) and misfit eps*I, u(k) = F(k) / ((2 pi)^2 (lam + 2 mu) |f|^2),
    so sum_ij fac sigma*_ij eps_ij(k) = K H(k) with K = 9 B^2 eps^2 / (lam + 2 mu) for k != 0 and no
    Nyquist component.  Term = H' (K (H - mean H) - 9 B eps^2 H).
    """
    n = int(np.prod(dims))
    B, nu = 60.0, 0.3
    mu = 3.0 * B * (1 - 2 * nu) / (2 * (1 + nu))
    lam = B - 2 * mu / 3
    eps = 0.02
    Cijkl = oracle_lib.isotropic(B, nu)
    m = _elastic_model(n, dims, Cijkl, eps * np.eye(3))
    s = _solver(m, dims, {}, 0.1)
    idx = np.arange(n)
    col = idx % dims[1]
    row = (idx // dims[1]) % dims[0]
    phi = 0.5 + 0.2 * np.sin(2 * np.pi * row / dims[0]) + 0.15 * np.cos(2 * np.pi * col / dims[1])
    s.set_lag(0, phi)
    res = s.eval_user_term(P.KIND_ELAST, 0)
    res = oracle_lib.fft(res, dims, inverse=True) / n
    H = indicator(phi)
    if len(dims) == 2:
        # 2D: only Dim=2 force components; the 3x3 solve stays block-diagonal for isotropic C
        K = 9 * B * B * eps * eps / (lam + 2 * mu) * 1.0
        # sigma* sums over i<Dim only: sum_{i<2} sigma_ii eps_ii = 3 B eps * (eps_00 + eps_11)
    else:
        K = 9 * B * B * eps * eps / (lam + 2 * mu)
    expect = indicator_deriv(phi) * (K * (H - H.mean()) - 9 * B * eps * eps * H)
    assert np.abs(res.real - expect).max() < 1e-9 * np.abs(expect).max()


@pytest.mark.parametrize("order", [3, 5, 10])
def test_vandeven(oracle_lib, order):
    """TestVandeven (vandeven_test.go:8-39)."""
    tab = oracle_lib.vandeven_table(order)
    for x, y in ((0.0, 1.0), (0.5, 0.5), (1.0, 0.0)):
        assert abs(oracle_lib.vandeven_eval(tab, x) - y) < 1e-4


@pytest.mark.parametrize("order", [2, 4, 8])
def test_vandeven_is_incomplete_beta(oracle_lib, order):
    """sigma(x) = 1 - I_x(p, p) (regularised incomplete beta), up to the trapezoid error."""
    from scipy.special import betainc
    tab = oracle_lib.vandeven_table(order)
    for x in np.linspace(0.0, 0.95, 17):
        assert abs(oracle_lib.vandeven_eval(tab, x) - (1.0 - betainc(order, order, x))) < 2e-5


def test_modal_filter_applied_with_two_over_pi(oracle_lib):
    """ApplyModalFilter (util.go:125-132): d(k) *= sigma(|f| * 2 / pi) after the update."""
    N = 16
    m = P.Model()
    m.add_field("conc", N * N)
    m.add_scalar("ZERO", 0.0)
    m.add_equation("dconc/dt = ZERO*conc")
    m.init()
    rng = np.random.default_rng(3)
    c0 = rng.standard_normal(N * N)
    s = _solver(m, [N, N], {"conc": c0}, 0.1, "euler", filter_order=4)
    s.step()
    from scipy.special import betainc
    k = np.fft.fftfreq(N)
    k = np.where(k == -0.5, 0.5, k)
    frad = np.sqrt(k[:, None] ** 2 + k[None, :] ** 2) * 2 / np.pi
    sigma = 1.0 - betainc(4, 4, np.minimum(frad, 1.0))
    ref = np.fft.ifft2(np.fft.fft2(c0.reshape(N, N)) * sigma).real.reshape(-1)
    assert np.abs(s.field("conc").real - ref).max() < 1e-4


def _ch_model(n, M=1.0, kappa=1.0):
    m = P.Model()
    m.add_field("conc", n)
    for s, v in (("m2", 2 * M), ("m6", 6 * M), ("m4", 4 * M), ("mk", M * kappa)):
        m.add_scalar(s, v)
    m.add_equation("dconc/dt = m2*LAP conc - m6*LAP conc^2 + m4*LAP conc^3 - mk*LAP^2 conc")
    m.init()
    return m


def test_ch_energy_single_mode(oracle_lib):
    """F = sum f(c) + kappa/2 sum |grad c|^2 against the single-mode closed form."""
    N, a, kk, kappa = 32, 0.1, 3, 1.7
    i = np.arange(N * N)
    col = i % N
    c = 0.5 + a * np.cos(2 * np.pi * kk * col / N)
    F = oracle_lib.ch_energy([N, N], c, [0.0, 0.0, 1.0, -2.0, 1.0], kappa)
    # bulk: f(0.5 + a cos) = (1/4 - a^2 cos^2)^2 averaged: 1/16 - a^2/4 + 3 a^4/8 per voxel
    bulk = N * N * (1.0 / 16 - a * a / 4 + 3 * a ** 4 / 8)
    grad = 0.5 * kappa * N * N * a * a * (2 * np.pi * kk / N) ** 2 / 2
    assert abs(F - (bulk + grad)) < 1e-12 * (bulk + grad)


def test_ch_mass_and_energy_decay(oracle_lib):
    N = 32
    m = _ch_model(N * N)
    rng = np.random.default_rng(0)
    c0 = 0.5 + 0.1 * (rng.random(N * N) - 0.5)
    s = _solver(m, [N, N], {"conc": c0}, 0.1)
    energies = []
    for _ in range(20):
        energies.append(oracle_lib.ch_energy([N, N], s.field("conc").real, [0, 0, 1, -2, 1], 1.0))
        s.stepper_propagate(10)
    c = s.field("conc").real
    assert abs(c.sum() - c0.sum()) < 1e-11 * c0.sum()
    assert all(b <= a + 1e-12 for a, b in zip(energies, energies[1:]))
