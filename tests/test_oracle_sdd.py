"""The oracle's SDD stepper against the reference's SDD tests (pf/sdd_test.go), restated."""
import math

import numpy as np
import pytest

from oracle import engine as E
from oracle import parser as P
from oracle.sdd import SDDOracle, householder, householder_denum, sherman_morrison_dot


@pytest.fixture(scope="module", autouse=True)
def _built():
    E.build()


def box_blur(x, dims, width=2):
    """pfutil.Blur with BoxKernel{Width} (pfutil/blur.go:15-67): the mean over the periodic
    (2 (width/2 + 1) + 1)^d box whose offsets are within width."""
    a = np.asarray(x, dtype=float).reshape(dims)
    cut = width // 2 + 1
    acc = np.zeros_like(a)
    cnt = 0
    for dy in range(-cut, cut + 1):
        for dx in range(-cut, cut + 1):
            if abs(dy) > width or abs(dx) > width:
                continue
            acc += np.roll(np.roll(a, -dy, axis=0), -dx, axis=1)
            cnt += 1
    return (acc / cnt).reshape(-1)


def test_dimer_length_time():
    """TestDimerLengthTime (sdd_test.go:59-66)."""
    m = P.Model()
    m.add_field("conc", 4)
    m.add_equation("dconc/dt = LAP conc")
    m.init()
    s = SDDOracle(m, [2, 2], {"conc": np.zeros(4, dtype=complex)}, 0.1)
    s.init_dimer_length = 5.0
    assert abs(s.dimer_length(s.required_dimer_length_time(0.5)) - 0.5) <= 1e-10


def test_diagonal_sherman_morrison():
    """TestDiagonalShermannMorrison (sdd_test.go:273-322): against a dense solve."""
    D = np.array([0.4, -0.2, 1.4])
    u = np.array([-1.0, 2.0, 3.0])
    v = np.array([2.0, 2.3, 1.2])
    b = np.array([4.0, 6.0, 8.2])
    A = np.diag(D) + np.outer(u, v)
    got = sherman_morrison_dot(1.0 / D + 0j, u + 0j, v + 0j, b + 0j)
    assert np.allclose(got.real, np.linalg.solve(A, b), atol=1e-10, rtol=0)


def test_householder_golden():
    """TestHouseHolder (sdd_test.go:409-451): the reference's numpy golden vector."""
    N = 4
    i = np.arange(N * N, dtype=float)
    vec = i.copy()
    orient = i * i - 3 * i
    orient /= np.sqrt(np.sum(orient ** 2))
    fv = np.fft.fft2(vec.reshape(N, N)).reshape(-1)
    fo = np.fft.fft2(orient.reshape(N, N)).reshape(-1)
    out = np.fft.ifft2(householder(fv, fo, 2.0, N * N).reshape(N, N)).reshape(-1).real
    expect = [0.0, 1.41446756, 2.41446756, 3.0, 3.17106489, 2.92766222, 2.26979199, 1.19745421,
              -0.28935113, -2.19062403, -4.50636448, -7.23657249, -10.38124806, -13.94039118, -17.91400186,
              -22.3020801]
    assert np.abs(out - expect).max() <= 1e-6


def test_householder_denum():
    """TestHouseHolderDenum (sdd_test.go:453-515): the Sherman-Morrison solve in k-space equals
    the dense real-space solve of ((1 - dt) I + dt sigma o o^T) x = b."""
    N, sigma, dt = 16, 1.0, 0.8
    i = np.arange(N * N, dtype=float)
    orient = 0.1 * (i * i - 4 * i)
    orient /= np.sqrt(np.sum(orient ** 2))
    x = 0.1 * i
    A = (1.0 - dt) * np.eye(N * N) + dt * sigma * np.outer(orient, orient)
    res = np.linalg.solve(A, x)
    fo = np.fft.fft2(orient.reshape(N, N)).reshape(-1)
    fx = np.fft.fft2(x.reshape(N, N)).reshape(-1)
    inv, u, v = householder_denum(np.ones(N * N, dtype=complex), fo, sigma, dt)
    out = np.fft.ifft2(sherman_morrison_dot(inv, u, v, fx).reshape(N, N)).reshape(-1).real
    assert np.abs(out - res).max() <= 1e-10


def test_double_well():
    """TestDoubleWell (sdd_test.go:68-108): the saddle of dconc/dt = conc - conc^3 between -1 and
    1.5 is 0."""
    N, dt = 4, 0.1
    m = P.Model()
    m.add_field("conc", N * N)
    m.add_equation("dconc/dt = conc - conc^3")
    m.init()
    init, final = np.full(N * N, -1.0), np.full(N * N, 1.5)
    s = SDDOracle(m, [N, N], {"conc": (0.5 * (init + final)).astype(complex)}, dt)
    s.init([init], [final])
    s.init_dimer_length = 0.1
    nsteps = int(s.required_dimer_length_time(1e-6 * s.dimer_length(0.0)) / dt) + 1
    for _ in range(nsteps):
        s.step()
    assert np.abs(s.field("conc")).max() <= 1e-6


def circle(N, radius):
    idx = np.arange(N * N)
    r, c = idx // N, idx % N
    out = np.full(N * N, -1.0)
    out[(r - N // 2) ** 2 + (c - N // 2) ** 2 <= radius * radius] = 1.0
    return out


@pytest.mark.slow
def test_classical_nucleation():
    """TestClassicalNucleation (sdd_test.go:165-271): the critical droplet's radius is
    2 sigma / rho within 0.3, sigma = sqrt(gamma / 2) 2/3."""
    N, gamma, rho, dt = 64, 0.5, 0.05, 0.7
    rc = 2.0 * math.sqrt(gamma / 2.0) * (2.0 / 3.0) / rho
    m = P.Model()
    m.add_field("phi", N * N)
    m.add_scalar("gamma", gamma)
    # (1 - phi^2)(phi + 3 rho / 4)
    a = 3.0 * rho / 4.0
    m.register_polynomial("MINUS_CHEM_POT", [(a, {}), (1.0, {"phi": 1.0}), (-a, {"phi": 2.0}), (-1.0, {"phi": 3.0})])
    m.add_equation("dphi/dt = MINUS_CHEM_POT + gamma*LAP phi")
    m.init()
    init = box_blur(circle(N, 10), [N, N])
    final = box_blur(circle(N, 15), [N, N])
    field = box_blur(circle(N, 12), [N, N])
    s = SDDOracle(m, [N, N], {"phi": field.astype(complex)}, dt)
    s.init_dimer_length = 1.0      # overwritten by Init, as in the reference
    s.min_dimer_length = 5e-6
    s.init([init], [final])
    for _ in range(int(1000.0 / dt)):
        s.step()
    phi = s.field("phi").real
    area = float(np.sum(0.5 * (1.0 + phi)))
    assert abs(math.sqrt(area / math.pi) - rc) <= 0.3
