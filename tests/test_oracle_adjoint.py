"""The oracle's FP64 reverse mode of Euler.Step against central finite differences.

The adjoint has no reference counterpart; the oracle's term-by-term transposed step
(or_euler_adjoint) is pinned here by central differences (h = 1e-6) of the oracle's
own forward solver, for the field gradient along seeded unit directions and for every
scalar of the model, on three classes: Cahn-Hilliard, Cahn-Hilliard with a Vandeven
filter, and a model with a field-linear explicit term.  Bar: 1e-6 relative error.
"""
import numpy as np
import pytest

from gopf_amd import workloads as W
from oracle import parser as P
from oracle.engine import OracleSolver

H = 1e-6
TOL = 1e-6


def build(eq, scal, n):
    m = P.Model()
    m.add_field("conc", n)
    for k, v in scal.items():
        m.add_scalar(k, v)
    m.add_equation(eq)
    m.init()
    return m


def loss(eq, scal, dims, c0, tgt, dt, steps, filt):
    o = OracleSolver(build(eq, scal, int(np.prod(dims))), dims, {"conc": c0}, dt, filter_order=filt)
    o.stepper_propagate(steps)
    return float(np.sum((o.field("conc").real - tgt) ** 2))


CASES = [
    ("ch", W.CH_EQUATION, W.ch_scalars(1.0, 1.0), [16, 16], 0.05, 40, 0),
    ("ch_filter", W.CH_EQUATION, W.ch_scalars(0.7, 1.2), [16, 16], 0.05, 30, 4),
    ("linear_explicit", "dconc/dt = mg*conc*conc + mk*LAP conc", {"mg": 0.2, "mk": 0.1}, [8, 16], 0.05, 25, 0),
    ("ch_3d", W.CH_EQUATION, W.ch_scalars(1.0, 1.0), [8, 8, 8], 0.05, 15, 0),
]


@pytest.mark.parametrize("name,eq,scal,dims,dt,steps,filt", CASES, ids=[c[0] for c in CASES])
def test_oracle_adjoint_matches_central_differences(name, eq, scal, dims, dt, steps, filt):
    n = int(np.prod(dims))
    c0 = W.ch_initial(5, n)
    t = OracleSolver(build(eq, scal, n), dims, {"conc": W.ch_initial(6, n)}, dt, filter_order=filt)
    t.stepper_propagate(steps // 2)
    tgt = t.field("conc").real
    o = OracleSolver(build(eq, scal, n), dims, {"conc": c0}, dt, filter_order=filt)
    J, g0, gs = o.adjoint(steps, tgt, list(scal))
    assert J == pytest.approx(loss(eq, scal, dims, c0, tgt, dt, steps, filt), rel=1e-14)
    # the solver is left at c_T
    ref = OracleSolver(build(eq, scal, n), dims, {"conc": c0}, dt, filter_order=filt)
    ref.stepper_propagate(steps)
    assert np.array_equal(o.field("conc").real, ref.field("conc").real)
    rng = np.random.default_rng(11)
    for _ in range(2):
        e = rng.standard_normal(n)
        e /= np.linalg.norm(e)
        fd = (loss(eq, scal, dims, c0 + H * e, tgt, dt, steps, filt)
              - loss(eq, scal, dims, c0 - H * e, tgt, dt, steps, filt)) / (2 * H)
        assert abs(fd - g0 @ e) <= TOL * abs(fd), (fd, g0 @ e)
    for k in scal:
        sp, sm = dict(scal), dict(scal)
        sp[k] += H
        sm[k] -= H
        fd = (loss(eq, sp, dims, c0, tgt, dt, steps, filt) - loss(eq, sm, dims, c0, tgt, dt, steps, filt)) / (2 * H)
        assert abs(fd - gs[k]) <= TOL * abs(fd), (k, fd, gs[k])


def test_central_differences_catch_a_dropped_nonlinear_term():
    """The check is sharp enough to see a plausible mistake: a reverse mode that drops the
    explicit g(c) terms (the chemical-potential derivative) misses the true central
    difference by far more than the bar."""
    dims, n, dt, steps = [16, 16], 256, 0.05, 20
    scal = W.ch_scalars(1.0, 1.0)
    c0 = W.ch_initial(5, n)
    tgt = W.ch_initial(6, n)
    _, g_ok, _ = OracleSolver(build(W.CH_EQUATION, scal, n), dims, {"conc": c0}, dt).adjoint(steps, tgt)
    planted = build(W.CH_EQUATION, scal, n)
    planted.rhs[0] = ([], planted.rhs[0][1])
    _, g_bad, _ = OracleSolver(planted, dims, {"conc": c0}, dt).adjoint(steps, tgt)
    e = np.random.default_rng(3).standard_normal(n)
    e /= np.linalg.norm(e)
    fd = (loss(W.CH_EQUATION, scal, dims, c0 + H * e, tgt, dt, steps, 0)
          - loss(W.CH_EQUATION, scal, dims, c0 - H * e, tgt, dt, steps, 0)) / (2 * H)
    assert abs(fd - g_ok @ e) <= TOL * abs(fd)
    assert abs(fd - g_bad @ e) > 1e-3 * abs(fd)
