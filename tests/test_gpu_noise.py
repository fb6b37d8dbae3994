"""Noise terms on the device against the oracle (the same Philox4x32-10 draws).

The reference draws from Go's global math/rand, which no run can reproduce; the device and the
oracle draw from the same counter-based generator keyed by the registered seed, so the device
is compared with the oracle value for value: max|gpu - oracle| / max|oracle| <= 1e-10 (the
draws themselves agree to a few ulps: log, cos and sqrt are libm's on the host).
"""
import numpy as np
import pytest

import gopf_amd as G
from oracle import parser as P
from oracle.engine import OracleSolver

pytestmark = pytest.mark.gpu


def maxrel(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


def kpz_models(N, seed):
    """examples/kardar_parisi_zhang: dheight/dt = LAP height + GRAD_SQ + WHITE_NOISE."""
    h0 = 0.01 * np.random.default_rng(1).random(N * N)
    pm, om = G.Model(), P.Model()
    pm.AddField("height", h0)
    om.add_field("height", N * N)
    pm.RegisterFunction("WHITE_NOISE", G.WhiteNoise(0.05, Seed=seed))
    om.register_white_noise("WHITE_NOISE", 0.05, seed)
    pm.RegisterExplicitTerm("GRAD_SQ", G.SquaredGradient("height", 1.0))
    om.register_explicit_term("GRAD_SQ", P.KIND_SQGRAD, field="height", factor=1.0)
    pm.AddEquation("dheight/dt = LAP height + GRAD_SQ + WHITE_NOISE")
    om.add_equation("dheight/dt = LAP height + GRAD_SQ + WHITE_NOISE")
    pm.Init()
    om.init()
    return pm, om, h0


@pytest.mark.parametrize("stepper", ["euler", "rk4"])
def test_white_noise_kpz(stepper):
    N, dt, steps = 64, 0.01, 20
    pm, om, h0 = kpz_models(N, 2024)
    s = G.Solver(pm, [N, N], dt, stepper=stepper)
    assert not s.fused
    o = OracleSolver(om, [N, N], {"height": h0.astype(complex)}, dt, stepper=stepper)
    s.Propagate(steps)
    o.solver_propagate(steps)
    assert maxrel(s.Field("height").real, o.field("height").real) <= 1e-10


def test_white_noise_seeds_and_statistics():
    """Another seed gives another path; one noisy Euler step of dx/dt = NOISE has the noise's
    variance 2 Strength dt^2 (Euler.Step: x += dt NOISE)."""
    N, dt = 512, 0.1
    out = {}
    for seed in (1, 2):
        m = G.Model()
        m.AddField("x", np.zeros(N * N))
        m.RegisterFunction("NOISE", G.WhiteNoise(2.0, Seed=seed))
        m.AddEquation("dx/dt = NOISE")
        m.Init()
        s = G.Solver(m, [N, N], dt)
        s.Propagate(1)
        out[seed] = s.Field("x").real
    x = out[1] / dt
    assert abs(x.std(ddof=1) - 2.0) <= 0.01 and abs(x.mean()) <= 0.02
    assert abs(np.corrcoef(out[1], out[2])[0, 1]) <= 0.01


def test_conservative_noise_parity_and_conservation():
    """TestConservativeNoise's model (noise_test.go:50-95) on the device: matches the oracle,
    stays real and keeps its integral at zero."""
    N, dt = 16, 0.1
    pm, om = G.Model(), P.Model()
    pm.AddField("myfield", np.zeros(N * N))
    om.add_field("myfield", N * N)
    pm.RegisterExplicitTerm("CONSERVATIVE_NOISE", G.NewConservativeNoise(1.0, 2, seed=99))
    om.register_conservative_noise("CONSERVATIVE_NOISE", 1.0, 2, 99)
    pm.AddEquation("dmyfield/dt = CONSERVATIVE_NOISE")
    om.add_equation("dmyfield/dt = CONSERVATIVE_NOISE")
    pm.Init()
    om.init()
    s = G.Solver(pm, [N, N], dt)
    o = OracleSolver(om, [N, N], {"myfield": np.zeros(N * N, dtype=complex)}, dt)
    s.Propagate(1000)
    o.solver_propagate(1000)
    f = s.Field("myfield").real
    assert maxrel(f, o.field("myfield").real) <= 1e-10
    assert abs(f.sum()) <= 1e-10 * np.abs(f).max() * N * N
    assert np.sum(np.abs(f) > np.sqrt(2.0)) > 0


def test_noise_source_is_not_fused():
    """A noise source is not a polynomial of the field: the single-field fused plan must not
    take it (it would drop the noise)."""
    N = 32
    m = G.Model()
    m.AddField("conc", np.zeros(N * N))
    m.RegisterFunction("NOISE", G.WhiteNoise(1.0, Seed=3))
    m.AddEquation("dconc/dt = LAP conc + NOISE")
    m.Init()
    s = G.Solver(m, [N, N], 0.1)
    assert not s.fused
    s.Propagate(1)
    assert np.abs(s.Field("conc").real).max() > 0.0
