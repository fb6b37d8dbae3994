"""The benchmarked workloads at their full sizes.

At these sizes the oracle affords a step or two, so each workload is checked against that
(plain FP64 path: max relative error <= 1e-10) and, over more steps, through properties that
hold at any size:
  * the FP32-bulk path against the FP64 path: relative L2 <= 1e-5;
  * the Cahn-Hilliard mass (the k = 0 mode) is conserved: drift <= 1e-12 relative in FP64;
  * the Cahn-Hilliard free energy does not increase (a gradient flow; the semi-implicit step
    is energy-stable at these time steps).
"""
import time

import numpy as np
import pytest

import gopf_amd as G
from gopf_amd import workloads as W
from oracle import parser as P

pytestmark = pytest.mark.gpu


def maxrel(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


def relL2(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def ch_models(c0, n):
    pm, om = G.Model(), P.Model()
    pm.AddField("conc", c0)
    om.add_field("conc", n)
    for k, v in W.ch_scalars(1.0, 1.0).items():
        pm.AddScalar(k, v)
        om.add_scalar(k, v)
    pm.AddEquation(W.CH_EQUATION)
    om.add_equation(W.CH_EQUATION)
    pm.Init()
    om.init()
    return pm, om


def test_ch3d_256_step_matches_oracle(oracle_lib):
    """256^3 Cahn-Hilliard, dt 0.05, the headline workload's initial field: one semi-implicit
    Euler step of the fused FP64 path against the oracle's step (euler.go:16-47)."""
    N, dt = 256, 0.05
    dims, n = [N] * 3, N ** 3
    c0 = W.ch_initial(1, n)
    pm, om = ch_models(c0, n)
    s = G.Solver(pm, dims, dt, precision="fp64")
    assert s.fused
    s.StepperPropagate(1)
    c_dev = s.Field(0).real
    del s
    from oracle.engine import OracleSolver
    t0 = time.perf_counter()
    o = OracleSolver(om, dims, {"conc": c0}, dt)
    o.stepper_propagate(1)
    c_or = o.field("conc").real
    err = maxrel(c_dev, c_or)
    print(f"256^3 one step: max rel {err:.3g} (oracle {time.perf_counter() - t0:.1f} s)", flush=True)
    assert err <= 1e-10
    # the step moved the field: the comparison is not of two initial fields
    assert maxrel(c_or, c0) > 1e-6


def test_ch3d_256_fp32_fp64_mass_energy():
    """256^3, 100 steps: FP32-bulk against FP64, the conserved mass, a non-increasing energy."""
    N, dt, steps = 256, 0.05, 100
    dims, n = [N] * 3, N ** 3
    c0 = W.ch_initial(1, n)
    out = {}
    for prec in ("fp64", "fp32"):
        pm, _ = ch_models(c0, n)
        s = G.Solver(pm, dims, dt, precision=prec)
        assert s.fused
        s.MonitorEnergy(W.CH_BULK_POLY, 1.0)
        s.StepperPropagate(steps)
        series, mass = s.EnergySeries()
        out[prec] = (s.Field(0).real, np.asarray(series), np.asarray(mass))
        del s
    c64, e64, m64 = out["fp64"]
    c32, e32, m32 = out["fp32"]
    err = relL2(c32, c64)
    drift64 = np.abs(m64 - c0.sum()).max() / c0.sum()
    drift32 = np.abs(m32 - c0.sum()).max() / c0.sum()
    print(f"256^3 {steps} steps: fp32 vs fp64 rel L2 {err:.3g}; mass drift fp64 {drift64:.3g} fp32 {drift32:.3g}; "
          f"energy {e64[0]:.6g} -> {e64[-1]:.6g}", flush=True)
    assert err <= 1e-5
    assert drift64 <= 1e-12
    assert drift32 <= 1e-9
    assert len(e64) == steps and len(e32) == steps
    assert np.all(np.diff(e64) <= 1e-12 * np.abs(e64[:-1]))
    assert e64[-1] < e64[0]
    assert np.abs(e32 - e64).max() <= 1e-5 * np.abs(e64).max()


def precipitate_models(conc0, phase0, n):
    Cm = G.CubicMaterial(110.0, 60.0, 30.0)
    misfit = 0.01 * np.eye(3)
    pm, om = G.Model(), P.Model()
    for name, data in (("conc", conc0), ("phase", phase0)):
        pm.AddField(name, data)
        om.add_field(name, n)
    for k, v in W.PRECIPITATE_SCALARS.items():
        pm.AddScalar(k, v)
        om.add_scalar(k, v)
    for name, poly in W.precipitate_polynomials().items():
        pm.RegisterPolynomial(name, poly)
        om.register_polynomial(name, poly)
    term = G.HomogeneousModulusLinElast("phase", Cm, misfit)
    pm.RegisterExplicitTerm("LIN_ELAST", term)
    om.register_explicit_term("LIN_ELAST", P.KIND_ELAST, field=term.FieldName, C=term.MatProp, misfit=term.Misfit)
    for e in W.PRECIPITATE_EQUATIONS:
        pm.AddEquation(e)
        om.add_equation(e)
    pm.Init()
    om.init()
    return pm, om


def test_precipitate_48_rk4_matches_oracle(oracle_lib):
    """The precipitate model at 48^3 (sphere radius 5), RK4, dt 0.01: two steps of the FP64
    general path against the oracle (rk4.go:29-127; the elastic field lags one step, so the
    second step exercises the lagged copy).  The oracle's elastic chain takes about 140 s per
    step at 192^3; 48^3 keeps it to seconds and has radix-3 transform lengths."""
    N, dt = 48, 0.01
    dims, n = [N] * 3, N ** 3
    conc0, phase0 = W.precipitate_initial(N, 5.0, seed=4)
    pm, om = precipitate_models(conc0, phase0, n)
    s = G.Solver(pm, dims, dt, stepper="rk4", precision="fp64")
    assert not s.fused
    s.Propagate(2)
    from oracle.engine import OracleSolver
    o = OracleSolver(om, dims, {"conc": conc0, "phase": phase0}, dt, "rk4")
    o.solver_propagate(2)
    for f in ("conc", "phase"):
        err = maxrel(s.Field(f).real, o.field(f).real)
        print(f"48^3 RK4 two steps, {f}: max rel {err:.3g}", flush=True)
        assert err <= 1e-10


def test_precipitate_192_fp32_fp64_and_cubic_symmetry():
    """192^3 precipitate, RK4, dt 0.01, 20 steps, from a noise-free centred sphere (radius 20)
    in a cubic material: the FP32-bulk path against FP64, and the FP64 fields invariant under
    the cube's symmetries (axis permutations and the reflections i -> N - i), which a
    frequency or stiffness index mixed up between axes would break."""
    N, dt, steps = 192, 0.01, 20
    dims, n = [N] * 3, N ** 3
    idx = np.arange(n)
    col, row, dep = idx % N, (idx // N) % N, idx // (N * N)
    r2 = (col - N / 2) ** 2 + (row - N / 2) ** 2 + (dep - N / 2) ** 2
    sphere = np.where(r2 < 400.0, 1.0, 0.0)
    out = {}
    for prec in ("fp64", "fp32"):
        pm, _ = precipitate_models(sphere, sphere, n)
        s = G.Solver(pm, dims, dt, stepper="rk4", precision=prec)
        s.Propagate(steps)
        out[prec] = {f: s.Field(f).real for f in ("conc", "phase")}
        del s
    for f in ("conc", "phase"):
        err = relL2(out["fp32"][f], out["fp64"][f])
        c = out["fp64"][f].reshape(N, N, N)
        moved = maxrel(c.reshape(-1), sphere)
        sym = 0.0
        for img in (c.transpose(1, 0, 2), c.transpose(2, 1, 0), c.transpose(0, 2, 1),
                    np.roll(c[::-1], 1, axis=0), np.roll(c[:, ::-1], 1, axis=1), np.roll(c[:, :, ::-1], 1, axis=2)):
            sym = max(sym, maxrel(img.reshape(-1), c.reshape(-1)))
        print(f"192^3 RK4 {steps} steps, {f}: fp32 vs fp64 rel L2 {err:.3g}; symmetry {sym:.3g}; moved {moved:.3g}",
              flush=True)
        assert err <= 1e-5
        assert sym <= 1e-10
        assert moved > 1e-4


def test_batch_384_sims_full_size():
    """The batched sweep at its full size (128 sims of 128^2, 192 of 256^2, 64 of 512^2,
    FP32-bulk, dt 0.1, 50 steps): every sim conserves its mass, and the first and last sim of
    each size class agree with a single FP64 solver of that sim (relative L2 <= 1e-5).

    The state's k = 0 mode is the FP64 mean, which the Cahn-Hilliard step leaves unchanged;
    the read-back field is an FP32 inverse transform plus that mean, and the rounding of its
    up to 512^2 values puts about 1e-10 to 1e-9 on their sum, hence the 1e-8 bound."""
    dt, steps = 0.1, 50
    m = G.Model()
    m.AddField("conc", n=16)
    for k, v in W.ch_scalars().items():
        m.AddScalar(k, v)
    m.AddEquation(W.CH_EQUATION)
    m.Init()
    b = G.Batch(m, W.C3_CLASSES, dt, precision="fp32")
    assert len(b) == 384
    inputs = []
    for s in range(len(b)):
        M, kappa, c0 = W.c3_sim_inputs(s, int(np.prod(b.shapes[s])))
        for k, v in W.ch_scalars(M, kappa).items():
            b.SetScalar(s, k, v)
        b.SetField(s, c0)
        inputs.append((M, kappa, c0.sum()))
    b.Propagate(steps)
    worst = 0.0
    for s in range(len(b)):
        worst = max(worst, abs(b.Field(s).real.sum() - inputs[s][2]) / inputs[s][2])
    print(f"384 sims, {steps} steps: worst mass drift {worst:.3g}", flush=True)
    assert worst <= 1e-8
    for s in (0, 127, 128, 319, 320, 383):
        dims = list(b.shapes[s])
        M, kappa, c0 = W.c3_sim_inputs(s, int(np.prod(dims)))
        sm = G.Model()
        sm.AddField("conc", c0)
        for k, v in W.ch_scalars(M, kappa).items():
            sm.AddScalar(k, v)
        sm.AddEquation(W.CH_EQUATION)
        sm.Init()
        single = G.Solver(sm, dims, dt, precision="fp64")
        single.StepperPropagate(steps)
        err = relL2(b.Field(s).real, single.Field(0).real)
        print(f"sim {s} {dims}: batch fp32 vs single fp64 rel L2 {err:.3g}", flush=True)
        assert err <= 1e-5
        del single
