"""The persistent 2D step (every step of a run in one cooperative launch, the passes separated
by device-wide barriers) against the per-pass launches and the oracle.

Bars: the persistent and per-pass paths compute each line with the same arithmetic up to the
row transform's schedule, so FP64 fields agree within 1e-12 max relative and FP32-bulk within
1e-6 relative L2; energy-series values within 1e-12 (FP64) and 1e-8 (FP32-bulk) relative (the
per-block partial sums group the voxels differently).  Against the oracle the FP64 bar of the solver tests holds: 1e-10.
"""
import numpy as np
import pytest

import gopf_amd as G
from oracle import parser as P

pytestmark = pytest.mark.gpu

CH_SCALARS = {"m2": 2.0, "m6": 6.0, "m4": 4.0, "mk": 1.0}
CH_EQ = "dconc/dt = m2*LAP conc - m6*LAP conc^2 + m4*LAP conc^3 - mk*LAP^2 conc"


def ch_model(c0):
    m = G.Model()
    m.AddField("conc", c0)
    for k, v in CH_SCALARS.items():
        m.AddScalar(k, v)
    m.AddEquation(CH_EQ)
    m.Init()
    return m


def ch_init(n, seed):
    rng = np.random.default_rng(seed)
    return 0.5 + 0.1 * (rng.random(n) - 0.5)


def run(c0, dims, precision, mode, threads, chunks, monitor_after=5):
    s = G.Solver(ch_model(c0), dims, 0.05, precision=precision)
    s.SetPersistent(mode, threads)
    s.StepperPropagate(monitor_after)
    s.MonitorEnergy(G.CH_BULK_POLY, 1.0)
    for k in chunks:
        s.StepperPropagate(k)
    e, m = s.EnergySeries()
    return s.Field(0).real, e, m, s.PersistentSteps(), s.GetTime()


@pytest.mark.parametrize("side,threads", [(32, 0), (64, 0), (64, 256), (128, 0), (128, 128), (256, 0),
                                          (256, 512)])
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_persistent_matches_per_pass(side, threads, precision):
    dims = [side, side]
    c0 = ch_init(side * side, 11)
    print("persist vs per-pass", side, threads, precision, flush=True)
    got = run(c0, dims, precision, 1, threads, [17, 2])
    ref = run(c0, dims, precision, 0, 0, [17, 2])
    assert got[3] == 5 + 17 + 2 and ref[3] == 0
    if precision == "fp64":
        assert np.abs(got[0] - ref[0]).max() <= 1e-12 * np.abs(ref[0]).max()
    else:
        assert np.linalg.norm(got[0] - ref[0]) / np.linalg.norm(ref[0]) <= 1e-6
    # series: FP32-bulk fields differ by FP32 rounding between the two row schedules
    tol = 1e-12 if precision == "fp64" else 1e-8
    assert got[1].size == ref[1].size == 19
    assert np.abs(got[1] - ref[1]).max() <= tol * np.abs(ref[1]).max()
    assert np.abs(got[2] - ref[2]).max() <= tol * np.abs(ref[2]).max()
    assert got[4] == ref[4]


@pytest.mark.parametrize("side,precision", [(128, "fp64"), (64, "fp32")])
def test_agent_fenced_barrier_matches_one_xcd_barrier(side, precision):
    """The barrier layout taken when the workgroups are not on one XCD (agent-scope fences,
    every block of the grid a participant) gives the one-XCD layout's results bit for bit."""
    dims = [side, side]
    c0 = ch_init(side * side, 16)
    a = run(c0, dims, precision, 2, 0, [17, 2])
    b = run(c0, dims, precision, 1, 0, [17, 2])
    assert a[3] == b[3] == 24
    assert np.array_equal(a[0], b[0])
    assert np.array_equal(a[1], b[1]) and np.array_equal(a[2], b[2])


def test_persistent_runs_split_across_calls_equal_one_run():
    """Launch boundaries move nothing: 7 + 13 + 3 steps equal 23 steps bit for bit, series
    included (the slot of each launch's last step is written after it)."""
    dims = [128, 128]
    c0 = ch_init(128 * 128, 12)
    a = run(c0, dims, "fp64", 1, 0, [7, 13, 3])
    b = run(c0, dims, "fp64", 1, 0, [23])
    assert np.array_equal(a[0], b[0])
    assert a[1].size == 23 and np.array_equal(a[1], b[1]) and np.array_equal(a[2], b[2])
    assert a[3] == b[3] == 28


@pytest.mark.parametrize("side,steps", [(32, 40), (256, 20)])
def test_persistent_fp64_matches_oracle(oracle_lib, side, steps):
    dims = [side, side]
    n = side * side
    c0 = ch_init(n, 13)
    om = P.Model()
    om.add_field("conc", n)
    for k, v in CH_SCALARS.items():
        om.add_scalar(k, v)
    om.add_equation(CH_EQ)
    om.init()
    print("persist oracle", side, steps, flush=True)
    s = G.Solver(ch_model(c0), dims, 0.05)
    s.SetPersistent(1)
    s.StepperPropagate(steps)
    assert s.PersistentSteps() == steps
    from oracle.engine import OracleSolver
    o = OracleSolver(om, dims, {"conc": c0}, 0.05)
    o.stepper_propagate(steps)
    ref = o.field("conc").real
    assert np.abs(s.Field(0).real - ref).max() <= 1e-10 * np.abs(ref).max()


def plain_run(eq, scalars, c0, dims, precision, mode, steps):
    m = G.Model()
    m.AddField("conc", c0)
    for k, v in scalars.items():
        m.AddScalar(k, v)
    m.AddEquation(eq)
    m.Init()
    s = G.Solver(m, dims, 0.05, precision=precision)
    s.SetPersistent(mode)
    s.StepperPropagate(steps)
    return s, s.Field(0).real


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_persistent_without_source_matches_per_pass(precision):
    """Pure diffusion has no explicit source and no monitor: the launch runs update passes
    only, one barrier per step.  Mass is conserved."""
    dims = [64, 64]
    c0 = ch_init(64 * 64, 17)
    s1, a = plain_run("dconc/dt = LAP conc", {}, c0, dims, precision, 1, 30)
    _, b = plain_run("dconc/dt = LAP conc", {}, c0, dims, precision, 0, 30)
    assert s1.fused and s1.PersistentSteps() == 30
    tol = 1e-12 if precision == "fp64" else 1e-6
    assert np.abs(a - b).max() <= tol * np.abs(b).max()
    assert abs(a.sum() - c0.sum()) <= 1e-9 * c0.sum()


@pytest.mark.parametrize("side", [64, 128])
def test_persistent_moving_mean_matches_per_pass(side):
    """A non-conserved model (an implicit linear term and an explicit cubic source without
    a Laplacian): the k = 0 mode moves every step through the tabulated m1(0), m2(0)."""
    eq = "dconc/dt = a*conc - b*conc^3 + LAP conc"
    sc = {"a": 0.5, "b": 1.0}
    dims = [side, side]
    c0 = ch_init(side * side, 18)
    s1, a = plain_run(eq, sc, c0, dims, "fp64", 1, 40)
    _, b = plain_run(eq, sc, c0, dims, "fp64", 0, 40)
    assert s1.fused and s1.PersistentSteps() == 40
    assert np.abs(a - b).max() <= 1e-12 * np.abs(b).max()
    assert abs(a.mean() - c0.mean()) > 1e-3


def test_persistent_settings_are_checked():
    c0 = ch_init(64 * 64, 14)
    s = G.Solver(ch_model(c0), [64, 64], 0.05)
    with pytest.raises(G.GpfError):
        s.SetPersistent(3)
    with pytest.raises(G.GpfError):
        s.SetPersistent(1, 96)
    # a grid without the persistent step (not square) runs the per-pass launches
    c1 = ch_init(64 * 128, 15)
    r = G.Solver(ch_model(c1), [64, 128], 0.05)
    r.SetPersistent(1)
    r.StepperPropagate(12)
    assert r.PersistentSteps() == 0
