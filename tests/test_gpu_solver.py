"""The HIP stepper against the oracle and the reference's known-answer tests.

Parity bar (max relative error = max|gpu - oracle| / max|oracle|):
  * plain FP64 path: <= 1e-10 after the stated steps
  * FP32-bulk path:  relative L2 <= 1e-5 (short runs), mass drift <= 1e-9
"""
import math

import numpy as np
import pytest

import gopf_amd as G
from gopf_amd import workloads as W
from oracle import parser as P

pytestmark = pytest.mark.gpu


def splitmix64_uniform(seed, n):
    """SplitMix64(seed) -> u = (x >> 11) * 2^-53 in [0, 1)."""
    out = np.empty(n)
    s = np.uint64(seed)
    with np.errstate(over="ignore"):
        for i in range(n):
            s = np.uint64(s + np.uint64(0x9E3779B97F4A7C15))
            z = s
            z = np.uint64((z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9))
            z = np.uint64((z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB))
            z = z ^ (z >> np.uint64(31))
            out[i] = float(z >> np.uint64(11)) * (1.0 / 9007199254740992.0)
    return out


def ch_init(n, seed):
    rng = np.random.default_rng(seed)
    return 0.5 + 0.1 * (rng.random(n) - 0.5)


def maxrel(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


def relL2(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def pair_models(fields, scalars, eqs, n, explicit=None, polys=None):
    pm, om = G.Model(), P.Model()
    for name, data in fields.items():
        pm.AddField(name, data)
        om.add_field(name, n)
    for s, v in scalars.items():
        pm.AddScalar(s, v)
        om.add_scalar(s, v)
    for name, (poly) in (polys or {}).items():
        pm.RegisterPolynomial(name, poly)
        om.register_polynomial(name, poly)
    for name, term in (explicit or {}).items():
        pm.RegisterExplicitTerm(name, term)
        if isinstance(term, G.SquaredGradient):
            om.register_explicit_term(name, P.KIND_SQGRAD, field=term.Field, factor=term.Factor)
        else:
            om.register_explicit_term(name, P.KIND_ELAST, field=term.FieldName, C=term.MatProp, misfit=term.Misfit)
    for e in eqs:
        pm.AddEquation(e)
        om.add_equation(e)
    pm.Init()
    om.init()
    return pm, om


# ---------------------------------------------------------------------------
# the reference's known-answer tests, on the device
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_exponential_decay(precision):
    """TestExponentialDecay (euler_test.go:10-49)."""
    N = 8
    m = G.Model()
    m.AddField("field", np.ones(N * N))
    m.AddScalar("rate", -1.0)
    m.AddEquation("dfield/dt = rate*field")
    m.Init()
    s = G.Solver(m, [N, N], 0.001, precision=precision)
    s.StepperPropagate(1000)
    f = s.Field(0)
    assert np.abs(f.real - math.exp(-1.0)).max() < 1e-3
    assert abs(s.GetTime() - 1.0) < 1e-10


def test_euler_square_decay():
    """TestEulerSquareDecary (euler_test.go:51-85)."""
    N = 8
    m = G.Model()
    m.AddField("field", np.ones(N * N))
    m.AddScalar("rate", -1.0)
    m.AddEquation("dfield/dt = rate*field^2")
    m.Init()
    s = G.Solver(m, [N, N], 0.001)
    assert s.fused
    s.StepperPropagate(1000)
    assert np.abs(s.Field(0).real - 0.5).max() < 1e-3


def test_rk4_simple_model():
    """TestSimpleModel (rk4_test.go:14-59)."""
    N = 8
    m = G.Model()
    m.AddField("field", np.ones(N * N))
    m.AddScalar("rate", -1.0)
    m.AddEquation("dfield/dt = rate*field^2")
    m.Init()
    s = G.Solver(m, [N, N], 0.1, stepper="rk4")
    assert not s.fused
    s.StepperPropagate(10)
    assert np.abs(s.Field(0).real - 0.5).max() < 1e-6
    assert abs(s.GetTime() - 1.0) < 1e-10


def test_rk4_with_implicit():
    """TestWithImplicit (rk4_test.go:61-100)."""
    N = 8
    m = G.Model()
    m.AddField("field", 0.5 * np.ones(N * N))
    m.AddScalar("rate", -1.0)
    m.AddEquation("dfield/dt = field + rate*field^2")
    m.Init()
    s = G.Solver(m, [N, N], 0.01, stepper="rk4")
    s.StepperPropagate(100)
    expect = math.exp(1.0) / (1.0 + math.exp(1.0))
    assert np.abs(s.Field(0).real - expect).max() < 1e-3


def test_solver_diffusion():
    """TestSolverDiffusion (solver_test.go:9-34)."""
    m = G.Model()
    c0 = np.zeros(256)
    c0[128] = 1.0
    m.AddField("conc", c0)
    m.AddEquation("dconc/dt = LAP conc")
    m.Init()
    s = G.Solver(m, [16, 16], 0.1)
    s.Solve(10, 10)
    c = s.Field(0).real
    assert abs(c.sum() - 1.0) < 1e-4
    assert (c < 1.0).all() and (c >= 0.0).all()


def test_square_grad_with_solver():
    """TestSquareGradWithSolver (squareGradientTerm_test.go:84-128)."""
    N = 16
    col = np.arange(N * N) % N
    x = col / N
    data = np.sin(2 * np.pi * x)
    grad_sq = (2 * np.pi * np.cos(2 * np.pi * x) / N) ** 2
    m = G.Model()
    m.AddField("field1", n=N * N)
    m.AddField("field2", data)
    m.AddScalar("ZERO", 0.0)
    m.RegisterExplicitTerm("GRAD_SQ_f2", G.SquaredGradient("field2"))
    m.AddEquation("dfield1/dt = GRAD_SQ_f2")
    m.AddEquation("dfield2/dt = ZERO*field1")
    m.Init()
    s = G.Solver(m, [N, N], 0.1)
    s.Solve(1, 10)
    assert np.abs(s.Field("field2").real - data).max() < 1e-10
    assert np.abs(s.Field("field1").real - 0.1 * 10 * grad_sq).max() < 1e-10


# ---------------------------------------------------------------------------
# parity with the oracle
# ---------------------------------------------------------------------------

CH_SCALARS = {"m2": 2.0, "m6": 6.0, "m4": 4.0, "mk": 1.0}
CH_EQ = "dconc/dt = m2*LAP conc - m6*LAP conc^2 + m4*LAP conc^3 - mk*LAP^2 conc"


@pytest.mark.parametrize("dims,steps,fused", [([64, 64], 200, True), ([64, 64], 200, False),
                                              ([32, 48], 100, True), ([16, 16, 16], 50, True),
                                              ([16, 16, 16], 50, False)])
def test_ch_fp64_parity(oracle_lib, dims, steps, fused):
    n = int(np.prod(dims))
    c0 = ch_init(n, 0)
    pm, om = pair_models({"conc": c0}, CH_SCALARS, [CH_EQ], n)
    print("ch fp64", dims, steps, fused, flush=True)
    s = G.Solver(pm, dims, 0.1, fused=fused)
    assert s.fused == fused
    s.StepperPropagate(steps)
    from oracle.engine import OracleSolver
    o = OracleSolver(om, dims, {"conc": c0}, 0.1)
    o.stepper_propagate(steps)
    assert maxrel(s.Field(0).real, o.field("conc").real) <= 1e-10


def test_ch_config1_128_1000_steps(oracle_lib):
    """2D CH 128^2, dt 0.1, 1000 semi-implicit Euler steps, FP64: final field and free-energy series."""
    dims = [128, 128]
    n = 128 * 128
    c0 = 0.5 + 0.1 * (splitmix64_uniform(0, n) - 0.5)
    pm, om = pair_models({"conc": c0}, CH_SCALARS, [CH_EQ], n)
    s = G.Solver(pm, dims, 0.1)
    assert s.fused
    s.MonitorEnergy(G.CH_BULK_POLY, 1.0)
    s.StepperPropagate(1000)
    from oracle.engine import OracleSolver
    o = OracleSolver(om, dims, {"conc": c0}, 0.1)
    energies = []
    for k in range(1000):
        if k % 100 == 0:
            energies.append((k, oracle_lib.ch_energy(dims, o.field("conc").real, G.CH_BULK_POLY, 1.0)))
        o.stepper_propagate(1)
    err = maxrel(s.Field(0).real, o.field("conc").real)
    print("C1 max rel err", err, flush=True)
    assert err <= 1e-10
    series, mass = s.EnergySeries()
    assert len(series) == 1000
    for k, e in energies:
        assert abs(series[k] - e) <= 1e-10 * abs(e)
    assert np.abs(mass - c0.sum()).max() <= 1e-9 * c0.sum()
    assert all(b <= a + 1e-12 for a, b in zip(series, series[1:]))


def test_ch_fp32_bulk_accuracy(oracle_lib):
    dims = [64, 64]
    n = 64 * 64
    c0 = ch_init(n, 3)
    pm, om = pair_models({"conc": c0}, CH_SCALARS, [CH_EQ], n)
    s = G.Solver(pm, dims, 0.1, precision="fp32")
    assert s.fused
    s.StepperPropagate(100)
    from oracle.engine import OracleSolver
    o = OracleSolver(om, dims, {"conc": c0}, 0.1)
    o.stepper_propagate(100)
    c = s.Field(0).real
    assert relL2(c, o.field("conc").real) <= 1e-5
    _, mass = s.Energy(0, G.CH_BULK_POLY, 1.0)
    assert abs(mass - c0.sum()) <= 1e-9 * c0.sum()


def test_vandeven_filter_parity(oracle_lib):
    dims = [32, 32]
    n = 32 * 32
    c0 = ch_init(n, 7)
    for fused in (True, False):
        pm, om = pair_models({"conc": c0}, CH_SCALARS, [CH_EQ], n)
        s = G.Solver(pm, dims, 0.1, filter_order=4, fused=fused)
        s.StepperPropagate(20)
        from oracle.engine import OracleSolver
        o = OracleSolver(om, dims, {"conc": c0}, 0.1, filter_order=4)
        o.stepper_propagate(20)
        assert maxrel(s.Field(0).real, o.field("conc").real) <= 1e-10


def test_two_field_rk4_parity(oracle_lib):
    """Two coupled fields, derived products, RK4 (rk4.go) against the oracle."""
    dims = [32, 32]
    n = 32 * 32
    rng = np.random.default_rng(11)
    a0 = 0.5 + 0.05 * rng.standard_normal(n)
    b0 = 0.3 + 0.05 * rng.standard_normal(n)
    # names shorter than 3 characters make the reference panic in ValidName (rhsBuilder.go:118)
    eqs = ["dphi/dt = -ka*LAP phi - phi^2*psi + LAP psi^2", "dpsi/dt = kb*LAP psi + phi*psi - psi^3 + phi"]
    for stepper in ("rk4", "euler"):
        pm, om = pair_models({"phi": a0, "psi": b0}, {"ka": 0.5, "kb": 1.0}, eqs, n)
        s = G.Solver(pm, dims, 0.05, stepper=stepper)
        s.StepperPropagate(30)
        from oracle.engine import OracleSolver
        o = OracleSolver(om, dims, {"phi": a0, "psi": b0}, 0.05, stepper)
        o.stepper_propagate(30)
        for f in ("phi", "psi"):
            assert maxrel(s.Field(f).real, o.field(f).real) <= 1e-10, (stepper, f)


@pytest.mark.parametrize("case", ["one_field_3d", "two_fields_odd_2d"])
def test_rk4_row_sync_parity(oracle_lib, case):
    """RK4 with every derived field a dense polynomial, so the stages' row pass joins the
    fields' inverse and the derived fields' forward transforms: one field (no second row
    in the transform; derived conc^2, conc^3) at 24^3, and two fields with three derived
    fields (psi^2, phi^2, phi*psi: an unpaired last one) on a 32 x 48 grid, against the
    oracle."""
    if case == "one_field_3d":
        dims, n = [24, 24, 24], 24 ** 3
        rng = np.random.default_rng(21)
        fields = {"conc": 0.5 + 0.05 * rng.standard_normal(n)}
        eqs = ["dconc/dt = ka*LAP conc - LAP conc^2 + kb*LAP conc^3 - conc^2*ka"]
        scal = {"ka": 0.3, "kb": 0.5}
        dt, steps = 0.02, 12
    else:
        dims, n = [32, 48], 32 * 48
        rng = np.random.default_rng(22)
        fields = {"phi": 0.5 + 0.05 * rng.standard_normal(n), "psi": 0.3 + 0.05 * rng.standard_normal(n)}
        eqs = ["dphi/dt = ka*LAP phi - ka*phi^2 + kb*LAP psi^2", "dpsi/dt = kb*LAP psi + phi*psi - psi^2"]
        scal = {"ka": 0.5, "kb": 1.0}
        dt, steps = 0.05, 20
    pm, om = pair_models(fields, scal, eqs, n)
    s = G.Solver(pm, dims, dt, stepper="rk4")
    assert not s.fused
    s.StepperPropagate(steps)
    from oracle.engine import OracleSolver
    o = OracleSolver(om, dims, fields, dt, "rk4")
    o.stepper_propagate(steps)
    for f in fields:
        err = maxrel(s.Field(f).real, o.field(f).real)
        print(case, f, err, flush=True)
        assert err <= 1e-10, (case, f)


def test_short_field_name_panics_like_the_reference():
    m = G.Model()
    m.AddField("a", n=16)
    m.AddEquation("da/dt = a^2")
    with pytest.raises(G.GpfError):
        m.Init()


def test_elastic_two_field_rk4_parity(oracle_lib):
    """Strain-precipitate model (examples/strain_single_precipitate) at 16^3 with cubic C, RK4."""
    N = 16
    dims = [N, N, N]
    n = N ** 3
    conc0, phase0 = W.precipitate_initial(N, 4.0, seed=4)
    Cm = G.CubicMaterial(110.0, 60.0, 30.0)
    misfit = 0.01 * np.eye(3)
    pm, om = pair_models({"conc": conc0, "phase": phase0}, W.PRECIPITATE_SCALARS, W.PRECIPITATE_EQUATIONS, n,
                         explicit={"LIN_ELAST": G.HomogeneousModulusLinElast("phase", Cm, misfit)},
                         polys=W.precipitate_polynomials())
    s = G.Solver(pm, dims, 0.01, stepper="rk4")
    s.Propagate(8)
    from oracle.engine import OracleSolver
    o = OracleSolver(om, dims, {"conc": conc0, "phase": phase0}, 0.01, "rk4")
    o.solver_propagate(8)
    for f in ("conc", "phase"):
        err = maxrel(s.Field(f).real, o.field(f).real)
        print("elastic", f, err, flush=True)
        assert err <= 1e-10


def test_energy_on_demand_matches_oracle(oracle_lib):
    dims = [32, 32]
    n = 32 * 32
    c0 = ch_init(n, 9)
    for fused in (True, False):
        pm, _ = pair_models({"conc": c0}, CH_SCALARS, [CH_EQ], n)
        s = G.Solver(pm, dims, 0.1, fused=fused)
        e, m = s.Energy(0, G.CH_BULK_POLY, 1.3)
        assert abs(e - oracle_lib.ch_energy(dims, c0, G.CH_BULK_POLY, 1.3)) <= 1e-12 * abs(e)
        assert abs(m - c0.sum()) <= 1e-12 * c0.sum()


@pytest.mark.parametrize("fused", [True, False])
def test_set_filter_matches_constructor_filter(fused):
    """TimeStepper.SetFilter with a Vandeven table equals the filter given at construction;
    SetFilter(None) removes it."""
    dims, n = [32, 32], 32 * 32
    c0 = ch_init(n, 3)
    runs = []
    for mode in ("ctor", "set", "none"):
        pm, _ = pair_models({"conc": c0}, CH_SCALARS, [CH_EQ], n)
        s = G.Solver(pm, dims, 0.1, fused=fused, filter_order=4 if mode == "ctor" else 0)
        if mode == "set":
            s.SetFilter(G.NewVandeven(4))
        s.StepperPropagate(10)
        runs.append(s.Field(0).real)
    assert np.array_equal(runs[0], runs[1])
    assert not np.array_equal(runs[0], runs[2])


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_graph_replayed_steps_equal_single_steps(precision):
    """StepperPropagate(n) replays captured 10-step graphs; the result and the energy series
    equal n single steps bit for bit, also after the monitor is switched on mid-run."""
    dims, n = [64, 64], 64 * 64
    c0 = ch_init(n, 4)
    out = []
    for graphs in (True, False):
        pm, _ = pair_models({"conc": c0}, CH_SCALARS, [CH_EQ], n)
        s = G.Solver(pm, dims, 0.05, precision=precision)
        s.SetPersistent(0)  # the per-pass launches (the persistent step has its own tests)

        def advance(k):
            if graphs:
                s.StepperPropagate(k)
            else:
                for _ in range(k):
                    s.StepperPropagate(1)
        advance(25)
        s.MonitorEnergy(G.CH_BULK_POLY, 1.0)
        advance(13)
        e, m = s.EnergySeries()
        out.append((s.Field(0).real, e, m, s.GetTime()))
    assert np.array_equal(out[0][0], out[1][0])
    assert np.array_equal(out[0][1], out[1][1]) and out[0][1].size == 13
    assert np.array_equal(out[0][2], out[1][2])
    assert out[0][3] == out[1][3]


@pytest.mark.parametrize("dims,steps", [([7, 16], 30), ([24, 96], 30), ([12, 24], 30), ([10, 20], 30),
                                        ([96, 192], 10), ([24, 24, 24], 10), ([12, 12, 12], 10),
                                        ([48, 48, 48], 5)])
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_fused_ch_sizes_match_oracle(oracle_lib, dims, steps, precision):
    """The fused step on row counts that leave an unpaired row, radix-3 lengths, lengths
    transformed by the brute-force DFT, and 3D cubes: FP64 within 1e-10 max relative,
    FP32-bulk within 1e-5 relative L2, of the oracle."""
    n = int(np.prod(dims))
    c0 = ch_init(n, 7)
    pm, om = pair_models({"conc": c0}, CH_SCALARS, [CH_EQ], n)
    print("fused sizes", dims, steps, precision, flush=True)
    s = G.Solver(pm, dims, 0.05, precision=precision, autotune=True)
    assert s.fused
    s.StepperPropagate(steps)
    from oracle.engine import OracleSolver
    o = OracleSolver(om, dims, {"conc": c0}, 0.05)
    o.stepper_propagate(steps)
    got, ref = s.Field(0).real, o.field("conc").real
    if precision == "fp64":
        assert maxrel(got, ref) <= 1e-10
    else:
        assert np.linalg.norm(got - ref) / np.linalg.norm(ref) <= 1e-5


@pytest.mark.parametrize("dims", [[80, 100], [160, 320], [16, 2048], [2048, 16], [112, 224], [80, 80, 80]])
def test_ch_radix5_radix7_and_2048_lengths_match_oracle(oracle_lib, dims):
    """The fused Cahn-Hilliard step on grids whose transform lengths use the radix-5 and
    radix-7 stages and the 2^11 schedule (rows, pencils and the update pass), 10 FP64 steps
    against the oracle (euler.go:16-47)."""
    import gopf_amd as G
    from gopf_amd import workloads as W
    from oracle import parser as P
    from oracle.engine import OracleSolver
    print("ch lengths", dims, flush=True)
    n = int(np.prod(dims))
    c0 = W.ch_initial(3, n)
    m = G.Model()
    m.AddField("conc", c0)
    om = P.Model()
    om.add_field("conc", n)
    for k, v in W.ch_scalars().items():
        m.AddScalar(k, v)
        om.add_scalar(k, v)
    m.AddEquation(W.CH_EQUATION)
    om.add_equation(W.CH_EQUATION)
    m.Init()
    om.init()
    s = G.Solver(m, dims, 0.1, precision="fp64")
    assert s.fused
    s.StepperPropagate(10)
    o = OracleSolver(om, dims, {"conc": c0}, 0.1)
    o.stepper_propagate(10)
    ref = o.field("conc").real
    err = np.abs(s.Field(0).real - ref).max() / np.abs(ref).max()
    assert err <= 1e-10, (dims, err)


@pytest.mark.parametrize("dims,precision", [([9, 15], "fp64"), ([15, 9], "fp64"), ([24, 33], "fp64"),
                                            ([15, 15, 15], "fp64"), ([9, 15], "fp32")])
def test_odd_fastest_extent_matches_oracle(oracle_lib, dims, precision):
    """An odd fastest extent, as FFTW plans it (pfutil/fftWrap.go:16-23): the half spectrum
    keeps columns kx = 0 .. nx/2, whose last one is then an ordinary frequency with a mirror
    at -kx.  The fused Cahn-Hilliard step (direct-DFT rows for lengths <= 64), 20 Euler
    steps against the oracle: <= 1e-10 FP64, <= 1e-5 relative L2 FP32."""
    from oracle.engine import OracleSolver
    print("odd fastest extent", dims, precision, flush=True)
    n = int(np.prod(dims))
    c0 = ch_init(n, 5)
    pm, om = pair_models({"conc": c0}, CH_SCALARS, [CH_EQ], n)
    s = G.Solver(pm, dims, 0.05, precision=precision)
    assert s.fused
    s.StepperPropagate(20)
    o = OracleSolver(om, dims, {"conc": c0}, 0.05)
    o.stepper_propagate(20)
    got, ref = s.Field(0).real, o.field("conc").real
    if precision == "fp64":
        assert maxrel(got, ref) <= 1e-10
    else:
        assert relL2(got, ref) <= 1e-5
    assert maxrel(ref, c0) > 1e-3


def test_odd_fastest_extent_two_fields_rk4_matches_oracle(oracle_lib):
    """The general path (RK4, two fields, dense derived fields in the stages' row pass) on a
    12 x 15 grid against the oracle, 20 steps, <= 1e-10."""
    from oracle.engine import OracleSolver
    dims, n = [12, 15], 12 * 15
    rng = np.random.default_rng(23)
    fields = {"phi": 0.5 + 0.05 * rng.standard_normal(n), "psi": 0.3 + 0.05 * rng.standard_normal(n)}
    eqs = ["dphi/dt = ka*LAP phi - ka*phi^2 + kb*LAP psi^2", "dpsi/dt = kb*LAP psi + phi*psi - psi^2"]
    scal = {"ka": 0.5, "kb": 1.0}
    pm, om = pair_models(fields, scal, eqs, n)
    s = G.Solver(pm, dims, 0.05, stepper="rk4")
    assert not s.fused
    s.StepperPropagate(20)
    o = OracleSolver(om, dims, fields, 0.05, "rk4")
    o.stepper_propagate(20)
    for f in fields:
        assert maxrel(s.Field(f).real, o.field(f).real) <= 1e-10, f


def test_odd_fastest_extent_beyond_64_is_refused():
    """Odd lengths above 64 have no compiled schedule (the direct DFT covers <= 64)."""
    m = G.Model()
    m.AddField("conc", ch_init(16 * 65, 3))
    for k, v in CH_SCALARS.items():
        m.AddScalar(k, v)
    m.AddEquation(CH_EQ)
    m.Init()
    with pytest.raises(G.GpfError) as ei:
        G.Solver(m, [16, 65], 0.1)
    assert ei.value.code == G.GPF_EINVAL and "odd" in str(ei.value)


