"""FP64 device reductions (gpf_reduce, gpf_solver_reduce): sum, sum of squares, min, max and
the count of non-finite values, the device form of the NaN scan a gopf driver runs over its
fields to stop an unstable run (examples/solverStability/main.go:58-67).

Checked against exact host sums (math.fsum) of the same values, numpy's min/max over the
finite values, and the field the solver returns; a NaN planted in a field and an Euler run
at an unstable dt must be reported."""
import math

import numpy as np
import pytest

import gopf_amd as G
from gopf_amd import workloads as W

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _expected(x: np.ndarray):
    v = x.astype(np.float64)
    fin = v[np.isfinite(v)]
    return {"sum": math.fsum(fin), "sum_sq": math.fsum(fin * fin),
            "min": float(fin.min()) if fin.size else math.inf, "max": float(fin.max()) if fin.size else -math.inf,
            "nonfinite": int(v.size - fin.size)}


def _check(got, want, rtol=1e-15):
    assert got["nonfinite"] == want["nonfinite"]
    assert got["min"] == want["min"] and got["max"] == want["max"]
    for k in ("sum", "sum_sq"):
        assert abs(got[k] - want[k]) <= rtol * max(1.0, abs(want[k])), (k, got[k], want[k])


@pytest.mark.parametrize("prec,dtype", [("fp64", np.float64), ("fp32", np.float32)])
@pytest.mark.parametrize("n", [0, 1, 65535, 65536, 1_000_003])
def test_reduce_matches_exact_sums(prec, dtype, n):
    print(f"gpf_reduce {prec} n={n}", flush=True)
    rng = np.random.default_rng(n + 7)
    x = (rng.standard_normal(n) * 3.0 + 0.5).astype(dtype)
    if n > 100:
        x[[3, n // 2, n - 1]] = [np.nan, np.inf, -np.inf]
    d = torch.from_numpy(x).to("cuda")
    got = G.Reduce(d.data_ptr(), n, prec)
    _check(got, _expected(x))


def test_solver_reduce_matches_the_returned_field():
    for prec in ("fp64", "fp32"):
        N = 64
        c0 = W.ch_initial(3, N * N)
        m = G.Model()
        m.AddField("conc", c0)
        for k, v in W.ch_scalars().items():
            m.AddScalar(k, v)
        m.AddEquation(W.CH_EQUATION)
        m.Init()
        s = G.Solver(m, [N, N], 0.1, precision=prec)
        s.StepperPropagate(7)
        c = s.Field("conc").real
        r = s.Reduce("conc")
        _check(r, _expected(c), rtol=1e-14)
        assert r["nonfinite"] == 0


def test_planted_nan_and_unstable_run_are_reported():
    N = 32
    rng = np.random.default_rng(0)
    m = G.Model()
    c0 = 2.0 * rng.random(N * N) - 1.0
    m.AddField("conc", c0)
    m.AddScalar("gamma", 0.02)
    m.RegisterPolynomial("CHEMPOT", [(1.0, {"conc": 1.0}), (-1.0, {"conc": 3.0})])
    m.AddEquation("dconc/dt = CHEMPOT + gamma*LAP conc")
    m.Init()
    s = G.Solver(m, [N, N], 0.1, stepper="rk4")
    bad = c0.copy()
    bad[17] = np.nan
    s.SetField("conc", bad)
    r = s.Reduce("conc")
    assert r["nonfinite"] == 1
    _check(r, _expected(bad), rtol=1e-14)
    # the solverStability example's Allen-Cahn model at dt = 3.4 (Euler blows up)
    e = G.Solver(m, [N, N], 3.4, stepper="euler")
    e.SetField("conc", c0)
    e.Propagate(100)
    r = e.Reduce("conc")
    host = e.Field("conc").real
    assert r["nonfinite"] == int(np.sum(~np.isfinite(host))) > 0
