"""Batched sweep of uneven single-field sims against independent solvers and the oracle.

Each sim of a batch must reproduce what one pf.Solver with that sim's scalars
computes (FP64: max relative error <= 1e-12 against the single-sim HIP path,
<= 1e-10 against the oracle); the FP32-bulk batch stays within 1e-5 relative L2
of the FP64 batch and conserves each sim's mass to 1e-9.
"""
import numpy as np
import pytest

import gopf_amd as G
from gopf_amd import workloads as W

pytestmark = pytest.mark.gpu

CLASSES = [((16, 16), 2), ((32, 32), 3), ((32, 64), 2)]


def ch_model(n):
    m = G.Model()
    m.AddField("conc", n=n)
    for k, v in W.ch_scalars().items():
        m.AddScalar(k, v)
    m.AddEquation(W.CH_EQUATION)
    m.Init()
    return m


def sim_params(nsims):
    rng = np.random.default_rng(42)
    M = np.exp(rng.uniform(np.log(0.1), np.log(10.0), nsims))
    kappa = np.exp(rng.uniform(np.log(0.1), np.log(10.0), nsims))
    return M, kappa


def make_batch(precision):
    b = G.Batch(ch_model(16 * 16), CLASSES, 0.02, precision=precision)
    M, kappa = sim_params(len(b))
    inits = []
    for s in range(len(b)):
        for k, v in W.ch_scalars(M[s], kappa[s]).items():
            b.SetScalar(s, k, v)
        c0 = W.ch_initial(1000 + s, int(np.prod(b.shapes[s])))
        b.SetField(s, c0)
        inits.append(c0)
    return b, M, kappa, inits


def test_batch_fp64_matches_single_solvers(oracle_lib):
    b, M, kappa, inits = make_batch("fp64")
    assert len(b) == 7
    b.Propagate(25)   # two replayed 10-step graphs and 5 eager steps
    b.Propagate(13)   # the other state-buffer parity
    from oracle import parser as P
    from oracle.engine import OracleSolver
    for s in range(len(b)):
        dims = list(b.shapes[s])
        n = int(np.prod(dims))
        m = G.Model()
        m.AddField("conc", inits[s])
        for k, v in W.ch_scalars(M[s], kappa[s]).items():
            m.AddScalar(k, v)
        m.AddEquation(W.CH_EQUATION)
        m.Init()
        single = G.Solver(m, dims, 0.02)
        single.StepperPropagate(38)
        got = b.Field(s).real
        ref = single.Field(0).real
        assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max(), s
        if s in (0, 3, 6):
            om = P.Model()
            om.add_field("conc", n)
            for k, v in W.ch_scalars(M[s], kappa[s]).items():
                om.add_scalar(k, v)
            om.add_equation(W.CH_EQUATION)
            om.init()
            o = OracleSolver(om, dims, {"conc": inits[s]}, 0.02)
            o.stepper_propagate(38)
            oref = o.field("conc").real
            assert np.abs(got - oref).max() <= 1e-10 * np.abs(oref).max(), s


def test_batch_fp32_close_to_fp64_and_conserves_mass():
    b64, _, _, inits = make_batch("fp64")
    b32, _, _, _ = make_batch("fp32")
    b64.Propagate(30)
    b32.Propagate(30)
    for s in range(len(b64)):
        a = b32.Field(s).real
        r = b64.Field(s).real
        assert np.linalg.norm(a - r) / np.linalg.norm(r) <= 1e-5, s
        assert abs(a.sum() - inits[s].sum()) <= 1e-9 * inits[s].sum(), s


@pytest.mark.parametrize("prec", ["fp32", "fp64"])
def test_grouped_launches_match_per_class_launches(prec):
    """One grouped row launch and one grouped update launch per step (every class in the same
    launch, 128 / 256 / 512 threads per block) against one launch per class and pass, with the
    same transform schedules: the per-line arithmetic is the same, so the fields must agree
    bit for bit after 25 steps (two replayed 10-step graphs and 5 direct steps)."""
    classes = [((64, 64), 3), ((128, 256), 2), ((256, 128), 2)]
    m = G.Model()
    m.AddField("conc", n=16)
    for k, v in W.ch_scalars().items():
        m.AddScalar(k, v)
    m.AddEquation(W.CH_EQUATION)
    m.Init()
    fields = {}
    for th in (0, 128, 256, 512):
        print(f"batch {prec} grouped threads {th}", flush=True)
        b = G.Batch(m, classes, 0.1, precision=prec, autotune=False)
        for s in range(len(b)):
            M, kappa, c0 = W.c3_sim_inputs(s, int(np.prod(b.shapes[s])))
            for k, v in W.ch_scalars(M, kappa).items():
                b.SetScalar(s, k, v)
            b.SetField(s, c0)
        b.SetLayout({"classes": [{"rows_tw": 8, "update_tw": 8, "rows_schedule": 0}] * len(classes),
                     "grouped_threads": th})
        assert b.Tuning()["grouped_threads"] == th
        b.Propagate(25)
        fields[th] = [b.Field(s).real for s in range(len(b))]
        del b
    for th in (128, 256, 512):
        for s in range(len(fields[0])):
            assert np.array_equal(fields[th][s], fields[0][s]), (th, s)
    # the step moved the fields
    c0 = W.c3_sim_inputs(0, 64 * 64)[2]
    assert np.abs(fields[0][0] - c0).max() > 1e-3


def test_batch_layout_refusals():
    m = G.Model()
    m.AddField("conc", n=16)
    for k, v in W.ch_scalars().items():
        m.AddScalar(k, v)
    m.AddEquation(W.CH_EQUATION)
    m.Init()
    b = G.Batch(m, [((96, 96), 2)], 0.1, precision="fp32", autotune=False)
    with pytest.raises(G.GpfError):   # 96 has no grouped launch
        b.SetLayout({"classes": [{"rows_tw": 8, "update_tw": 8, "rows_schedule": 0}], "grouped_threads": 256})
    with pytest.raises(G.GpfError):   # lines per block must be a power of two
        b.SetLayout({"classes": [{"rows_tw": 6, "update_tw": 8, "rows_schedule": 0}], "grouped_threads": 0})
