"""Device adjoint of the fused Euler step (gpf_solver_adjoint).

No reference counterpart: the gradients are pinned by the oracle's term-by-term FP64
reverse mode (itself pinned by central differences on the CPU) and by central
differences (h = 1e-6) of the device's own FP64 forward solver.

Bars: FP64 against the oracle, loss within 1e-12 and gradients within 1e-10 relative
(max norm for the field gradient); against central differences 1e-6 relative; FP32
against FP64 (128^2, 200 steps) within 1e-4 relative (loss), 1e-3 relative L2 (field
gradient) and 1e-3 relative (scalars).
"""
import numpy as np
import pytest

import gopf_amd as G
from gopf_amd import workloads as W

pytestmark = pytest.mark.gpu


def gpu_solver(eq, scal, dims, c0, dt, precision="fp64", filt=0):
    m = G.Model()
    m.AddField("conc", c0)
    for k, v in scal.items():
        m.AddScalar(k, v)
    m.AddEquation(eq)
    m.Init()
    s = G.Solver(m, dims, dt, precision=precision, filter_order=filt)
    assert s.fused
    return s


def oracle_adjoint(eq, scal, dims, c0, dt, steps, tgt, filt=0):
    from oracle import parser as P
    from oracle.engine import OracleSolver
    m = P.Model()
    m.add_field("conc", int(np.prod(dims)))
    for k, v in scal.items():
        m.add_scalar(k, v)
    m.add_equation(eq)
    m.init()
    return OracleSolver(m, dims, {"conc": c0}, dt, filter_order=filt).adjoint(steps, tgt, list(scal))


CASES = [
    ("ch_2d", W.CH_EQUATION, W.ch_scalars(1.0, 1.0), [32, 32], 0.05, 50, 7, 0),
    ("ch_2d_rect_filter", W.CH_EQUATION, W.ch_scalars(0.7, 1.2), [32, 64], 0.05, 30, 30, 4),
    ("linear_explicit", "dconc/dt = mg*conc*conc + mk*LAP conc", {"mg": 0.2, "mk": 0.1}, [16, 32], 0.05, 25, 4, 0),
    ("ch_3d", W.CH_EQUATION, W.ch_scalars(1.0, 1.0), [16, 16, 16], 0.05, 12, 5, 0),
]


@pytest.mark.parametrize("name,eq,scal,dims,dt,steps,every,filt", CASES, ids=[c[0] for c in CASES])
def test_adjoint_fp64_matches_oracle(oracle_lib, name, eq, scal, dims, dt, steps, every, filt):
    n = int(np.prod(dims))
    print(f"# adjoint {name} dims={dims} steps={steps} every={every}", flush=True)
    c0 = W.ch_initial(5, n)
    tgt = W.ch_initial(6, n)
    s = gpu_solver(eq, scal, dims, c0, dt, filt=filt)
    J, g0, gs = s.Adjoint(steps, tgt, list(scal), checkpoint_every=every)
    Jo, go, gso = oracle_adjoint(eq, scal, dims, c0, dt, steps, tgt, filt)
    assert abs(J - Jo) <= 1e-12 * abs(Jo)
    assert np.abs(g0 - go).max() <= 1e-10 * np.abs(go).max()
    for k in scal:
        assert abs(gs[k] - gso[k]) <= 1e-10 * abs(gso[k]), (k, gs[k], gso[k])
    # the solver is left at c_T and its time advanced
    ref = gpu_solver(eq, scal, dims, c0, dt, filt=filt)
    ref.StepperPropagate(steps)
    assert np.array_equal(s.Field(0).real, ref.Field(0).real)
    assert s.GetTime() == pytest.approx(steps * dt)


@pytest.mark.parametrize("wanted", [[], ["mk"]], ids=["field_only", "implicit_scalar"])
@pytest.mark.parametrize("name,eq,scal,dims,dt,steps,every,filt", CASES, ids=[c[0] for c in CASES])
def test_adjoint_fused_reverse_matches_oracle(oracle_lib, name, eq, scal, dims, dt, steps, every, filt, wanted):
    """Gradients for the field and for a scalar of the implicit part only: no scalar needs the
    source-derivative sums, so each reverse step's inverse rows, lam = p + g'(c) q and the
    next forward rows are one row pass."""
    n = int(np.prod(dims))
    print(f"# adjoint {name} dims={dims} steps={steps} every={every} scalars={wanted}", flush=True)
    c0 = W.ch_initial(5, n)
    tgt = W.ch_initial(6, n)
    s = gpu_solver(eq, scal, dims, c0, dt, filt=filt)
    J, g0, gs = s.Adjoint(steps, tgt, wanted, checkpoint_every=every)
    Jo, go, gso = oracle_adjoint_subset(eq, scal, dims, c0, dt, steps, tgt, wanted, filt)
    assert abs(J - Jo) <= 1e-12 * abs(Jo)
    assert np.abs(g0 - go).max() <= 1e-10 * np.abs(go).max()
    for k in wanted:
        assert abs(gs[k] - gso[k]) <= 1e-10 * abs(gso[k]), (k, gs[k], gso[k])


def oracle_adjoint_subset(eq, scal, dims, c0, dt, steps, tgt, wanted, filt=0):
    from oracle import parser as P
    from oracle.engine import OracleSolver
    m = P.Model()
    m.add_field("conc", int(np.prod(dims)))
    for k, v in scal.items():
        m.add_scalar(k, v)
    m.add_equation(eq)
    m.init()
    return OracleSolver(m, dims, {"conc": c0}, dt, filter_order=filt).adjoint(steps, tgt, list(wanted))


def test_adjoint_checkpoint_interval_does_not_change_the_gradient():
    dims, n = [32, 32], 1024
    c0 = W.ch_initial(5, n)
    tgt = W.ch_initial(6, n)
    res = []
    for every in (1, 9, 40, 100):
        s = gpu_solver(W.CH_EQUATION, W.ch_scalars(), dims, c0, 0.05)
        res.append(s.Adjoint(40, tgt, ["mk"], checkpoint_every=every))
    for J, g0, gs in res[1:]:
        assert J == res[0][0]
        assert np.array_equal(g0, res[0][1])
        assert gs["mk"] == res[0][2]["mk"]


def test_adjoint_rejects_unsupported_inputs():
    dims, n = [16, 16], 256
    s = gpu_solver(W.CH_EQUATION, W.ch_scalars(), dims, W.ch_initial(5, n), 0.05)
    with pytest.raises(G.GpfError):
        s.Adjoint(5, np.zeros(n), ["nosuch"])
    with pytest.raises(G.GpfError):
        s.Adjoint(0, np.zeros(n))
    with pytest.raises(ValueError):
        s.Adjoint(5, np.zeros(n + 1))


def forward_loss(scal, dims, c0, dt, steps, tgt):
    s = gpu_solver(W.CH_EQUATION, scal, dims, c0, dt)
    s.StepperPropagate(steps)
    r = s.Field(0).real - tgt
    return float(np.sum(r * r))


def central4(f, h):
    """Fourth-order central difference (8 (f(h) - f(-h)) - (f(2h) - f(-2h))) / 12h."""
    return (8.0 * (f(h) - f(-h)) - (f(2 * h) - f(-2 * h))) / (12.0 * h)


def test_adjoint_matches_central_differences_adjoint_config():
    """512^2 Cahn-Hilliard, 2000 steps (dt = 0.1), checkpoints every 100 steps; the target
    is a forward run at kappa = 1.3 from seed 6; gradients with respect to kappa (mk = M kappa)
    and to c_0 along 16 seeded unit directions, against fourth-order central differences
    with h = 1e-6 of the FP64 forward solver.

    After 2000 steps J(kappa) is curved enough that the second-order difference with
    h = 1e-6 is itself off by 2.6e-4 (its error falls as h^2: 2.5e-2 at h = 1e-5), hence the
    fourth-order stencil.  Along c_0 the differences carry about 1e-3 of rounding noise (the
    spread between the second- and fourth-order estimates), which is 1e-6 relative only for
    directional derivatives above 1e3: the 16 directions are held to 1e-6 as a vector
    (relative L2 over the directions) and to 1e-5 one by one."""
    N, steps, dt, h = 512, 2000, 0.1, 1e-6
    dims, n = [N, N], N * N
    c0 = W.ch_initial(5, n)
    ts = gpu_solver(W.CH_EQUATION, W.ch_scalars(1.0, 1.3), dims, W.ch_initial(6, n), dt)
    ts.StepperPropagate(steps)
    tgt = ts.Field(0).real
    scal = W.ch_scalars(1.0, 1.0)
    s = gpu_solver(W.CH_EQUATION, scal, dims, c0, dt)
    J, g0, gs = s.Adjoint(steps, tgt, ["mk"], checkpoint_every=100)
    dkappa = 1.0 * gs["mk"]
    fdk = central4(lambda d: forward_loss(W.ch_scalars(1.0, 1.0 + d), dims, c0, dt, steps, tgt), h)
    rk = abs(fdk - dkappa) / abs(fdk)
    print(f"# J={J:.12e} dJ/dkappa adjoint={dkappa:.12e} fd={fdk:.12e} rel={rk:.3e}", flush=True)
    assert rk <= 1e-6
    rng = np.random.default_rng(5)
    ad, fd = [], []
    for d in range(16):
        e = rng.standard_normal(n)
        e /= np.linalg.norm(e)
        fd.append(central4(lambda t: forward_loss(scal, dims, c0 + t * e, dt, steps, tgt), h))
        ad.append(float(g0 @ e))
        print(f"# direction {d}: adjoint={ad[-1]:.12e} fd={fd[-1]:.12e} "
              f"rel={abs(fd[-1] - ad[-1]) / abs(fd[-1]):.3e}", flush=True)
    ad, fd = np.array(ad), np.array(fd)
    vec = np.linalg.norm(ad - fd) / np.linalg.norm(fd)
    print(f"# 16 directions: relative L2 {vec:.3e}", flush=True)
    assert vec <= 1e-6
    assert np.all(np.abs(ad - fd) <= 1e-5 * np.abs(fd))


def test_adjoint_fp32_close_to_fp64():
    dims, n, steps = [128, 128], 128 * 128, 200
    c0 = W.ch_initial(5, n)
    tgt = W.ch_initial(6, n)
    r = {}
    for p in ("fp64", "fp32"):
        s = gpu_solver(W.CH_EQUATION, W.ch_scalars(), dims, c0, 0.1, precision=p)
        r[p] = s.Adjoint(steps, tgt, ["mk", "m2"], checkpoint_every=50)
    J64, g64, s64 = r["fp64"]
    J32, g32, s32 = r["fp32"]
    rel = np.linalg.norm(g32 - g64) / np.linalg.norm(g64)
    print(f"# fp32 vs fp64: J {abs(J32 - J64) / J64:.3e} grad {rel:.3e} "
          f"mk {abs(s32['mk'] - s64['mk']) / abs(s64['mk']):.3e}", flush=True)
    assert abs(J32 - J64) <= 1e-4 * J64
    assert rel <= 1e-3
    for k in ("mk", "m2"):
        assert abs(s32[k] - s64[k]) <= 1e-3 * abs(s64[k])
