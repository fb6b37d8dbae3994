"""The SDD stepper on the device: the reference's known answers (pf/sdd_test.go) and parity
with the oracle's restatement (oracle/sdd.py, real fields), FP64.

Parity bar: max|gpu - oracle| / max|oracle| <= 1e-9 over the stated steps."""
import math

import numpy as np
import pytest

import gopf_amd as G
from oracle import parser as P
from oracle.sdd import SDDOracle

pytestmark = pytest.mark.gpu


def maxrel(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


def box_blur(x, N, width=2):
    a = np.asarray(x, dtype=float).reshape(N, N)
    cut = width // 2 + 1
    acc = np.zeros_like(a)
    cnt = 0
    for dy in range(-cut, cut + 1):
        for dx in range(-cut, cut + 1):
            if abs(dy) <= width and abs(dx) <= width:
                acc += np.roll(np.roll(a, -dy, axis=0), -dx, axis=1)
                cnt += 1
    return (acc / cnt).reshape(-1)


def circle(N, radius):
    idx = np.arange(N * N)
    r, c = idx // N, idx % N
    out = np.full(N * N, -1.0)
    out[(r - N // 2) ** 2 + (c - N // 2) ** 2 <= radius * radius] = 1.0
    return out


def double_well(N=4, dt=0.1):
    init, final = np.full(N * N, -1.0), np.full(N * N, 1.5)
    c0 = 0.5 * (init + final)
    pm, om = G.Model(), P.Model()
    pm.AddField("conc", c0)
    om.add_field("conc", N * N)
    pm.AddEquation("dconc/dt = conc - conc^3")
    om.add_equation("dconc/dt = conc - conc^3")
    pm.Init()
    om.init()
    s = G.Solver(pm, [N, N], dt, stepper="sdd")
    s.SDDInit([init], [final])
    s.SDDSetParams(InitDimerLength=0.1)
    o = SDDOracle(om, [N, N], {"conc": c0.astype(complex)}, dt)
    o.init([init], [final])
    o.init_dimer_length = 0.1
    return s, o


def test_double_well():
    """TestDoubleWell (sdd_test.go:68-108): the saddle between -1 and 1.5 is 0."""
    s, o = double_well()
    nsteps = int(math.log(1e6) / 0.1) + 1      # RequiredDimerLengthTime(1e-6 l0) / dt + 1
    for k in range(nsteps):
        s.Propagate(1)
        o.step()
        if k == 49:
            assert maxrel(s.Field("conc").real, o.field("conc").real) <= 1e-9
    assert np.abs(s.Field("conc").real).max() <= 1e-6
    mon = s.SDDMonitor()
    assert mon["MaxForce"] <= 1e-3


def test_surface_two_fields():
    """Test2DSurface (sdd_test.go:110-163) on a 2 x 2 grid of uniform fields: the saddle of
    E = (x^2 - 1)^2 + y^2 is (0, 0) and the orientation ends along x."""
    N, dt = 2, 0.001
    pm = G.Model()
    pm.AddField("xCrd", np.full(N * N, 0.2))
    pm.AddField("yCrd", np.full(N * N, 0.3))
    pm.AddScalar("FOUR", 4.0)
    pm.AddScalar("TWO", 2.0)
    pm.AddEquation("dxCrd/dt = FOUR*xCrd - FOUR*xCrd^3")
    pm.AddEquation("dyCrd/dt = -TWO*yCrd")
    pm.Init()
    s = G.Solver(pm, [N, N], dt, stepper="sdd")
    s.SDDSetInitialOrientation([0.1] * (N * N) + [0.5] * (N * N))
    l0 = s.SDDDimerLength(0.0)
    steps = int(math.log(l0 / 1e-5) / dt)
    s.Propagate(steps)
    x, y = s.Field("xCrd").real, s.Field("yCrd").real
    assert np.abs(x).max() <= 1e-3 and np.abs(y).max() <= 1e-3
    v = s.SDDOrientation()
    assert np.abs(v[:N * N] - 0.5).max() <= 1e-3 and np.abs(v[N * N:]).max() <= 1e-3


def nucleation_models(N=64, dt=0.7):
    gamma, rho = 0.5, 0.05
    a = 3.0 * rho / 4.0
    poly = [(a, {"phi": 0.0}), (1.0, {"phi": 1.0}), (-a, {"phi": 2.0}), (-1.0, {"phi": 3.0})]
    init = box_blur(circle(N, 10), N)
    final = box_blur(circle(N, 15), N)
    field = box_blur(circle(N, 12), N)
    pm, om = G.Model(), P.Model()
    pm.AddField("phi", field)
    om.add_field("phi", N * N)
    pm.AddScalar("gamma", gamma)
    om.add_scalar("gamma", gamma)
    pm.RegisterPolynomial("MINUS_CHEM_POT", poly)
    om.register_polynomial("MINUS_CHEM_POT", poly)
    pm.AddEquation("dphi/dt = MINUS_CHEM_POT + gamma*LAP phi")
    om.add_equation("dphi/dt = MINUS_CHEM_POT + gamma*LAP phi")
    pm.Init()
    om.init()
    s = G.Solver(pm, [N, N], dt, stepper="sdd")
    s.SDDInit([init], [final])
    s.SDDSetParams(MinDimerLength=5e-6)
    o = SDDOracle(om, [N, N], {"phi": field.astype(complex)}, dt)
    o.min_dimer_length = 5e-6
    o.init([init], [final])
    rc = 2.0 * math.sqrt(gamma / 2.0) * (2.0 / 3.0) / rho
    return s, o, rc


def test_classical_nucleation_parity():
    s, o, _ = nucleation_models()
    for _ in range(40):
        s.Propagate(1)
        o.step()
    assert maxrel(s.Field("phi").real, o.field("phi").real) <= 1e-9
    v = s.SDDOrientation()
    assert np.abs(v - o.orientation).max() / np.abs(o.orientation).max() <= 1e-8
    m = s.SDDMonitor()
    assert abs(m["MaxTorque"] - o.monitor["MaxTorque"]) <= 1e-8 * max(1.0, o.monitor["MaxTorque"])
    assert abs(m["FieldNorm"] - o.monitor["FieldNorm"]) <= 1e-9 * o.monitor["FieldNorm"]


def test_classical_nucleation_radius():
    """TestClassicalNucleation (sdd_test.go:165-271): the critical droplet's radius is 2 sigma / rho
    within 0.3."""
    s, _, rc = nucleation_models()
    s.Propagate(int(1000.0 / 0.7))
    phi = s.Field("phi").real
    assert np.all(np.isfinite(phi))
    Rc = math.sqrt(float(np.sum(0.5 * (1.0 + phi))) / math.pi)
    assert abs(Rc - rc) <= 0.3, (Rc, rc)


def test_sdd_errors():
    m = G.Model()
    m.AddField("conc", np.zeros(16))
    m.AddEquation("dconc/dt = conc - conc^3")
    m.Init()
    s = G.Solver(m, [4, 4], 0.1, stepper="sdd")
    with pytest.raises(G.GpfError):
        s.Propagate(1)                                # not initialised (sdd.go:162-164)
    with pytest.raises(G.GpfError):
        s.SDDSetInitialOrientation(np.ones(15))       # wrong length (sdd.go:407-409)
    with pytest.raises(G.GpfError):
        G.Solver(m, [4, 4], 0.1, stepper="sdd", filter_order=4)   # SetFilter panics (sdd.go:422-426)
    with pytest.raises(G.GpfError):
        G.Solver(m, [4, 4], 0.1, stepper="sdd", precision="fp32")
