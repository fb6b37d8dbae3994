"""The oracle's VolumeConservingLP and SpectralViscosity against the reference's tests
(pf/volumeConserving_test.go, pf/spectralViscosity_test.go), restated."""
import numpy as np
import pytest

from oracle import engine as E
from oracle import parser as P


@pytest.fixture(scope="module", autouse=True)
def _built():
    E.build()


def test_interpolation():
    """TestInterpolation (spectralViscosity_test.go:10-35)."""
    L = E.lib()
    for k, expect in ((0.19, 0.0), (0.66, 1.0), (0.4, 1.0 / 8.0)):
        assert abs(L.or_sv_interpolant(k, 0.6) - expect) <= 1e-10


def test_spectral_viscosity_term():
    """TestSpectralViscosityTerm (spectralViscosity_test.go:37-72): one implicit term whose
    value is -Eps Q(|f|) |f|^2."""
    N = 16
    m = P.Model()
    m.add_field("conc", N * N)
    m.register_implicit_term("SPECTRAL_VISC", P.KIND_SVISC, eps=1.0, threshold=0.25, power=2)
    m.add_equation("dconc/dt = SPECTRAL_VISC")
    m.init()
    terms, denum = m.rhs[0]
    assert len(denum) == 1 and len(terms) == 0
    o = E.OracleSolver(m, [N, N], {"conc": np.zeros(N * N, dtype=complex)}, 0.1)
    res = o.get_denum(0)
    for i in range(N * N):
        f = np.array(E.freq([N, N], i))
        f2 = float(f @ f)
        expect = -1.0 * E.lib().or_sv_interpolant(np.sqrt(f2), 0.25) * f2
        assert abs(res[i].real - expect) <= 1e-10 and abs(res[i].imag) <= 1e-10


def test_spectral_viscosity_closed_form():
    """Q's three branches at sample radii (Tadmor's interpolant, threshold m): 0 below m/3,
    2x^2 - 3x^3 with x = 3f/2m - 1/2 between, 1 above m."""
    L = E.lib()
    m = 0.3
    for f in (0.0, 0.05, 0.099):
        assert L.or_sv_interpolant(f, m) == 0.0
    for f in (0.31, 0.5):
        assert L.or_sv_interpolant(f, m) == 1.0
    for f in (0.15, 0.2, 0.25):
        x = 1.5 * f / m - 0.5
        assert abs(L.or_sv_interpolant(f, m) - (2 * x * x - 3 * x ** 3)) <= 1e-14


def _volcons_model(N):
    m = P.Model()
    m.add_field("myfield", N * N)
    m.add_field("indicator", N * N)
    m.register_explicit_term("CONS", P.KIND_VOLCONS, field="myfield", indicator="indicator", dt=0.01,
                             num_nodes=N * N)
    m.add_equation("dmyfield/dt = CONS")
    m.add_equation("dindicator/dt = LAP indicator")
    m.init()
    return m


def test_vol_conserve():
    """TestVolConserve (volumeConserving_test.go:15-62): dmyfield/dt = rate + correction keeps
    the volume, apart from the two steps before the multiplier starts to act."""
    N, dt, rate = 16, 0.01, 0.01
    m = _volcons_model(N)
    indicator = np.zeros(N * N, dtype=complex)
    indicator[N * N // 2 + 1:] = 1.0
    myfield = np.ones(N * N, dtype=complex)
    o = E.OracleSolver(m, [N, N], {"myfield": myfield, "indicator": indicator}, dt)
    ind_k = np.fft.fft2(indicator.reshape(N, N)).reshape(-1)
    for _ in range(100):
        o.set_field("indicator", indicator)        # IFFT'd indicator when the term is evaluated
        correction = o.eval_user_term(P.KIND_VOLCONS, 0)
        myfield = myfield + dt * (rate + correction)
        o.set_field("myfield", myfield)
        o.set_field("indicator", ind_k)            # FFT'd again before OnStepFinished
        o.on_step_finished()
    volume = float(np.sum(myfield.real))
    expect_change = 2.0 * dt * N * N * rate
    assert abs(volume - expect_change - 256.0) <= 1e-10


def test_vol_conserve_planted_mistake_fails():
    """Updating the multiplier from the first step on (no IsFirstUpdate) misses the reference's
    volume by dt N^2 rate: the known answer catches it."""
    N, dt, rate = 16, 0.01, 0.01
    indicator = np.zeros(N * N)
    indicator[N * N // 2 + 1:] = 1.0
    mult, current, vol = 0.0, float(N * N), None
    field = np.ones(N * N)
    for _ in range(100):
        field = field + dt * (rate + mult * indicator)
        fi = float(field.sum())
        mult -= (fi - current) / (dt * indicator.sum())
        current = fi
    vol = float(field.sum())
    assert abs(vol - 2.0 * dt * N * N * rate - 256.0) > 1e-3


def test_volcons_indicator_must_exist():
    m = P.Model()
    m.add_field("phase", 16)
    m.register_explicit_term("CONS", P.KIND_VOLCONS, field="phase", indicator="NOPE", dt=0.1, num_nodes=16)
    m.add_equation("dphase/dt = CONS")
    with pytest.raises(P.ParseError):
        m.init()


def _hess_data(N):
    """GetData (tensorialHessian_test.go:11-41): a separable bump and its second derivatives / N^2."""
    i = np.arange(N * N)
    y = (i % N) / N
    x = (i // N) / N
    px = 16.0 * (x * x - 2.0 * x ** 3 + x ** 4)
    py = 16.0 * (y * y - 2.0 * y ** 3 + y ** 4)
    d = {"data": px * py}
    d["dx2"] = 16.0 * (2.0 - 12.0 * x + 12 * x * x) * py / (N * N)
    d["dy2"] = px * 16.0 * (2.0 - 12.0 * y + 12.0 * y * y) / (N * N)
    d["dxdy"] = 16.0 * (2.0 * x - 6.0 * x * x + 4.0 * x ** 3) * 16.0 * (2.0 * y - 6.0 * y * y + 4.0 * y ** 3) / (N * N)
    return d


@pytest.mark.parametrize("K,res,tol", [([1.0, 0.0, 0.0, 0.0], "dx2", 1e-3), ([0.0, 0.0, 0.0, 1.0], "dy2", 1e-3),
                                       ([0.0, 0.5, 0.5, 0.0], "dxdy", 1e-6)])
def test_tensorial_hessian(K, res, tol):
    """TestTensorialHessian (tensorialHessian_test.go:43-103): the multiplier times FFT(H),
    inverse-transformed, is the closed-form second derivative."""
    N = 64
    d = _hess_data(N)
    m = P.Model()
    m.add_field("myfield", N * N)
    m.register_implicit_term("HESSIAN", P.KIND_HESS, K=K)
    m.add_equation("dmyfield/dt = HESSIAN")
    m.init()
    o = E.OracleSolver(m, [N, N], {"myfield": np.zeros(N * N, dtype=complex)}, 0.1)
    mult = o.eval_user_term(P.KIND_HESS, 0)
    Hk = np.fft.fft2(d["data"].reshape(N, N)).reshape(-1)
    out = np.fft.ifft2((mult * Hk).reshape(N, N)).reshape(-1)
    assert np.abs(out.real - d[res]).max() <= tol and np.abs(out.imag).max() <= tol


def test_hessian_with_model():
    """TestHessianWithModel (tensorialHessian_test.go:105-150): HESSIAN is the one implicit term."""
    N = 16
    m = P.Model()
    m.add_field("conc1", N * N)
    m.add_field("conc2", N * N)
    m.register_implicit_term("HESSIAN", P.KIND_HESS, K=[1.0, 2.0, 2.0, 2.0])
    m.add_equation("dconc1/dt = HESSIAN")
    m.add_equation("dconc2/dt = -conc2")
    m.init()
    terms, denum = m.rhs[0]
    assert len(denum) == 1 and len(terms) == 0
