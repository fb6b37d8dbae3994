"""gopf_amd -- MI355X spectral time step of gopf behind its own interfaces.

Python view of the C ABI in ``include/gpf.h`` (ctypes; no torch types).  The
classes keep the reference's names so code reads like gopf's own tests:

  Model            pf.Model        (pf/model.go:116-260)
  Solver           pf.Solver + TimeStepper (pf/solver.go:29-134, euler.go, rk4.go)
  FourierTransform pfutil.FFTWWrapper (pfutil/fftWrap.go:8-95)
  NewVandeven      pf.NewVandeven  (pf/vandeven.go:13-40)

All compute runs in ``libgopf_amd.so`` (hand-written HIP for gfx950); loading
fails loudly when the library is missing.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence, Tuple, Union

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libgopf_amd.so")

GPF_OK, GPF_EINVAL, GPF_EPARSE, GPF_EDEVICE, GPF_ESTATE = 0, 1, 2, 3, 4
GPF_FP64, GPF_FP32 = 0, 1
GPF_EULER, GPF_RK4, GPF_IMPLICIT_EULER, GPF_SDD = 0, 1, 2, 3
GPF_IO_FLOAT64, GPF_IO_UINT8 = 0, 1
KIND_CONCRETE, KIND_SQGRAD, KIND_ELAST = 0, 1, 2


class GpfError(RuntimeError):
    """A failed library call; the reference panics in the same situations."""

    def __init__(self, code: int, msg: str):
        super().__init__(f"[gpf error {code}] {msg}")
        self.code = code


class _Options(C.Structure):
    _fields_ = [("precision", C.c_int), ("stepper", C.c_int), ("filter_order", C.c_int),
                ("fused", C.c_int), ("device", C.c_int), ("autotune", C.c_int)]


class _Reductions(C.Structure):
    _fields_ = [("sum", C.c_double), ("sum_sq", C.c_double), ("min", C.c_double), ("max", C.c_double),
                ("nonfinite", C.c_long)]


def _red_dict(r: "_Reductions") -> Dict[str, float]:
    return {"sum": r.sum, "sum_sq": r.sum_sq, "min": r.min, "max": r.max, "nonfinite": int(r.nonfinite)}


class _Term(C.Structure):
    _fields_ = [("kind", C.c_int), ("coef_re", C.c_double), ("coef_im", C.c_double), ("brick", C.c_int),
                ("has_lap", C.c_int), ("lap", C.c_int), ("user", C.c_int)]


_lib = None
_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_lp = C.POINTER(C.c_long)
_vp = C.c_void_p


def lib():
    """Load libgopf_amd.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: run `python -m gopf_amd.build` (the HIP library is required)")
        L = C.CDLL(LIB_PATH)
        _declare(L)
        _lib = L
    return _lib


def _declare(L) -> None:
    sig = {
        "gpf_last_error": (C.c_char_p, []),
        "gpf_model_create": (C.c_int, [C.POINTER(_vp)]),
        "gpf_model_destroy": (None, [_vp]),
        "gpf_model_add_field": (C.c_int, [_vp, C.c_char_p, _dp, C.c_long]),
        "gpf_model_add_scalar": (C.c_int, [_vp, C.c_char_p, C.c_double, C.c_double]),
        "gpf_model_add_equation": (C.c_int, [_vp, C.c_char_p]),
        "gpf_model_register_square_gradient": (C.c_int, [_vp, C.c_char_p, C.c_char_p, C.c_double, C.c_int]),
        "gpf_model_register_homogeneous_modulus": (C.c_int, [_vp, C.c_char_p, C.c_char_p, _dp, _dp]),
        "gpf_model_register_volume_conserving": (C.c_int, [_vp, C.c_char_p, C.c_char_p, C.c_char_p, C.c_double,
                                                          C.c_long]),
        "gpf_model_register_spectral_viscosity": (C.c_int, [_vp, C.c_char_p, C.c_double, C.c_double, C.c_int]),
        "gpf_model_register_tensorial_hessian": (C.c_int, [_vp, C.c_char_p, C.c_char_p, _dp, C.c_int]),
        "gpf_model_register_white_noise": (C.c_int, [_vp, C.c_char_p, C.c_double, C.c_ulonglong]),
        "gpf_model_register_conservative_noise": (C.c_int, [_vp, C.c_char_p, C.c_double, C.c_int, C.c_ulonglong]),
        "gpf_model_register_polynomial": (C.c_int, [_vp, C.c_char_p, C.c_int, _dp, _dp, C.c_int]),
        "gpf_model_init": (C.c_int, [_vp]),
        "gpf_model_num_bricks": (C.c_int, [_vp, _ip, _ip]),
        "gpf_model_brick_name": (C.c_int, [_vp, C.c_int, C.c_char_p, C.c_int]),
        "gpf_model_num_terms": (C.c_int, [_vp, C.c_int, _ip, _ip]),
        "gpf_model_get_term": (C.c_int, [_vp, C.c_int, C.c_int, C.c_int, C.POINTER(_Term)]),
        "gpf_options_default": (None, [C.POINTER(_Options)]),
        "gpf_solver_create": (C.c_int, [_vp, _ip, C.c_int, C.c_double, C.POINTER(_Options), C.POINTER(_vp)]),
        "gpf_solver_destroy": (None, [_vp]),
        "gpf_solver_propagate": (C.c_int, [_vp, C.c_int]),
        "gpf_solver_stepper_propagate": (C.c_int, [_vp, C.c_int]),
        "gpf_solver_prepare": (C.c_int, [_vp, C.c_int, _dp]),
        "gpf_solver_get_time": (C.c_int, [_vp, _dp]),
        "gpf_solver_get_field": (C.c_int, [_vp, C.c_int, _dp, C.c_long]),
        "gpf_solver_set_field": (C.c_int, [_vp, C.c_int, _dp, C.c_long]),
        "gpf_solver_energy": (C.c_int, [_vp, C.c_int, C.c_int, _dp, C.c_double, _dp, _dp]),
        "gpf_solver_monitor_energy": (C.c_int, [_vp, C.c_int, _dp, C.c_double]),
        "gpf_solver_reduce": (C.c_int, [_vp, C.c_int, C.POINTER(_Reductions)]),
        "gpf_reduce": (C.c_int, [_vp, C.c_long, C.c_int, C.c_int, _vp, C.POINTER(_Reductions)]),
        "gpf_reduce_noalloc": (C.c_int, [_vp, C.c_long, C.c_int, C.c_int, _vp, _vp]),
        "gpf_solver_energy_series": (C.c_int, [_vp, _dp, _dp, C.c_long, _lp]),
        "gpf_solver_synchronize": (C.c_int, [_vp]),
        "gpf_solver_is_fused": (C.c_int, [_vp]),
        "gpf_solver_enable_timing": (C.c_int, [_vp, C.c_int]),
        "gpf_solver_pass_times": (C.c_int, [_vp, C.c_int, _dp, _lp, C.c_char_p, C.c_int]),
        "gpf_solver_stream": (_vp, [_vp]),
        "gpf_solver_set_persistent": (C.c_int, [_vp, C.c_int, C.c_int]),
        "gpf_solver_persistent_steps": (C.c_int, [_vp, _lp]),
        "gpf_solver_tuning": (C.c_int, [_vp, _ip, C.c_int]),
        "gpf_solver_set_tuning": (C.c_int, [_vp, _ip, C.c_int]),
        "gpf_solver_set_filter": (C.c_int, [_vp, _dp, C.c_long]),
        "gpf_solver_set_newton_krylov": (C.c_int, [_vp, C.c_int, C.c_double, C.c_double, C.c_int, C.c_int]),
        "gpf_solver_newton_stats": (C.c_int, [_vp, _ip, _ip, _ip, _dp]),
        "gpf_solver_sdd_init": (C.c_int, [_vp, _dp, _dp, C.c_long]),
        "gpf_solver_sdd_set_orientation": (C.c_int, [_vp, _dp, C.c_long]),
        "gpf_solver_sdd_set_params": (C.c_int, [_vp, C.c_double, C.c_double, C.c_double, C.c_double, C.c_double]),
        "gpf_solver_sdd_monitor": (C.c_int, [_vp, _dp]),
        "gpf_solver_sdd_orientation": (C.c_int, [_vp, _dp, C.c_long]),
        "gpf_solver_sdd_dimer_length": (C.c_int, [_vp, C.c_double, _dp]),
        "gpf_solver_save_fields": (C.c_int, [_vp, C.c_char_p, C.c_int, C.c_int]),
        "gpf_solver_flush_io": (C.c_int, [_vp]),
        "gpf_solver_save_csv": (C.c_int, [_vp, C.c_char_p, C.c_int]),
        "gpf_save_float64": (C.c_int, [C.c_char_p, _dp, C.c_long]),
        "gpf_load_float64": (C.c_int, [C.c_char_p, _dp, C.c_long, _lp]),
        "gpf_save_csv": (C.c_int, [C.c_char_p, C.POINTER(C.c_char_p), C.c_int, _dp, C.c_long, _ip, C.c_int]),
        "gpf_write_xdmf": (C.c_int, [C.c_char_p, C.POINTER(C.c_char_p), C.c_int, C.c_char_p, C.c_int, _ip, C.c_int]),
        "gpf_create_xdmf": (C.c_int, [C.POINTER(C.c_char_p), C.c_int, C.c_char_p, C.c_int, _ip, C.c_int,
                                      C.c_char_p, C.c_char_p, C.c_long, _lp]),
        "gpf_solver_adjoint": (C.c_int, [_vp, C.c_int, C.c_int, _dp, C.c_long, _dp, _dp, C.c_int,
                                         C.POINTER(C.c_char_p), _dp]),
        "gpf_device_copy_bandwidth": (C.c_int, [C.c_long, C.c_int, C.c_int, _dp]),
        "gpf_device_bandwidth": (C.c_int, [C.c_long, C.c_int, C.c_int, C.c_int, _dp]),
        "gpf_batch_create": (C.c_int, [_vp, C.c_int, _ip, _ip, C.c_double, C.POINTER(_Options), C.POINTER(_vp)]),
        "gpf_batch_destroy": (None, [_vp]),
        "gpf_batch_num_sims": (C.c_int, [_vp]),
        "gpf_batch_set_scalar": (C.c_int, [_vp, C.c_int, C.c_char_p, C.c_double]),
        "gpf_batch_set_field": (C.c_int, [_vp, C.c_int, _dp, C.c_long]),
        "gpf_batch_get_field": (C.c_int, [_vp, C.c_int, _dp, C.c_long]),
        "gpf_batch_propagate": (C.c_int, [_vp, C.c_int]),
        "gpf_batch_synchronize": (C.c_int, [_vp]),
        "gpf_batch_tuning": (C.c_int, [_vp, _ip, C.c_int]),
        "gpf_batch_set_layout": (C.c_int, [_vp, _ip, C.c_int]),
        "gpf_fft_create": (C.c_int, [_ip, C.c_int, C.c_int, C.c_int, C.POINTER(_vp)]),
        "gpf_fft_destroy": (None, [_vp]),
        "gpf_fft_forward": (C.c_int, [_vp, _dp, C.c_long]),
        "gpf_fft_inverse": (C.c_int, [_vp, _dp, C.c_long]),
        "gpf_fft_freq": (C.c_int, [_ip, C.c_int, C.c_long, _dp]),
        "gpf_fft_conjugate_node": (C.c_long, [_ip, C.c_int, C.c_long]),
        "gpf_vandeven_table": (C.c_int, [C.c_int, _dp]),
        "gpf_vandeven_eval": (C.c_double, [_dp, C.c_double]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args


def _check(rc: int) -> None:
    if rc != 0:
        raise GpfError(rc, lib().gpf_last_error().decode())


def _ints(v: Sequence[int]):
    arr = (C.c_int * len(v))(*[int(x) for x in v])
    return arr


def _cbuf(data, n: Optional[int] = None) -> np.ndarray:
    a = np.ascontiguousarray(np.asarray(data, dtype=np.complex128).reshape(-1))
    if n is not None and a.size != n:
        raise ValueError("length mismatch")
    return a


def _dptr(a: np.ndarray):
    return a.ctypes.data_as(_dp)


# ---------------------------------------------------------------------------
# registered terms
# ---------------------------------------------------------------------------

@dataclass
class VolumeConservingLP:
    """pf.VolumeConservingLP (volumeConserving.go:8-61): Multiplier * Indicator, with the
    multiplier tuned after every step so that the sum of Field stays constant."""
    Field: str
    Indicator: str
    Dt: float
    NumNodes: int


def NewVolumeConservingLP(fieldName: str, indicator: str, dt: float, numNodes: int) -> VolumeConservingLP:
    return VolumeConservingLP(fieldName, indicator, dt, numNodes)


@dataclass
class SpectralViscosity:
    """pf.SpectralViscosity (spectralViscosity.go:10-58): the implicit multiplier
    -Eps Q(|f|) |f|^Power."""
    Eps: float
    DissipationThreshold: float
    Power: int


@dataclass
class WhiteNoise:
    """pf.WhiteNoise (noise.go:10-22): Gaussian noise with <Y Y'> = 2 Strength delta delta.
    Register it with Model.RegisterFunction(name, noise); Seed keys the device generator."""
    Strength: float
    Seed: int = 0


@dataclass
class ConservativeNoise:
    """pf.ConservativeNoise (noise.go:24-100): the discrete divergence of Dim Gaussian current
    fields.  Register it with Model.RegisterExplicitTerm(name, noise)."""
    Strength: float
    Dim: int
    Seed: int = 0


def NewConservativeNoise(strength: float, dim: int, seed: int = 0) -> ConservativeNoise:
    return ConservativeNoise(strength, dim, seed)


@dataclass
class TensorialHessian:
    """pf.TensorialHessian (tensorialHessian.go:19-66): the implicit multiplier of
    sum_ij K_ij d^2/dx_i dx_j; K row-major with 4 (2D) or 9 (3D) entries."""
    Field: str
    K: Sequence[float]


@dataclass
class SquaredGradient:
    """pf.SquaredGradient (squareGradientTerm.go:15-37): Factor * |grad Field|^2."""
    Field: str
    Factor: float = 1.0


@dataclass
class HomogeneousModulusLinElast:
    """pf.HomogeneousModulusLinElast (homoLinElast.go:33-43, 137-150)."""
    FieldName: str
    MatProp: np.ndarray          # 81 values C_ijkl
    Misfit: np.ndarray           # 3x3


def CubicMaterial(c11: float, c12: float, c44: float) -> np.ndarray:
    """elasticity.CubicMaterial (rank4.go:113-128)."""
    t = np.zeros((3, 3, 3, 3))
    for i in range(3):
        t[i, i, i, i] = c11
    for i in range(3):
        for j in range(i + 1, 3):
            t[i, i, j, j] = c12
            t[j, j, i, i] = c12
            t[i, j, i, j] = c44
            t[j, i, j, i] = c44
    return t.reshape(81)


def Isotropic(bulk: float, poisson: float) -> np.ndarray:
    """elasticity.Isotropic (rank4.go:84-110)."""
    shear = 3.0 * bulk * (1.0 - 2.0 * poisson) / (2.0 * (1.0 + poisson))
    t = np.zeros((3, 3, 3, 3))
    for i in range(3):
        for j in range(3):
            for k in range(3):
                for l in range(3):
                    v = 0.0
                    if i == j and k == l:
                        v += bulk - 2.0 * shear / 3.0
                    if i == k and j == l:
                        v += shear
                    if i == l and j == k:
                        v += shear
                    t[i, j, k, l] = v
    return t.reshape(81)


# ---------------------------------------------------------------------------
# pf.Model
# ---------------------------------------------------------------------------

class Model:
    """pf.Model held by the library (model.go:116-260)."""

    def __init__(self) -> None:
        h = _vp()
        _check(lib().gpf_model_create(C.byref(h)))
        self._h = h
        self._fields: List[str] = []

    def __del__(self):
        try:
            if getattr(self, "_h", None):
                lib().gpf_model_destroy(self._h)
                self._h = None
        except Exception:
            pass

    # -- building --------------------------------------------------------------
    def AddField(self, name: str, data=None, n: Optional[int] = None) -> None:
        """NewField + AddField (model.go:44-57, 141-144)."""
        if data is None:
            if n is None:
                raise ValueError("need data or n")
            _check(lib().gpf_model_add_field(self._h, name.encode(), None, n))
        else:
            buf = _cbuf(data)
            _check(lib().gpf_model_add_field(self._h, name.encode(), _dptr(buf), buf.size))
        self._fields.append(name)

    def AddScalar(self, name: str, value: complex) -> None:
        v = complex(value)
        _check(lib().gpf_model_add_scalar(self._h, name.encode(), v.real, v.imag))

    def AddEquation(self, eq: str) -> None:
        _check(lib().gpf_model_add_equation(self._h, eq.encode()))

    def RegisterExplicitTerm(self, name: str, term) -> None:
        self._register(name, term, implicit=False)

    def RegisterImplicitTerm(self, name: str, term) -> None:
        self._register(name, term, implicit=True)

    def RegisterFunction(self, name: str, fn) -> None:
        """Model.RegisterFunction (model.go:400-412) for the device-evaluable functions: a
        WhiteNoise (its Generate), or use RegisterPolynomial for polynomials of the fields."""
        if isinstance(fn, WhiteNoise):
            _check(lib().gpf_model_register_white_noise(self._h, name.encode(), float(fn.Strength),
                                                        int(fn.Seed) & 0xFFFFFFFFFFFFFFFF))
            return
        raise TypeError("RegisterFunction takes a WhiteNoise here; register polynomials with RegisterPolynomial")

    def _register(self, name: str, term, implicit: bool) -> None:
        if isinstance(term, ConservativeNoise):
            if implicit:
                raise GpfError(1, "ConservativeNoise is an explicit term")
            _check(lib().gpf_model_register_conservative_noise(self._h, name.encode(), float(term.Strength),
                                                               int(term.Dim), int(term.Seed) & 0xFFFFFFFFFFFFFFFF))
            return
        if isinstance(term, VolumeConservingLP):
            if implicit:
                raise GpfError(1, "VolumeConservingLP is an explicit term")
            _check(lib().gpf_model_register_volume_conserving(self._h, name.encode(), term.Field.encode(),
                                                              term.Indicator.encode(), float(term.Dt),
                                                              int(term.NumNodes)))
            return
        if isinstance(term, TensorialHessian):
            if not implicit:
                raise GpfError(1, "TensorialHessian runs as an implicit term")
            K = np.ascontiguousarray(np.asarray(term.K, dtype=np.float64).reshape(-1))
            _check(lib().gpf_model_register_tensorial_hessian(self._h, name.encode(), str(term.Field or "").encode(),
                                                              _dptr(K), K.size))
            return
        if isinstance(term, SpectralViscosity):
            if not implicit:
                raise GpfError(1, "SpectralViscosity runs as an implicit term")
            _check(lib().gpf_model_register_spectral_viscosity(self._h, name.encode(), float(term.Eps),
                                                               float(term.DissipationThreshold), int(term.Power)))
            return
        if isinstance(term, SquaredGradient):
            _check(lib().gpf_model_register_square_gradient(self._h, name.encode(), term.Field.encode(),
                                                            float(term.Factor), 1 if implicit else 0))
        elif isinstance(term, HomogeneousModulusLinElast):
            if implicit:
                raise GpfError(1, "HomogeneousModulusLinElast is an explicit term")
            Cm = np.ascontiguousarray(np.asarray(term.MatProp, dtype=np.float64).reshape(81))
            mis = np.ascontiguousarray(np.asarray(term.Misfit, dtype=np.float64).reshape(9))
            _check(lib().gpf_model_register_homogeneous_modulus(self._h, name.encode(), term.FieldName.encode(),
                                                                _dptr(Cm), _dptr(mis)))
        else:
            raise TypeError(f"term type {type(term).__name__} cannot run on the device")

    def RegisterPolynomial(self, name: str, terms: List[Tuple[float, Dict[str, float]]]) -> None:
        """RegisterFunction (model.go:400-412) for a polynomial of the fields' real parts."""
        nf = len(self._fields)
        coefs = np.array([float(c) for c, _ in terms], dtype=np.float64)
        pw = np.zeros((len(terms), nf), dtype=np.float64)
        for t, (_, powers) in enumerate(terms):
            for f, p in powers.items():
                pw[t, self._fields.index(f)] = p
        pw = np.ascontiguousarray(pw.reshape(-1))
        _check(lib().gpf_model_register_polynomial(self._h, name.encode(), len(terms), _dptr(coefs),
                                                   _dptr(pw) if pw.size else None, nf))

    def Init(self) -> None:
        _check(lib().gpf_model_init(self._h))

    # -- introspection -----------------------------------------------------------
    def AllFieldNames(self) -> List[str]:
        nf, nd = C.c_int(), C.c_int()
        _check(lib().gpf_model_num_bricks(self._h, C.byref(nf), C.byref(nd)))
        out = []
        buf = C.create_string_buffer(256)
        for i in range(nf.value + nd.value):
            _check(lib().gpf_model_brick_name(self._h, i, buf, 256))
            out.append(buf.value.decode())
        return out

    def RHS(self, eq: int) -> Tuple[List[_Term], List[_Term]]:
        nt, nd = C.c_int(), C.c_int()
        _check(lib().gpf_model_num_terms(self._h, eq, C.byref(nt), C.byref(nd)))
        res = ([], [])
        for den, count in ((0, nt.value), (1, nd.value)):
            for i in range(count):
                t = _Term()
                _check(lib().gpf_model_get_term(self._h, eq, den, i, C.byref(t)))
                res[den].append(t)
        return res

    def NumFields(self) -> int:
        return len(self._fields)

    def FieldIndex(self, name: str) -> int:
        return self._fields.index(name)

    def FieldNames(self) -> List[str]:
        """The names of Model.Fields in order."""
        return list(self._fields)


# ---------------------------------------------------------------------------
# pf.Solver
# ---------------------------------------------------------------------------

class Solver:
    """pf.Solver with a device TimeStepper (solver.go:29-134)."""

    def __init__(self, model: Model, domainSize: Sequence[int], dt: float, stepper: str = "euler",
                 precision: str = "fp64", filter_order: int = 0, fused: bool = True, device: int = 0,
                 autotune: bool = False):
        o = _Options()
        lib().gpf_options_default(C.byref(o))
        o.precision = GPF_FP32 if precision in ("fp32", "float32") else GPF_FP64
        steppers = {"euler": GPF_EULER, "rk4": GPF_RK4, "implicitEuler": GPF_IMPLICIT_EULER, "sdd": GPF_SDD}
        if stepper not in steppers:
            raise GpfError(1, "Unknown stepper scheme")
        o.stepper = steppers[stepper]
        o.filter_order = int(filter_order)
        o.fused = 1 if fused else 0
        o.device = int(device)
        o.autotune = 1 if autotune else 0
        h = _vp()
        self.dims = [int(d) for d in domainSize]
        self.n = int(np.prod(self.dims))
        _check(lib().gpf_solver_create(model._h, _ints(self.dims), len(self.dims), float(dt), C.byref(o), C.byref(h)))
        self._h = h
        self.model = model
        self.Dt = dt

    def __del__(self):
        try:
            if getattr(self, "_h", None):
                lib().gpf_solver_destroy(self._h)
                self._h = None
        except Exception:
            pass

    def Propagate(self, nsteps: int) -> None:
        _check(lib().gpf_solver_propagate(self._h, int(nsteps)))

    def StepperPropagate(self, nsteps: int) -> None:
        _check(lib().gpf_solver_stepper_propagate(self._h, int(nsteps)))

    def Prepare(self, nsteps: int = 0) -> float:
        """Autotuner, monitor buffers and the replayed step graph ahead of the first step
        (no reference counterpart); the fields and the time do not change.  Returns the host
        milliseconds it took."""
        ms = C.c_double()
        _check(lib().gpf_solver_prepare(self._h, int(nsteps), C.byref(ms)))
        return ms.value

    def Solve(self, nepochs: int, nsteps: int, callbacks=()) -> None:
        """Solver.Solve (solver.go:106-120) with host callbacks cb(solver, epoch); files of
        asynchronous snapshot callbacks (Float64IO, Uint8IO) are written when it returns."""
        for epoch in range(nepochs):
            self.Propagate(nsteps)
            for cb in callbacks:
                cb(self, epoch)
        self.FlushIO()

    def FlushIO(self) -> None:
        """Waits until every queued field snapshot is on disk."""
        _check(lib().gpf_solver_flush_io(self._h))

    def GetTime(self) -> float:
        t = C.c_double()
        _check(lib().gpf_solver_get_time(self._h, C.byref(t)))
        return t.value

    def _idx(self, f: Union[int, str]) -> int:
        return self.model.FieldIndex(f) if isinstance(f, str) else int(f)

    def Field(self, f: Union[int, str]) -> np.ndarray:
        out = np.zeros(self.n, dtype=np.complex128)
        _check(lib().gpf_solver_get_field(self._h, self._idx(f), _dptr(out), self.n))
        return out

    def SetField(self, f: Union[int, str], data) -> None:
        buf = _cbuf(data, self.n)
        _check(lib().gpf_solver_set_field(self._h, self._idx(f), _dptr(buf), self.n))

    def Energy(self, f: Union[int, str], poly: Sequence[float], kappa: float) -> Tuple[float, float]:
        p = np.ascontiguousarray(np.asarray(poly, dtype=np.float64))
        e, m = C.c_double(), C.c_double()
        _check(lib().gpf_solver_energy(self._h, self._idx(f), p.size, _dptr(p), float(kappa), C.byref(e), C.byref(m)))
        return e.value, m.value

    def Reduce(self, f: Union[int, str]) -> Dict[str, float]:
        """FP64 sum, sum of squares, min, max and the count of non-finite values of a field
        (the device form of a driver's NaN scan, examples/solverStability/main.go:58-67)."""
        r = _Reductions()
        _check(lib().gpf_solver_reduce(self._h, self._idx(f), C.byref(r)))
        return _red_dict(r)

    def MonitorEnergy(self, poly: Sequence[float], kappa: float) -> None:
        p = np.ascontiguousarray(np.asarray(poly, dtype=np.float64))
        _check(lib().gpf_solver_monitor_energy(self._h, p.size, _dptr(p), float(kappa)))

    def EnergySeries(self) -> Tuple[np.ndarray, np.ndarray]:
        n = C.c_long()
        _check(lib().gpf_solver_energy_series(self._h, None, None, 0, C.byref(n)))
        e = np.zeros(n.value)
        m = np.zeros(n.value)
        _check(lib().gpf_solver_energy_series(self._h, _dptr(e), _dptr(m), n.value, C.byref(n)))
        return e, m

    def Synchronize(self) -> None:
        _check(lib().gpf_solver_synchronize(self._h))

    @property
    def fused(self) -> bool:
        return bool(lib().gpf_solver_is_fused(self._h))

    def EnableTiming(self, on: bool = True) -> None:
        _check(lib().gpf_solver_enable_timing(self._h, 1 if on else 0))

    def PassTimes(self) -> Dict[str, Tuple[float, int]]:
        ms = np.zeros(16)
        cnt = np.zeros(16, dtype=np.int64)
        names = C.create_string_buffer(16 * 64)
        np_ = lib().gpf_solver_pass_times(self._h, 16, _dptr(ms), cnt.ctypes.data_as(_lp), names, 64)
        if np_ < 0:
            _check(-np_)
        out = {}
        for i in range(np_):
            nm = names.raw[i * 64:(i + 1) * 64].split(b"\0", 1)[0].decode()
            out[nm] = (float(ms[i]), int(cnt[i]))
        return out

    def Tuning(self) -> Dict[str, int]:
        v = (C.c_int * 10)()
        lib().gpf_solver_tuning(self._h, v, 10)
        return {"rows_fused": v[0], "pencils_fwd": v[1], "pencil_update": v[2], "pencils_inv": v[3],
                "update_axis": "zy"[2 - v[4]] if v[4] in (1, 2) else "y", "rows_schedule": v[5],
                "pencils_fwd_lines_per_thread": v[6], "pencils_inv_lines_per_thread": v[7],
                "pencils_fwd_schedule": v[8], "pencils_inv_schedule": v[9]}

    TUNING_KEYS = ("rows_fused", "pencils_fwd", "pencil_update", "pencils_inv", "update_axis", "rows_schedule",
                   "pencils_fwd_lines_per_thread", "pencils_inv_lines_per_thread", "pencils_fwd_schedule",
                   "pencils_inv_schedule")

    def SetTuning(self, t: Dict[str, object]) -> None:
        """Applies a launch layout returned by Tuning() instead of the autotuner."""
        v = [int({"z": 2, "y": 1}[t[k]]) if k == "update_axis" else int(t[k]) for k in self.TUNING_KEYS]
        _check(lib().gpf_solver_set_tuning(self._h, _ints(v), len(v)))

    def Stream(self) -> int:
        return int(lib().gpf_solver_stream(self._h) or 0)

    def SetPersistent(self, mode: int = -1, threads: int = 0) -> None:
        """Persistent step of square 2D grids (side 32-256): -1 auto (sides up to 128), 0 off, 1 on,
        2 on with the agent-fenced barrier; threads per block (0: 64)."""
        _check(lib().gpf_solver_set_persistent(self._h, int(mode), int(threads)))

    def PersistentSteps(self) -> int:
        n = C.c_long(0)
        _check(lib().gpf_solver_persistent_steps(self._h, C.byref(n)))
        return int(n.value)

    def SetFilter(self, filt) -> None:
        """TimeStepper.SetFilter (pf/euler.go:62-64) with a Vandeven filter (its table), or None."""
        if filt is None:
            _check(lib().gpf_solver_set_filter(self._h, None, 0))
            return
        tab = np.ascontiguousarray(np.asarray(filt.Data, dtype=np.float64))
        _check(lib().gpf_solver_set_filter(self._h, _dptr(tab), tab.size))

    def SetNewtonKrylov(self, Maxiter: int = 50, StepSize: float = 1e-3, Tol: float = 1e-7, Stencil: int = 6,
                        Restart: int = 30) -> None:
        """ImplicitEuler.NonlinSolver = &nonlin.NewtonKrylov{...} (pf/implicitEuler.go:26-28, 221-229);
        Restart is the GMRES restart length."""
        _check(lib().gpf_solver_set_newton_krylov(self._h, int(Maxiter), float(StepSize), float(Tol), int(Stencil),
                                                  int(Restart)))

    def NewtonStats(self) -> Dict[str, float]:
        """Newton and GMRES iterations, convergence and max-norm residual of the last ImplicitEuler step."""
        a, b, c = C.c_int(), C.c_int(), C.c_int()
        r = C.c_double()
        _check(lib().gpf_solver_newton_stats(self._h, C.byref(a), C.byref(b), C.byref(c), C.byref(r)))
        return {"newton_iters": a.value, "krylov_iters": b.value, "converged": bool(c.value), "residual": r.value}

    # -- pf.SDD (sdd.go) ---------------------------------------------------------
    def _flat(self, fields) -> np.ndarray:
        arrs = [np.asarray(f).real.astype(np.float64).reshape(-1) for f in fields]
        return np.ascontiguousarray(np.concatenate(arrs))

    def SDDInit(self, init, final) -> None:
        """SDD.Init (sdd.go:382-402): init and final are lists with one array per field."""
        a, b = self._flat(init), self._flat(final)
        _check(lib().gpf_solver_sdd_init(self._h, _dptr(a), _dptr(b), a.size))

    def SDDSetInitialOrientation(self, orient) -> None:
        """SDD.SetInitialOrientation (sdd.go:404-420)."""
        v = np.ascontiguousarray(np.asarray(orient, dtype=np.float64).reshape(-1))
        _check(lib().gpf_solver_sdd_set_orientation(self._h, _dptr(v), v.size))

    def SDDSetParams(self, Alpha: float = 0.5, Orientation: float = 1.0, DimerLength: float = 1.0,
                     InitDimerLength: float = -1.0, MinDimerLength: float = 0.0) -> None:
        """SDD.Alpha, TimeConstants{Orientation, DimerLength}, InitDimerLength (negative keeps
        the value from SDDInit), MinDimerLength (sdd.go:84-130)."""
        _check(lib().gpf_solver_sdd_set_params(self._h, float(Alpha), float(Orientation), float(DimerLength),
                                               float(InitDimerLength), float(MinDimerLength)))

    def SDDMonitor(self) -> Dict[str, float]:
        out = np.zeros(6)
        _check(lib().gpf_solver_sdd_monitor(self._h, _dptr(out)))
        keys = ["MaxForce", "ForcePowerSpectrum", "MaxTorque", "FieldNorm", "FieldNormChange", "TorqueDotOrientation"]
        return dict(zip(keys, out.tolist()))

    def SDDOrientation(self) -> np.ndarray:
        out = np.zeros(self.n * self.model.NumFields())
        _check(lib().gpf_solver_sdd_orientation(self._h, _dptr(out), out.size))
        return out

    def SDDDimerLength(self, t: float) -> float:
        v = C.c_double()
        _check(lib().gpf_solver_sdd_dimer_length(self._h, float(t), C.byref(v)))
        return v.value

    def Adjoint(self, nsteps: int, target, scalars: Sequence[str] = (), checkpoint_every: int = 100
                ) -> Tuple[float, np.ndarray, Dict[str, float]]:
        """Reverse-mode gradient of J = sum_x (c_T - target)^2 over nsteps Euler steps.

        Returns (J, dJ/dc_0, {scalar: dJ/dscalar}).  No reference counterpart: it
        differentiates Euler.Step (pf/euler.go:16-47) on the fused single-field class,
        recomputing states from checkpoints taken every checkpoint_every steps.  The
        solver is left at c_T, as after StepperPropagate(nsteps).
        """
        tgt = np.ascontiguousarray(np.asarray(target).real.astype(np.float64).reshape(-1))
        if tgt.size != self.n:
            raise ValueError("length mismatch")
        names = [str(x) for x in scalars]
        cn = (C.c_char_p * max(len(names), 1))(*[x.encode() for x in names])
        gs = np.zeros(max(len(names), 1))
        gf = np.zeros(self.n)
        loss = C.c_double()
        _check(lib().gpf_solver_adjoint(self._h, int(nsteps), int(checkpoint_every), _dptr(tgt), self.n,
                                        C.byref(loss), _dptr(gf), len(names), cn, _dptr(gs)))
        return loss.value, gf, {k: float(gs[i]) for i, k in enumerate(names)}


# ---------------------------------------------------------------------------
# pfutil.FFTWWrapper
# ---------------------------------------------------------------------------

class FourierTransform:
    """FourierTransform implemented on the device (fftWrap.go:8-95)."""

    def __init__(self, dims: Sequence[int], precision: str = "fp64", device: int = 0):
        self.Dimensions = [int(d) for d in dims]
        self.n = int(np.prod(self.Dimensions))
        h = _vp()
        prec = GPF_FP32 if precision in ("fp32", "float32") else GPF_FP64
        _check(lib().gpf_fft_create(_ints(self.Dimensions), len(self.Dimensions), prec, int(device), C.byref(h)))
        self._h = h

    def __del__(self):
        try:
            if getattr(self, "_h", None):
                lib().gpf_fft_destroy(self._h)
                self._h = None
        except Exception:
            pass

    def FFT(self, data: np.ndarray) -> np.ndarray:
        """In place on a complex128 numpy array, like FFTWWrapper.FFT; also returns it."""
        if data.dtype != np.complex128 or not data.flags.c_contiguous:
            raise TypeError("data must be a contiguous complex128 array")
        _check(lib().gpf_fft_forward(self._h, data.ctypes.data_as(_dp), data.size))
        return data

    def IFFT(self, data: np.ndarray) -> np.ndarray:
        if data.dtype != np.complex128 or not data.flags.c_contiguous:
            raise TypeError("data must be a contiguous complex128 array")
        _check(lib().gpf_fft_inverse(self._h, data.ctypes.data_as(_dp), data.size))
        return data

    def Freq(self, i: int) -> List[float]:
        return Freq(self.Dimensions, i)

    def ConjugateNode(self, i: int) -> int:
        return int(lib().gpf_fft_conjugate_node(_ints(self.Dimensions), len(self.Dimensions), int(i)))


class Batch:
    """Many independent single-field sims of uneven 2D sizes sharing one model.

    The reference runs such a sweep as one pf.Solver per sim (pf/solver.go:40-63);
    here every size class advances in one launch per pass, steps replay a HIP graph.
    """

    def __init__(self, model: Model, classes: Sequence[Tuple[Sequence[int], int]], dt: float,
                 precision: str = "fp32", autotune: bool = True, device: int = 0):
        o = _Options()
        lib().gpf_options_default(C.byref(o))
        o.precision = GPF_FP32 if precision in ("fp32", "float32") else GPF_FP64
        o.autotune = 1 if autotune else 0
        o.device = int(device)
        shapes = []
        counts = []
        self.shapes = []
        for shape, count in classes:
            shapes += [int(shape[0]), int(shape[1])]
            counts.append(int(count))
            self.shapes += [tuple(shape)] * int(count)
        h = _vp()
        _check(lib().gpf_batch_create(model._h, len(counts), _ints(shapes), _ints(counts), float(dt), C.byref(o),
                                      C.byref(h)))
        self._h = h

    def __del__(self):
        try:
            if getattr(self, "_h", None):
                lib().gpf_batch_destroy(self._h)
                self._h = None
        except Exception:
            pass

    def __len__(self) -> int:
        return int(lib().gpf_batch_num_sims(self._h))

    def SetScalar(self, sim: int, name: str, value: float) -> None:
        _check(lib().gpf_batch_set_scalar(self._h, int(sim), name.encode(), float(value)))

    def SetField(self, sim: int, data) -> None:
        n = int(np.prod(self.shapes[sim]))
        buf = _cbuf(data, n)
        _check(lib().gpf_batch_set_field(self._h, int(sim), _dptr(buf), n))

    def Field(self, sim: int) -> np.ndarray:
        n = int(np.prod(self.shapes[sim]))
        out = np.zeros(n, dtype=np.complex128)
        _check(lib().gpf_batch_get_field(self._h, int(sim), _dptr(out), n))
        return out

    def Propagate(self, nsteps: int) -> None:
        _check(lib().gpf_batch_propagate(self._h, int(nsteps)))

    def Synchronize(self) -> None:
        _check(lib().gpf_batch_synchronize(self._h))

    def SetLayout(self, layout: Dict[str, object]) -> None:
        """Applies a layout returned by Tuning() instead of the tuner."""
        v = []
        for c in layout["classes"]:
            v += [int(c["rows_tw"]), int(c["update_tw"]), int(c["rows_schedule"])]
        v.append(int(layout["grouped_threads"]))
        _check(lib().gpf_batch_set_layout(self._h, _ints(v), len(v)))

    def Tuning(self) -> Dict[str, object]:
        """The launch layout: per class the row and update lines per block and the row
        transform schedule; grouped_threads > 0 when one launch per pass advances every class."""
        v = (C.c_int * 64)()
        k = lib().gpf_batch_tuning(self._h, v, 64)
        vals = [v[i] for i in range(min(k, 64))]
        ncls = (len(vals) - 1) // 3
        return {"classes": [{"rows_tw": vals[3 * c], "update_tw": vals[3 * c + 1], "rows_schedule": vals[3 * c + 2]}
                            for c in range(ncls)],
                "grouped_threads": vals[-1] if vals else 0}


def Reduce(data_ptr: int, n: int, precision: str = "fp64", device: int = 0, stream: int = 0) -> Dict[str, float]:
    """gpf_reduce over n device values at data_ptr (float for "fp32", double for "fp64")."""
    r = _Reductions()
    prec = GPF_FP32 if precision in ("fp32", "float32") else GPF_FP64
    _check(lib().gpf_reduce(_vp(int(data_ptr)), int(n), prec, int(device), _vp(int(stream)) if stream else None,
                            C.byref(r)))
    return _red_dict(r)


def ReduceNoAlloc(data_ptr: int, n: int, precision: str, device: int, stream: int, result_ptr: int) -> None:
    """gpf_reduce_noalloc: the result's five doubles are written to device memory at
    result_ptr on `stream`, which is not synchronised."""
    prec = GPF_FP32 if precision in ("fp32", "float32") else GPF_FP64
    _check(lib().gpf_reduce_noalloc(_vp(int(data_ptr)), int(n), prec, int(device),
                                    _vp(int(stream)) if stream else None, _vp(int(result_ptr))))


def CopyBandwidth(nbytes: int, reps: int = 20, device: int = 0) -> float:
    """Measured device copy bandwidth in GB/s (read + write bytes per second)."""
    ms = C.c_double()
    _check(lib().gpf_device_copy_bandwidth(int(nbytes), int(reps), int(device), C.byref(ms)))
    return 2.0 * nbytes / (ms.value * 1e-3) / 1e9


def DeviceBandwidth(nbytes: int, kind: str = "copy", reps: int = 10, device: int = 0) -> float:
    """Measured device bandwidth in GB/s: "copy" (nontemporal, read + write bytes),
    "read" or "write"."""
    k = {"copy": 0, "read": 1, "write": 2}[kind]
    ms = C.c_double()
    _check(lib().gpf_device_bandwidth(int(nbytes), int(reps), k, int(device), C.byref(ms)))
    unit = 4 * 256 * 4 * 4
    moved = (int(nbytes) // unit) * unit * (2 if k == 0 else 1)
    return moved / (ms.value * 1e-3) / 1e9


def Freq(dims: Sequence[int], i: int) -> List[float]:
    """FFTWWrapper.Freq (fftWrap.go:57-74) without constructing a plan."""
    out = (C.c_double * 3)()
    _check(lib().gpf_fft_freq(_ints(dims), len(dims), int(i), out))
    return [out[a] for a in range(len(dims))]


def ConjugateNode(dims: Sequence[int], i: int) -> int:
    return int(lib().gpf_fft_conjugate_node(_ints(dims), len(dims), int(i)))


class Vandeven:
    """pf.Vandeven (vandeven.go:8-40)."""

    def __init__(self, order: int):
        self.Data = np.zeros(1000)
        _check(lib().gpf_vandeven_table(int(order), _dptr(self.Data)))
        self.order = order

    def Eval(self, x: float) -> float:
        return float(lib().gpf_vandeven_eval(_dptr(self.Data), float(x)))


def NewVandeven(order: int) -> Vandeven:
    return Vandeven(order)


def cahn_hilliard_model(n: int, M: float = 1.0, kappa: float = 1.0, data=None) -> Model:
    """The Cahn-Hilliard model with f = c^2 (1 - c)^2 as gopf equation strings.

    f'(c) = 2c - 6c^2 + 4c^3; dc/dt = M LAP f'(c) - M kappa LAP^2 c, written in the
    form of examples/cahnHilliard/main.go:33 with scalars m2 = 2M, m6 = 6M, m4 = 4M, mk = M kappa.
    """
    m = Model()
    m.AddField("conc", data if data is not None else None, n=n)
    for name, v in (("m2", 2 * M), ("m6", 6 * M), ("m4", 4 * M), ("mk", M * kappa)):
        m.AddScalar(name, v)
    m.AddEquation("dconc/dt = m2*LAP conc - m6*LAP conc^2 + m4*LAP conc^3 - mk*LAP^2 conc")
    m.Init()
    return m


CH_BULK_POLY = [0.0, 0.0, 1.0, -2.0, 1.0]   # f(c) = c^2 - 2 c^3 + c^4 = c^2 (1 - c)^2


from .fileio import (BatchFieldDB, CsvIO, FieldDB, Float64IO, LoadCSV, LoadFloat64, NewFloat64IO, NewUint8IO,  # noqa: E402
                     SaveCsv, SaveFloat64, Uint8IO, CreateXDMF, WriteXDMF)
