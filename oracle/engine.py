"""ctypes glue for the FP64 C oracle (TEST INFRASTRUCTURE ONLY).

Only tests, ``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline`` leg
import this module; the product library never loads the oracle.

``OracleSolver`` couples the parser restatement (``oracle/parser.py``) with the
C engine (``oracle/pf_oracle.c``) and exposes the reference's stepping entry
points: ``TimeStepper.Step`` (``step``), ``Euler/RK4.Propagate``
(``stepper_propagate``) and ``Solver.Propagate`` (``solver_propagate``).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from typing import Dict, List, Optional, Sequence

import numpy as np

from . import parser as P

_HERE = os.path.dirname(os.path.abspath(__file__))
# GPF_ORACLE_LIB selects another build of the same source (the sanitizer test's)
LIB_PATH = os.environ.get("GPF_ORACLE_LIB", os.path.join(_HERE, "liboracle.so"))
_lib = None


def build(force: bool = False, path: Optional[str] = None, sanitize: bool = False) -> str:
    """Compile oracle/pf_oracle.c into oracle/liboracle.so (or `path`) with gcc; sanitize
    adds AddressSanitizer and UndefinedBehaviorSanitizer (UB aborts)."""
    src = os.path.join(_HERE, "pf_oracle.c")
    out = path or LIB_PATH
    if force or not os.path.exists(out) or os.path.getmtime(out) < os.path.getmtime(src):
        san = ["-fsanitize=address,undefined", "-fno-sanitize-recover=undefined", "-fno-omit-frame-pointer", "-g"]
        cmd = ["gcc", "-O2", "-ffp-contract=off", "-fno-fast-math", "-fPIC", "-shared"] + (san if sanitize else []) + \
              ["-o", out, src, "-lm"]
        subprocess.check_call(cmd)
    return out


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = C.CDLL(LIB_PATH)
        _declare(_lib)
    return _lib


_ip = C.POINTER(C.c_int)
_dp = C.POINTER(C.c_double)


def _declare(L) -> None:
    L.or_fft.argtypes = [C.c_int, _ip, _dp, C.c_int]
    L.or_freq.argtypes = [C.c_int, _ip, C.c_long, _dp]
    L.or_conjugate_node.argtypes = [C.c_int, _ip, C.c_long]
    L.or_conjugate_node.restype = C.c_long
    L.or_vandeven_table.argtypes = [C.c_int, _dp]
    L.or_vandeven_eval.argtypes = [_dp, C.c_double]
    L.or_vandeven_eval.restype = C.c_double
    L.or_cubic_material.argtypes = [C.c_double, C.c_double, C.c_double, _dp]
    L.or_isotropic.argtypes = [C.c_double, C.c_double, _dp]
    L.or_energy_density.argtypes = [_dp, _dp]
    L.or_energy_density.restype = C.c_double
    L.or_model_create.restype = C.c_void_p
    L.or_model_create.argtypes = [
        C.c_int, _ip, C.c_int, C.c_int,
        _ip, _ip, _dp, _ip, _ip, _dp,
        C.c_int, _ip, _ip, _dp, _ip, _ip, _ip, _ip,
        _ip, _ip, _dp, _ip, _ip, _ip, _ip,
        C.c_int, _ip, _dp,
        C.c_int, _ip, _dp, _dp,
        C.c_int, C.c_double, C.c_int]
    for name in ("or_model_destroy", "or_sync_derived", "or_step"):
        getattr(L, name).argtypes = [C.c_void_p]
    L.or_set_brick.argtypes = [C.c_void_p, C.c_int, _dp]
    L.or_get_brick.argtypes = [C.c_void_p, C.c_int, _dp]
    L.or_set_lag.argtypes = [C.c_void_p, C.c_int, _dp]
    L.or_inject_displacement.argtypes = [C.c_void_p, C.c_int, _dp]
    L.or_get_rhs.argtypes = [C.c_void_p, C.c_int, _dp]
    L.or_get_denum.argtypes = [C.c_void_p, C.c_int, _dp]
    L.or_fft_brick.argtypes = [C.c_void_p, C.c_int, C.c_int]
    L.or_stepper_propagate.argtypes = [C.c_void_p, C.c_int]
    L.or_solver_propagate.argtypes = [C.c_void_p, C.c_int]
    L.or_get_time.argtypes = [C.c_void_p]
    L.or_get_time.restype = C.c_double
    L.or_current_step.argtypes = [C.c_void_p]
    L.or_current_step.restype = C.c_long
    L.or_eval_user_term.argtypes = [C.c_void_p, C.c_int, C.c_int, _dp]
    L.or_add_volcons.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_double, C.c_long]
    L.or_add_svisc.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_int]
    L.or_add_hessian.argtypes = [C.c_void_p, _dp, C.c_int]
    L.or_set_noise.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_ulonglong]
    L.or_add_cnoise.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_int)]
    L.or_philox.argtypes = [C.POINTER(C.c_uint), C.POINTER(C.c_uint), C.POINTER(C.c_uint)]
    L.or_noise_normal.argtypes = [C.c_long, C.c_uint, C.c_uint, C.c_ulonglong]
    L.or_noise_normal.restype = C.c_double
    L.or_volcons_multiplier.argtypes = [C.c_void_p, C.c_int]
    L.or_volcons_multiplier.restype = C.c_double
    L.or_on_step_finished.argtypes = [C.c_void_p]
    L.or_sv_interpolant.argtypes = [C.c_double, C.c_double]
    L.or_sv_interpolant.restype = C.c_double
    L.or_ch_energy.argtypes = [C.c_int, _ip, _dp, C.c_int, _dp, C.c_double]
    L.or_ch_energy.restype = C.c_double
    L.or_euler_adjoint.argtypes = [C.c_void_p, C.c_int, _dp, _dp, _dp, _dp, _dp]
    L.or_euler_adjoint.restype = C.c_int


def _i(a) -> C.Array:
    a = np.ascontiguousarray(np.asarray(a, dtype=np.int32))
    return a, a.ctypes.data_as(_ip)


def _d(a) -> C.Array:
    a = np.ascontiguousarray(np.asarray(a, dtype=np.float64))
    return a, a.ctypes.data_as(_dp)


# ---------------------------------------------------------------------------
# stand-alone helpers
# ---------------------------------------------------------------------------

def fft(data: np.ndarray, dims: Sequence[int], inverse: bool = False) -> np.ndarray:
    """FFTWWrapper.FFT / IFFT restatement on a flat complex128 array (unnormalised)."""
    buf = np.ascontiguousarray(np.array(data, dtype=np.complex128).reshape(-1))
    dk, dp = _i(list(dims))
    lib().or_fft(len(dims), dp, buf.ctypes.data_as(_dp), 1 if inverse else -1)
    return buf


def freq(dims: Sequence[int], i: int) -> List[float]:
    out = np.zeros(3)
    dk, dp = _i(list(dims))
    lib().or_freq(len(dims), dp, i, out.ctypes.data_as(_dp))
    return list(out[: len(dims)])


def conjugate_node(dims: Sequence[int], i: int) -> int:
    dk, dp = _i(list(dims))
    return int(lib().or_conjugate_node(len(dims), dp, i))


def vandeven_table(order: int) -> np.ndarray:
    out = np.zeros(1000)
    lib().or_vandeven_table(order, out.ctypes.data_as(_dp))
    return out


def vandeven_eval(table: np.ndarray, x: float) -> float:
    t = np.ascontiguousarray(table, dtype=np.float64)
    return float(lib().or_vandeven_eval(t.ctypes.data_as(_dp), x))


def cubic_material(c11: float, c12: float, c44: float) -> np.ndarray:
    out = np.zeros(81)
    lib().or_cubic_material(c11, c12, c44, out.ctypes.data_as(_dp))
    return out


def isotropic(bulk: float, poisson: float) -> np.ndarray:
    out = np.zeros(81)
    lib().or_isotropic(bulk, poisson, out.ctypes.data_as(_dp))
    return out


def energy_density(Cijkl: np.ndarray, misfit: np.ndarray) -> float:
    a, ap = _d(Cijkl)
    b, bp = _d(np.asarray(misfit).reshape(9))
    return float(lib().or_energy_density(ap, bp))


def ch_energy(dims: Sequence[int], c: np.ndarray, poly: Sequence[float], kappa: float) -> float:
    """F = sum_x sum_p poly[p] c^p + kappa/2 sum_x |grad c|^2 (Parseval, FP64 compensated)."""
    dk, dp = _i(list(dims))
    cc, cp = _d(np.asarray(c, dtype=np.float64).reshape(-1))
    pp, ppp = _d(list(poly))
    return float(lib().or_ch_energy(len(dims), dp, cp, len(poly), ppp, kappa))


# ---------------------------------------------------------------------------
# model engine
# ---------------------------------------------------------------------------

class OracleSolver:
    """The reference's Model + stepper, evaluated by the C oracle."""

    EULER, RK4 = 0, 1

    def __init__(self, model: P.Model, dims: Sequence[int], fields: Dict[str, np.ndarray],
                 dt: float, stepper: str = "euler", filter_order: int = 0) -> None:
        if not model.rhs:
            model.init()
        self.model = model
        self.dims = [int(d) for d in dims]
        self.n = int(np.prod(self.dims))
        self.nfields = len(model.fields)
        self.nderived = len(model.derived)
        names = model.field_names()
        # derived descriptors
        dkind, dts, dcoef, dfs, dff, dfp = [], [0], [], [0], [], []
        for d in model.derived:
            dkind.append(d.kind)
            for coef, facs in d.terms:
                dcoef.append(coef)
                for f, p in facs:
                    dff.append(f)
                    dfp.append(p)
                dfs.append(len(dff))
            dts.append(len(dcoef))
        # equations

        def flat(lists):
            start, kind, coef, brick, hl, lap, user = [0], [], [], [], [], [], []
            for L in lists:
                for t in L:
                    kind.append(t.kind)
                    coef += [t.coef.real, t.coef.imag]
                    brick.append(t.brick)
                    hl.append(1 if t.has_lap else 0)
                    lap.append(t.lap)
                    user.append(t.user)
                start.append(len(kind))
            return start, kind, coef, brick, hl, lap, user

        T = flat([r[0] for r in model.rhs])
        D = flat([r[1] for r in model.rhs])
        sq = model.user_tables[P.KIND_SQGRAD]
        el = model.user_tables[P.KIND_ELAST]
        sq_brick = [model.brick_index(s["field"]) for s in sq]
        sq_factor = [float(s.get("factor", 1.0)) for s in sq]
        el_field = [names.index(e["field"]) for e in el]
        el_C = [float(x) for e in el for x in np.asarray(e["C"]).reshape(81)]
        el_mis = [float(x) for e in el for x in np.asarray(e["misfit"]).reshape(9)]
        self._keep = []

        def I(a):
            arr, p = _i(a if len(a) else [0])
            self._keep.append(arr)
            return p

        def Dd(a):
            arr, p = _d(a if len(a) else [0.0])
            self._keep.append(arr)
            return p

        self.h = lib().or_model_create(
            len(self.dims), I(self.dims), self.nfields, self.nderived,
            I(dkind), I(dts), Dd(dcoef), I(dfs), I(dff), Dd(dfp),
            len(model.rhs), I(T[0]), I(T[1]), Dd(T[2]), I(T[3]), I(T[4]), I(T[5]), I(T[6]),
            I(D[0]), I(D[1]), Dd(D[2]), I(D[3]), I(D[4]), I(D[5]), I(D[6]),
            len(sq), I(sq_brick), Dd(sq_factor),
            len(el), I(el_field), Dd(el_C), Dd(el_mis),
            self.EULER if stepper == "euler" else self.RK4, float(dt), int(filter_order))
        for v in model.user_tables.get(P.KIND_VOLCONS, []):
            lib().or_add_volcons(self.h, model.brick_index(v["field"]), model.brick_index(v["indicator"]),
                                 float(v["dt"]), int(v["num_nodes"]))
        for v in model.user_tables.get(P.KIND_SVISC, []):
            lib().or_add_svisc(self.h, float(v["eps"]), float(v["threshold"]), int(v["power"]))
        for d, dd in enumerate(model.derived):
            if dd.kind == 2:
                lib().or_set_noise(self.h, d, float(dd.noise_std), int(dd.noise_seed))
        for v in model.user_tables.get(P.KIND_CNOISE, []):
            b = [model.brick_index(nm) for nm in v["names"]] + [-1, -1, -1]
            arr = (C.c_int * 3)(*b[:3])
            lib().or_add_cnoise(self.h, len(v["names"]), arr)
        for v in model.user_tables.get(P.KIND_HESS, []):
            K = np.zeros(9)
            k = np.asarray(v["K"], dtype=np.float64).reshape(-1)
            K[:k.size] = k
            lib().or_add_hessian(self.h, K.ctypes.data_as(_dp), int(k.size))
        for name, data in fields.items():
            self.set_field(name, data)

    def __del__(self):
        try:
            if getattr(self, "h", None):
                lib().or_model_destroy(self.h)
                self.h = None
        except Exception:
            pass

    # -- data ----------------------------------------------------------------
    def set_brick(self, b: int, data: np.ndarray) -> None:
        buf = np.ascontiguousarray(np.asarray(data, dtype=np.complex128).reshape(self.n))
        lib().or_set_brick(self.h, b, buf.ctypes.data_as(_dp))

    def brick(self, b: int) -> np.ndarray:
        out = np.zeros(self.n, dtype=np.complex128)
        lib().or_get_brick(self.h, b, out.ctypes.data_as(_dp))
        return out

    def set_field(self, name: str, data: np.ndarray) -> None:
        self.set_brick(self.model.brick_index(name), data)

    def field(self, name: str) -> np.ndarray:
        return self.brick(self.model.brick_index(name))

    def set_lag(self, e: int, data: np.ndarray) -> None:
        buf = np.ascontiguousarray(np.asarray(data, dtype=np.float64).reshape(self.n))
        lib().or_set_lag(self.h, e, buf.ctypes.data_as(_dp))

    def inject_displacement(self, e: int, disp: np.ndarray) -> None:
        buf = np.ascontiguousarray(np.asarray(disp, dtype=np.complex128).reshape(3 * self.n))
        lib().or_inject_displacement(self.h, e, buf.ctypes.data_as(_dp))

    # -- evaluation ------------------------------------------------------------
    def sync_derived(self) -> None:
        lib().or_sync_derived(self.h)

    def fft_brick(self, b: int, inverse: bool = False) -> None:
        lib().or_fft_brick(self.h, b, 1 if inverse else -1)

    def get_rhs(self, eq: int) -> np.ndarray:
        out = np.zeros(self.n, dtype=np.complex128)
        lib().or_get_rhs(self.h, eq, out.ctypes.data_as(_dp))
        return out

    def get_denum(self, eq: int) -> np.ndarray:
        out = np.zeros(self.n, dtype=np.complex128)
        lib().or_get_denum(self.h, eq, out.ctypes.data_as(_dp))
        return out

    def eval_user_term(self, kind: int, user: int) -> np.ndarray:
        out = np.zeros(self.n, dtype=np.complex128)
        lib().or_eval_user_term(self.h, kind, user, out.ctypes.data_as(_dp))
        return out

    def on_step_finished(self) -> None:
        lib().or_on_step_finished(self.h)

    def volcons_multiplier(self, v: int = 0) -> float:
        return float(lib().or_volcons_multiplier(self.h, v))

    # -- stepping --------------------------------------------------------------
    def step(self) -> None:
        lib().or_step(self.h)

    def stepper_propagate(self, nsteps: int) -> None:
        lib().or_stepper_propagate(self.h, nsteps)

    def solver_propagate(self, nsteps: int) -> None:
        lib().or_solver_propagate(self.h, nsteps)

    def get_time(self) -> float:
        return float(lib().or_get_time(self.h))

    # -- reverse mode (no reference counterpart) ---------------------------------
    def adjoint(self, nsteps: int, target: np.ndarray, scalars: Sequence[str] = ()):
        """Gradient of J = sum (c_T - target)^2 after nsteps Euler steps from the current field.

        Returns (J, dJ/dc_0, {scalar: dJ/dscalar}); the scalar gradients chain the C
        oracle's per-term coefficient gradients through each term's (scalar, power)
        factors.  Leaves c_T in the field.
        """
        tgt = np.ascontiguousarray(np.asarray(target).real.astype(np.float64).reshape(self.n))
        terms, denum = self.model.rhs[0]
        gt = np.zeros(max(len(terms), 1))
        gd = np.zeros(max(len(denum), 1))
        g0 = np.zeros(self.n)
        loss = C.c_double()
        rc = lib().or_euler_adjoint(self.h, int(nsteps), tgt.ctypes.data_as(_dp), C.byref(loss),
                                    g0.ctypes.data_as(_dp), gt.ctypes.data_as(_dp), gd.ctypes.data_as(_dp))
        if rc != 0:
            raise ValueError("the oracle adjoint covers one field, Euler, concrete terms")
        out = {}
        for name in scalars:
            out[name] = (sum(gt[i] * t.coef_derivative(self.model.scalars, name) for i, t in enumerate(terms))
                         + sum(gd[i] * t.coef_derivative(self.model.scalars, name) for i, t in enumerate(denum)))
        return loss.value, g0, out
