"""Shrinking-Dimer-Dynamics stepper on the oracle's model engine (TEST INFRASTRUCTURE ONLY).

Restates pf.SDD (pf/sdd.go:84-460) statement by statement on full complex spectra:

* ``householder`` (sdd.go:310-326) and ``householder_denum`` + ``sherman_morrison_dot``
  (sdd.go:328-345, 437-460);
* ``Step`` (sdd.go:159-301): dimer ends at x -+ l v / 2, the Householder-reflected force on
  the centre with the Sherman-Morrison solve of the implicit part, the torque on the
  orientation and its update, and the monitor values;
* ``dimer_length`` / ``required_dimer_length_time`` (sdd.go:356-368), ``init`` /
  ``set_initial_orientation`` (sdd.go:382-420), ``field_norm`` (sdd.go:371-380).

Like the reference, the weighted force of every field is built from the first field's
right-hand sides (sdd.go:212-216 index rhsStart[j], rhsEnd[j] with j < numNodes).

The fields are projected onto their real parts after every step (``real_fields``).  The
dimer climbs along its orientation, and without the projection the imaginary roundoff of
numpy's inverse FFT grows along that ascent direction by about 1.5x per step in
TestClassicalNucleation, overflowing within 170 steps.  With it, the restatement
reproduces that test's known answer.  The device keeps real fields throughout.
"""
from __future__ import annotations

import math
from typing import Dict, List, Sequence

import numpy as np

from . import parser as P
from .engine import OracleSolver


def householder(data: np.ndarray, ft_orient: np.ndarray, sigma: float, num_nodes: int) -> np.ndarray:
    """data - sigma o <data, o> / numNodes (sdd.go:313-326), on full spectra."""
    dot = np.sum(data * np.conj(ft_orient)) / num_nodes
    return data - sigma * ft_orient * dot


def householder_denum(denum: np.ndarray, ft_orient: np.ndarray, sigma: float, dt: float):
    """(invDiagonal, u, v) of (I - dt (D - sigma o o^T))^-1 (sdd.go:328-345)."""
    inv = 1.0 / (1.0 - dt * denum)
    v = sigma * denum * dt * np.conj(ft_orient) / len(denum)
    return inv, ft_orient, v


def sherman_morrison_dot(inv: np.ndarray, u: np.ndarray, v: np.ndarray, vec: np.ndarray) -> np.ndarray:
    """diagonalShermannMorrison.dot (sdd.go:443-460)."""
    denum = 1.0 + np.sum(u * inv * v)
    s = np.sum(v * inv * vec)
    return inv * vec - inv * u * s / denum


class SDDOracle:
    """pf.SDD driving an OracleSolver's bricks."""

    def __init__(self, model: P.Model, dims: Sequence[int], fields: Dict[str, np.ndarray], dt: float,
                 real_fields: bool = True) -> None:
        self.o = OracleSolver(model, dims, fields, dt)
        self.real_fields = real_fields
        self.model = model
        self.dims = [int(d) for d in dims]
        self.n = self.o.n
        self.nf = len(model.fields)
        self.dt = dt
        self.tau_orientation = 1.0
        self.tau_length = 1.0
        self.alpha = 0.5
        self.current_step = 0
        self.min_dimer_length = 0.0
        self.init_dimer_length = 0.0
        self.orientation = np.zeros(self.nf * self.n)
        self.initialized = False
        self.monitor = {"MaxForce": 0.0, "ForcePowerSpectrum": 0.0, "MaxTorque": 0.0, "FieldNorm": 0.0,
                        "FieldNormChange": 0.0}

    # -- set-up ------------------------------------------------------------
    def init(self, init: List[np.ndarray], final: List[np.ndarray]) -> None:
        v = np.concatenate([np.real(np.asarray(b) - np.asarray(a)) for a, b in zip(init, final)])
        self.init_dimer_length = math.sqrt(float(np.sum(v * v)))
        self.orientation = v / self.init_dimer_length
        self.initialized = True

    def set_initial_orientation(self, orient) -> None:
        orient = np.asarray(orient, dtype=np.float64)
        if orient.size != self.orientation.size:
            raise ValueError("Inconsistent length of the passed orientaiton vector")
        self.init_dimer_length = math.sqrt(float(np.sum(orient * orient)))
        self.orientation = orient / self.init_dimer_length
        self.initialized = True

    def get_time(self) -> float:
        return self.current_step * self.dt

    def dimer_length(self, t: float) -> float:
        length = self.init_dimer_length * math.exp(-t / self.tau_length)
        return self.min_dimer_length if length < self.min_dimer_length else length

    def required_dimer_length_time(self, length: float) -> float:
        return self.tau_length * math.log(self.init_dimer_length / length)

    # -- helpers ------------------------------------------------------------
    def _fft(self, a: np.ndarray) -> np.ndarray:
        return np.fft.fftn(a.reshape(self.dims)).reshape(-1)

    def _ifft(self, a: np.ndarray) -> np.ndarray:
        return np.fft.ifftn(a.reshape(self.dims)).reshape(-1)

    def _fields(self) -> List[np.ndarray]:
        return [self.o.brick(i) for i in range(self.nf)]

    def _set_fields(self, fields: List[np.ndarray]) -> None:
        for i, f in enumerate(fields):
            self.o.set_brick(i, f)

    def _fft_model(self) -> None:
        """SDD.fft (sdd.go:137-146): SyncDerivedFields, then FFT of fields and derived fields."""
        self.o.sync_derived()
        for b in range(self.nf + len(self.model.derived)):
            self.o.fft_brick(b)

    def _ifft_model(self) -> None:
        """SDD.ifft (sdd.go:149-154)."""
        for i in range(self.nf):
            self.o.set_brick(i, self._ifft(self.o.brick(i)))

    def _shift(self, scale: float) -> None:
        """ShiftFieldsAlongDimer (sdd.go:347-355)."""
        for i in range(self.nf):
            f = self.o.brick(i) + scale * self.orientation[i * self.n:(i + 1) * self.n]
            self.o.set_brick(i, f)

    def _rhs(self) -> np.ndarray:
        return np.concatenate([self.o.get_rhs(i) for i in range(self.nf)])

    # -- Step (sdd.go:159-301) ------------------------------------------------
    def step(self) -> None:
        if not self.initialized:
            raise RuntimeError("SDD: The method have to be initialized first. See SDD.Init")
        if self.dt < 1e-16:
            raise RuntimeError("Timestep not set in SDD.")
        n, nf, dt = self.n, self.nf, self.dt
        fnorm = float(sum(np.sum(np.abs(f) ** 2) for f in self._fields()))
        self.monitor["FieldNormChange"] = self.monitor["FieldNorm"] - fnorm
        self.monitor["FieldNorm"] = fnorm
        ft_orient = np.concatenate([self._fft(self.orientation[i * n:(i + 1) * n].astype(complex))
                                    for i in range(nf)])
        length = self.dimer_length(self.get_time())
        self._shift(-0.5 * length)
        self._fft_model()
        rhs_start = self._rhs()
        self._ifft_model()
        self._shift(length)
        self._fft_model()
        rhs_end = self._rhs()
        self._ifft_model()
        self._shift(-0.5 * length)
        self._fft_model()
        self.monitor["MaxForce"] = 0.0
        power = 0.0
        for i in range(nf):
            d = self.o.brick(i)
            orig = d.copy()
            work = self.alpha * rhs_start[:n] + (1.0 - self.alpha) * rhs_end[:n]
            active = ft_orient[i * n:(i + 1) * n]
            work = householder(work, active, 2.0, n)
            d = d + dt * work
            inv, u, v = householder_denum(self.o.get_denum(i), active, 2.0, dt)
            d = sherman_morrison_dot(inv, u, v, d)
            self.o.set_brick(i, d)
            diff = np.abs((d - orig) / dt)
            power += float(np.sum(diff * diff / n))
            self.monitor["MaxForce"] = max(self.monitor["MaxForce"], float(diff.max()))
        self.monitor["ForcePowerSpectrum"] = math.sqrt(power / (nf * n))
        torque = rhs_start - rhs_end
        for i in range(nf):
            torque[i * n:(i + 1) * n] -= self.o.get_denum(i) * ft_orient[i * n:(i + 1) * n] * length
        torque = householder(torque, ft_orient, 1.0, n)
        tau = np.concatenate([self._ifft(torque[i * n:(i + 1) * n]) for i in range(nf)]).real
        self.rhs_dot_orientation = float(np.sum(tau * self.orientation))
        self.orientation = self.orientation - dt * tau / (self.dimer_length(self.get_time()) * self.tau_orientation)
        self.monitor["MaxTorque"] = float(np.abs(tau).max())
        self.orientation = self.orientation / math.sqrt(float(np.sum(self.orientation ** 2)))
        self._ifft_model()
        if self.real_fields:
            for i in range(nf):
                self.o.set_brick(i, self.o.brick(i).real.astype(complex))
        self.current_step += 1

    def field(self, name: str) -> np.ndarray:
        return self.o.field(name)
