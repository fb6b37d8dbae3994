"""Synthetic workload definitions shared by the benchmark and the tests.

Initial conditions use SplitMix64(seed) -> u = (x >> 11) * 2^-53 in [0, 1), and
the Cahn-Hilliard initial field c0 = 0.5 + 0.1 (u - 0.5), i.e. 0.5 + U(-0.05, 0.05).
"""
from __future__ import annotations

import numpy as np

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64_uniform(seed: int, n: int) -> np.ndarray:
    """n uniforms in [0, 1) from SplitMix64 started at `seed` (vectorised)."""
    with np.errstate(over="ignore"):
        k = np.arange(1, n + 1, dtype=np.uint64)
        z = np.uint64(seed) + k * _GOLDEN
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        z = z ^ (z >> np.uint64(31))
    return (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def ch_initial(seed: int, n: int) -> np.ndarray:
    return 0.5 + 0.1 * (splitmix64_uniform(seed, n) - 0.5)


CH_BULK_POLY = [0.0, 0.0, 1.0, -2.0, 1.0]   # f(c) = c^2 (1 - c)^2
CH_EQUATION = "dconc/dt = m2*LAP conc - m6*LAP conc^2 + m4*LAP conc^3 - mk*LAP^2 conc"


def ch_scalars(M: float = 1.0, kappa: float = 1.0) -> dict:
    return {"m2": 2.0 * M, "m6": 6.0 * M, "m4": 4.0 * M, "mk": M * kappa}
