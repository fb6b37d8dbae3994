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


# batched sweep: 128 sims of 128^2, 192 of 256^2, 64 of 512^2
C3_CLASSES = [((128, 128), 128), ((256, 256), 192), ((512, 512), 64)]


def c3_sim_inputs(s: int, n: int):
    """Sim s of the sweep: SplitMix64(1000 + s); (M, kappa) = exp(U[ln 0.1, ln 10]) from the
    first two draws, the initial field 0.5 + U(-0.05, 0.05) from the next n."""
    u = splitmix64_uniform(1000 + s, n + 2)
    lo, hi = np.log(0.1), np.log(10.0)
    M = float(np.exp(lo + (hi - lo) * u[0]))
    kappa = float(np.exp(lo + (hi - lo) * u[1]))
    return M, kappa, 0.5 + 0.1 * (u[2:] - 0.5)


# ---------------------------------------------------------------------------
# two-field strain precipitate (examples/strain_single_precipitate/main.go:14-113,
# without CONSERVE_PREC_VOL): conc and phase, cubic elasticity, RK4
# ---------------------------------------------------------------------------
PRECIPITATE_EQUATIONS = ["dconc/dt = CHEMICALPOT + kappa*LAP conc",
                         "dphase/dt = DERIV_PHASE_ORDER + LIN_ELAST + kappa*LAP phase"]
PRECIPITATE_SCALARS = {"kappa": 0.1}


def precipitate_initial(N: int, radius: float, seed: int = 4):
    """conc = phase = 1 inside a centred sphere, 0 outside, plus U(-0.01, 0.01) (numpy PCG64 seed)."""
    n = N ** 3
    idx = np.arange(n)
    col, row, dep = idx % N, (idx // N) % N, idx // (N * N)
    r2 = (col - N / 2) ** 2 + (row - N / 2) ** 2 + (dep - N / 2) ** 2
    rng = np.random.default_rng(seed)
    inside = np.where(r2 < radius * radius, 1.0, 0.0)
    conc0 = inside + rng.uniform(-0.01, 0.01, n)
    phase0 = inside + rng.uniform(-0.01, 0.01, n)
    return conc0, phase0


def precipitate_polynomials(A: float = 0.1, B: float = 0.1, W: float = 0.1, M: float = 1.0):
    """The example's registered functions as polynomials of the fields' real parts:
    CHEMICALPOT = -M (A c (1 - H) - B (1 - c) H) with H = 3x^2 - 2x^3, and
    DERIV_PHASE_ORDER = -(-A/2 c^2 H' + B/2 (1 - c)^2 H' + W L'), H' = 6x - 6x^2,
    L' = 2x - 6x^2 + 4x^3 (x = phase)."""
    chem = [(-M * A, {"conc": 1}), (3 * M * (A - B), {"conc": 1, "phase": 2}),
            (-2 * M * (A - B), {"conc": 1, "phase": 3}), (3 * M * B, {"phase": 2}), (-2 * M * B, {"phase": 3})]
    dpo = []
    for coef, cp in ((-0.5 * A, {"conc": 2}), (0.5 * B, {}), (-B, {"conc": 1}), (0.5 * B, {"conc": 2})):
        dpo.append((-coef * 6.0, dict(cp, phase=1)))
        dpo.append((coef * 6.0, dict(cp, phase=2)))
    dpo += [(-2 * W, {"phase": 1}), (6 * W, {"phase": 2}), (-4 * W, {"phase": 3})]

    def norm(terms):
        out = []
        for c, p in terms:
            q = {"conc": 0.0, "phase": 0.0}
            q.update({k: float(v) for k, v in p.items()})
            out.append((c, q))
        return out

    return {"CHEMICALPOT": norm(chem), "DERIV_PHASE_ORDER": norm(dpo)}
