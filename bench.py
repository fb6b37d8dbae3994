#!/usr/bin/env python3
"""Benchmark: 3D Cahn-Hilliard 256^3, semi-implicit Euler, on one MI355X.

One step = one semi-implicit Euler step of dc/dt = M LAP f'(c) - M kappa LAP^2 c,
f = c^2 (1 - c)^2, M = kappa = 1, dt = 0.05, over the whole 256^3 grid (pf.Euler.Step,
pf/euler.go:16-47, with the model string of examples/cahnHilliard/main.go:33 form).

Headline value: FP32-bulk voxel-steps/s (FP64 k = 0 mode, FP64 denominators and
reductions).  The plain FP64 path is timed beside it; the FP32-bulk field is
compared with the FP64 field after the same steps; the dominant kernel's
achieved HBM bandwidth is measured with HIP events on the solver's stream; the
FP64 CPU oracle is timed on the host for one step of the same workload.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "voxel-steps/s, 256³ CH: FP32-bulk vs plain FP64 at stated err; % HBM peak"
HBM_PEAK_GBS = 8000.0


def pass_bytes(name: str, nx: int, ny: int, nz: int, s: int) -> float:
    """Algorithmic HBM bytes of one launch of each pass of the fused 3D step.

    Arrays are half spectra of (nx/2 + 1) x ny x nz complex values of 2s bytes.  The state
    S is the spectrum of c; the work array W holds S inverse-transformed along y between
    steps:
      pencils_inv   inverse z of W: reads and writes W                 2 arrays
      rows_fused    x inverse, g(c), x forward: reads and writes W     2 arrays
      pencils_fwd   forward z of W: reads and writes W                 2 arrays
      pencil_update forward y of W, update of S, inverse y of S:
                    reads W and S, writes S and W                      4 arrays
    10 arrays per step: 10 (nx/2 + 1)/nx * 2s ~ 10s bytes per voxel-step.
    """
    arr = (nx // 2 + 1) * ny * nz * 2 * s
    return {"rows_fused": 2, "pencils_fwd": 2, "pencil_update": 4, "pencils_inv": 2}[name] * arr


def build_solver(G, W, n, dims, c0, precision, dt):
    m = G.Model()
    m.AddField("conc", c0)
    for k, v in W.ch_scalars(1.0, 1.0).items():
        m.AddScalar(k, v)
    m.AddEquation(W.CH_EQUATION)
    m.Init()
    s = G.Solver(m, dims, dt, precision=precision, autotune=True)
    if not s.fused:
        raise RuntimeError("the Cahn-Hilliard model did not compile to the fused path")
    return s


def time_steps(s, steps):
    s.Synchronize()
    t0 = time.perf_counter()
    s.StepperPropagate(steps)
    s.Synchronize()
    return time.perf_counter() - t0


def cpu_baseline(n_side: int, dt: float, c0: np.ndarray):
    """FP64 oracle (restatement of the reference's Euler step with complex128 FFTs) on one core."""
    from oracle import engine as E
    from oracle import parser as P
    from gopf_amd import workloads as W
    E.build()
    dims = [n_side] * 3
    n = n_side ** 3
    m = P.Model()
    m.add_field("conc", n)
    for k, v in W.ch_scalars(1.0, 1.0).items():
        m.add_scalar(k, v)
    m.add_equation(W.CH_EQUATION)
    m.init()
    o = E.OracleSolver(m, dims, {"conc": c0}, dt)
    t0 = time.perf_counter()
    o.stepper_propagate(1)
    el = time.perf_counter() - t0
    return {"value": n / el, "unit": "voxel-steps/s", "cores": 1, "kind": "port",
            "sample": f"1 Euler step of the {n_side}^3 CH workload by the FP64 C oracle "
                      f"(3 forward + 1 inverse complex128 3D FFT, cmplx.Pow derived fields), {el:.1f} s"}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--dt", type=float, default=0.05)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle timing")
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64 comparison run")
    ap.add_argument("--err-steps", type=int, default=1000, help="steps of the FP32-vs-FP64 accuracy run")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU: the multi-GPU slab decomposition is "
                                   "design-only (DESIGN.md); run with --gpus 1"}))
        return 2
    import gopf_amd as G
    from gopf_amd import workloads as W

    N = args.n
    dims = [N, N, N]
    n = N ** 3
    c0 = W.ch_initial(1, n)
    copy_bytes = 1 << 30
    copy_gbs = G.CopyBandwidth(copy_bytes, reps=10)
    print(f"# device copy of {copy_bytes} B: {copy_gbs:.0f} GB/s", file=sys.stderr, flush=True)
    print(f"# building FP32-bulk solver {dims}", file=sys.stderr, flush=True)
    s32 = build_solver(G, W, n, dims, c0, "fp32", args.dt)
    s32.StepperPropagate(args.warmup)   # the first step runs the lines-per-block autotuner
    el32 = time_steps(s32, args.steps)
    # per-pass device times over a second timed region
    tsteps = min(args.steps, 200)
    s32.EnableTiming(True)
    s32.StepperPropagate(tsteps)
    passes = s32.PassTimes()
    s32.EnableTiming(False)
    tuning = s32.Tuning()
    print(f"# fp32: {1e3 * el32 / args.steps:.4f} ms/step; passes {passes}", file=sys.stderr, flush=True)
    res = {}
    if not args.no_fp64:
        s64 = build_solver(G, W, n, dims, c0, "fp64", args.dt)
        s64.StepperPropagate(args.warmup)
        el64 = time_steps(s64, args.steps)
        del s64
        print(f"# fp64: {1e3 * el64 / args.steps:.4f} ms/step", file=sys.stderr, flush=True)
        # accuracy at the stated step count, from the same initial field
        del s32
        e32 = build_solver(G, W, n, dims, c0, "fp32", args.dt)
        e32.StepperPropagate(args.err_steps)
        f32 = e32.Field(0).real
        _, m32 = e32.Energy(0, W.CH_BULK_POLY, 1.0)
        del e32
        e64 = build_solver(G, W, n, dims, c0, "fp64", args.dt)
        e64.StepperPropagate(args.err_steps)
        f64 = e64.Field(0).real
        _, m64 = e64.Energy(0, W.CH_BULK_POLY, 1.0)
        del e64
        m0 = float(np.sum(c0))
        res = {
            "fp64": {"value": n * args.steps / el64, "ms_per_step": 1e3 * el64 / args.steps},
            "speedup_fp32_vs_fp64": el64 / el32,
            "err_vs_fp64": {"steps": args.err_steps,
                            "rel_l2": float(np.linalg.norm(f32 - f64) / np.linalg.norm(f64)),
                            "max_rel": float(np.abs(f32 - f64).max() / np.abs(f64).max())},
            "mass_drift": {"fp32": abs(m32 - m0) / m0, "fp64": abs(m64 - m0) / m0},
        }
        print(f"# err after {args.err_steps} steps: {res['err_vs_fp64']}", file=sys.stderr, flush=True)
    value = n * args.steps / el32
    # roofline of the dominant kernel
    best = max(passes.items(), key=lambda kv: kv[1][0])
    name, (ms, cnt) = best
    avg_s = ms / cnt / 1e3
    algo = pass_bytes(name, N, N, N, 4)
    achieved = algo / avg_s / 1e9
    step_bytes = sum(pass_bytes(k, N, N, N, 4) for k in passes)
    roof = {"bound": "hbm", "kernel": name, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": None, "bytes_per_launch": algo,
            "avg_launch_ms": ms / cnt, "lines_per_block": tuning,
            "passes_ms_per_launch": {k: v[0] / v[1] for k, v in passes.items()},
            "step_effective_GBs": step_bytes * args.steps / el32 / 1e9,
            "step_frac": step_bytes * args.steps / el32 / 1e9 / HBM_PEAK_GBS,
            "measured_copy_GBs": copy_gbs, "frac_of_measured_copy": achieved / copy_gbs}
    cpu = None
    if not args.no_cpu:
        print("# timing the CPU oracle (one step)", file=sys.stderr, flush=True)
        cpu = cpu_baseline(N, args.dt, c0)
    out = {
        "metric": METRIC, "value": value, "unit": "voxel-steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": 1e3 * el32 / args.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"cahn_hilliard_3d_{N}", "grid": dims, "dt": args.dt, "M": 1.0, "kappa": 1.0,
                   "stepper": "semi-implicit Euler", "init": "0.5 + U(-0.05, 0.05), SplitMix64 seed 1",
                   "precision": "fp32 bulk; fp64 k=0 mode, denominators, reductions"},
        "roofline": roof, "cpu_baseline": cpu,
    }
    out.update(res)
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
