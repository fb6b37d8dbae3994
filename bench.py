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


TUNING_FILE = os.path.join(ROOT, "gopf_amd", "tuning", "gfx950.json")


def pmc_traffic(workload: str, kernel: str, layout: dict):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 --pmc summary
    (profiles/pmc_traffic.json: FETCH_SIZE scaled by the copy-kernel calibration, plus
    WRITE_SIZE), and None with the reason when the summary does not cover this workload or
    was measured under another launch layout than `layout`."""
    path = os.path.join(ROOT, "profiles", "pmc_traffic.json")
    try:
        with open(path) as f:
            d = json.load(f)
        w = d["workloads"][workload]
        if w.get("layout") != layout:
            return None, "the committed counters were measured under another launch layout"
        return float(w["kernels"][kernel]["hbm_bytes"]), None
    except (OSError, KeyError, ValueError, TypeError):
        return None, "no committed counter summary for this workload and kernel"


def stored_layout(workload: str, precision: str):
    """The launch layout persisted for this workload and precision (autotuned once on an
    MI355X, then reused, so that every run and every profile uses the same kernels)."""
    try:
        with open(TUNING_FILE) as f:
            return json.load(f)[workload][precision]
    except (OSError, KeyError, ValueError):
        return None


def build_solver(G, W, n, dims, c0, precision, dt, layout=None):
    m = G.Model()
    m.AddField("conc", c0)
    for k, v in W.ch_scalars(1.0, 1.0).items():
        m.AddScalar(k, v)
    m.AddEquation(W.CH_EQUATION)
    m.Init()
    s = G.Solver(m, dims, dt, precision=precision, autotune=layout is None)
    if not s.fused:
        raise RuntimeError("the Cahn-Hilliard model did not compile to the fused path")
    if layout is not None:
        s.SetTuning(layout)
    return s


def time_steps(s, steps):
    s.Synchronize()
    t0 = time.perf_counter()
    s.StepperPropagate(steps)
    s.Synchronize()
    return time.perf_counter() - t0


class StreamTimer:
    """HIP events recorded on the solver's own stream around the timed steps."""

    def __init__(self):
        import ctypes
        self.C = ctypes
        self.hip = ctypes.CDLL("libamdhip64.so")
        self.e0, self.e1 = ctypes.c_void_p(), ctypes.c_void_p()
        for e in (self.e0, self.e1):
            if self.hip.hipEventCreate(ctypes.byref(e)) != 0:
                raise RuntimeError("hipEventCreate failed")

    def time(self, s, steps):
        """(host seconds, device seconds) of `steps` steps of solver s."""
        C = self.C
        stream = C.c_void_p(s.Stream())
        s.Synchronize()
        t0 = time.perf_counter()
        if self.hip.hipEventRecord(self.e0, stream) != 0:
            raise RuntimeError("hipEventRecord failed")
        s.StepperPropagate(steps)
        if self.hip.hipEventRecord(self.e1, stream) != 0:
            raise RuntimeError("hipEventRecord failed")
        s.Synchronize()
        host = time.perf_counter() - t0
        if self.hip.hipEventSynchronize(self.e1) != 0:
            raise RuntimeError("hipEventSynchronize failed")
        ms = C.c_float()
        if self.hip.hipEventElapsedTime(C.byref(ms), self.e0, self.e1) != 0:
            raise RuntimeError("hipEventElapsedTime failed")
        return host, ms.value / 1e3


def host_cpu():
    """Host core count and CPU model, for the CPU baseline's record."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count()
    return {"nproc": avail, "cpu_model": model}


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
    out = {"value": n / el, "unit": "voxel-steps/s", "cores": 1, "kind": "port",
           "label": "restatement, not reference (the Go+FFTW reference cannot be built here)",
           "sample": f"1 Euler step of the {n_side}^3 CH workload by the FP64 C oracle "
                     f"(3 forward + 1 inverse complex128 3D FFT, cmplx.Pow derived fields), {el:.1f} s"}
    out.update(host_cpu())
    return out


# ---------------------------------------------------------------------------
# the other workloads (each prints its own JSON line with --config)
# ---------------------------------------------------------------------------

def _oracle_ch(n_dims, c0, dt, scal):
    from oracle import engine as E
    from oracle import parser as P
    from gopf_amd import workloads as W
    E.build()
    m = P.Model()
    m.add_field("conc", int(np.prod(n_dims)))
    for k, v in scal.items():
        m.add_scalar(k, v)
    m.add_equation(W.CH_EQUATION)
    m.init()
    return E.OracleSolver(m, n_dims, {"conc": c0}, dt)


def cpu_baseline_ch2d(N: int, steps: int, dt: float, c0, scal, what: str):
    """FP64 oracle on one core: `steps` Euler steps of an N^2 CH sim."""
    o = _oracle_ch([N, N], c0, dt, scal)
    t0 = time.perf_counter()
    o.stepper_propagate(steps)
    el = time.perf_counter() - t0
    return {"value": N * N * steps / el, "unit": "voxel-steps/s", "cores": 1, "kind": "port",
            "sample": f"{what}: {steps} Euler steps of one {N}^2 CH sim by the FP64 C oracle, {el:.1f} s"}


def run_c1(G, W, args):
    """128^2 CH, FP64, seed 0, dt 0.1, M = kappa = 1, 1000 steps with the free-energy series.
    The warm-up (W steps, at least 2000: about 25 ms, so that the timed 13 ms start at the
    device's steady clocks) runs before the field is reset to the initial condition."""
    N, dt, steps = 128, 0.1, args.steps
    dims = [N, N]
    c0 = W.ch_initial(0, N * N)
    m = G.Model()
    m.AddField("conc", c0)
    for k, v in W.ch_scalars(1.0, 1.0).items():
        m.AddScalar(k, v)
    m.AddEquation(W.CH_EQUATION)
    m.Init()
    s = G.Solver(m, dims, dt, precision="fp64", autotune=True)
    s.SetPersistent({"auto": -1, "on": 1, "off": 0}[args.persistent])
    s.StepperPropagate(max(args.warmup, 2000))   # tunes the launch configuration, warms the clocks
    s.SetField(0, c0)
    e0, m0 = s.Energy(0, W.CH_BULK_POLY, 1.0)
    s.MonitorEnergy(W.CH_BULK_POLY, 1.0)
    p0 = s.PersistentSteps()
    el, dev = StreamTimer().time(s, steps)
    persistent = s.PersistentSteps() - p0 == steps
    e, mass = s.EnergySeries()
    series = np.concatenate([[e0], e])
    out = {"metric": "steps/s, 128^2 CH FP64 with the per-step free-energy series", "value": steps / el,
           "unit": "steps/s", "ms_per_step": 1e3 * el / steps, "ms_per_step_events": 1e3 * dev / steps,
           "voxel_steps_per_s": N * N * steps / el,
           "dtype": "f64", "config": {"workload": "cahn_hilliard_2d_128", "grid": dims, "dt": dt, "steps": steps,
                                      "init": "0.5 + U(-0.05, 0.05), SplitMix64 seed 0",
                                      "launch": "persistent" if persistent else "per-pass graphs",
                                      "warmup_steps": max(args.warmup, 2000)},
           "energy": {"values": int(series.size), "first": float(series[0]), "last": float(series[-1]),
                      "max_increase": float(np.max(np.diff(series))), "mass_drift": float(abs(mass[-1] - m0) / m0)}}
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline_ch2d(N, min(steps, 200), dt, c0, W.ch_scalars(1.0, 1.0), "sample")
    return out


def run_c3(G, W, args):
    """384 CH sims: 128 of 128^2, 192 of 256^2, 64 of 512^2, FP32-bulk, dt 0.1, 500 steps."""
    classes = W.C3_CLASSES
    dt, steps = 0.1, args.steps
    m = G.Model()
    m.AddField("conc", n=16)
    for k, v in W.ch_scalars().items():
        m.AddScalar(k, v)
    m.AddEquation(W.CH_EQUATION)
    m.Init()
    b = G.Batch(m, classes, dt, precision="fp32", autotune=True)
    vox = 0
    for s in range(len(b)):
        n = int(np.prod(b.shapes[s]))
        M, kappa, c0 = W.c3_sim_inputs(s, n)
        for k, v in W.ch_scalars(M, kappa).items():
            b.SetScalar(s, k, v)
        b.SetField(s, c0)
        vox += n
    b.Propagate(args.warmup)          # compiles, tunes and captures the step graph
    b.Synchronize()
    t0 = time.perf_counter()
    b.Propagate(steps)
    b.Synchronize()
    el = time.perf_counter() - t0
    out = {"metric": "voxel-steps/s, 384 batched 2D CH sims of uneven sizes, FP32-bulk", "value": vox * steps / el,
           "unit": "voxel-steps/s", "ms_per_step": 1e3 * el / steps, "dtype": "f32",
           "config": {"workload": "batched_ch_2d_384", "classes": [[list(c), k] for c, k in classes],
                      "voxels_per_step": vox, "dt": dt, "steps": steps,
                      "params": "(M, kappa) = exp(U[ln 0.1, ln 10]), SplitMix64 seed 1000 + sim"},
           "layout": b.Tuning()}
    if not args.no_cpu:
        M, kappa, c0 = W.c3_sim_inputs(128, 256 * 256)
        out["cpu_baseline"] = cpu_baseline_ch2d(256, 20, dt, c0, W.ch_scalars(M, kappa),
                                                "one 256^2 sim of the sweep (sim 128)")
    return out


def precipitate_model(G, W, conc0, phase0):
    m = G.Model()
    m.AddField("conc", conc0)
    m.AddField("phase", phase0)
    for k, v in W.PRECIPITATE_SCALARS.items():
        m.AddScalar(k, v)
    for name, poly in W.precipitate_polynomials().items():
        m.RegisterPolynomial(name, poly)
    m.RegisterExplicitTerm("LIN_ELAST", G.HomogeneousModulusLinElast("phase", G.CubicMaterial(110.0, 60.0, 30.0),
                                                                     0.01 * np.eye(3)))
    for e in W.PRECIPITATE_EQUATIONS:
        m.AddEquation(e)
    m.Init()
    return m


def cpu_baseline_c4(W, N: int, steps: int):
    """FP64 oracle on one core: the precipitate model at N^3, RK4 with the full 14-FFT
    elastic evaluation of the reference."""
    from oracle import engine as E
    from oracle import parser as P
    E.build()
    conc0, phase0 = W.precipitate_initial(N, 20.0 * N / 192.0, seed=4)
    m = P.Model()
    n = N ** 3
    m.add_field("conc", n)
    m.add_field("phase", n)
    for k, v in W.PRECIPITATE_SCALARS.items():
        m.add_scalar(k, v)
    for name, poly in W.precipitate_polynomials().items():
        m.register_polynomial(name, poly)
    import gopf_amd as G
    m.register_explicit_term("LIN_ELAST", P.KIND_ELAST, field="phase", C=G.CubicMaterial(110.0, 60.0, 30.0),
                             misfit=0.01 * np.eye(3))
    for e in W.PRECIPITATE_EQUATIONS:
        m.add_equation(e)
    m.init()
    o = E.OracleSolver(m, [N, N, N], {"conc": conc0, "phase": phase0}, 0.01, "rk4")
    t0 = time.perf_counter()
    o.solver_propagate(steps)
    el = time.perf_counter() - t0
    return {"value": n * steps / el, "unit": "voxel-steps/s", "cores": 1, "kind": "port",
            "sample": f"{steps} RK4 steps of the precipitate model at {N}^3 (sphere radius scaled to "
                      f"{20.0 * N / 192.0:.1f}) by the FP64 C oracle, {el:.1f} s; voxel-steps/s carries over "
                      f"to 192^3 up to the FFT's log factor"}


def c4_traffic(N: int, s: int) -> dict:
    """Algorithmic HBM bytes of one step of the two-field precipitate RK4 step, every array
    counted once per pass that reads or writes it.  R = a real array (N^3 s bytes), H = a
    half spectrum ((N/2+1) N^2 complex values), K = a per-k table in the field precision
    ((N/2+1) N^2 values of s bytes: the elastic multiplier, and each RK4 stage's factor).
    Per step: the derived fields' transform (rows with the polynomials inline, y, z) 2R+4H
    read / 6H written; the elastic term (rows of phi with the indicator inline, y, z; the
    K(k) multiply; z, y inverse; the fused row pass of H'(phi)(S - 2 e* H(phi)), which reads
    phi; y, z) 2R+8H+K / 9H; eight RK4 stage passes (conc 13H / 8H, phase 17H / 8H, and the
    stage factor table, one K per pass); three between-stage syncs (z, y inverse of both
    fields, the fused row pass, y, z forward of both derived fields) 30H / 30H; the last
    inverse (z, y, rows) 6H / 4H+2R; the lagged-field copy R / R.
    FP64 (s = 8) does each stage's update in the fields' first inverse pencil pass: the
    corrected spectra are neither read back by that pass (8H) nor, before the last stage,
    written (6H)."""
    R = N ** 3 * s
    H = (N // 2 + 1) * N * N * 2 * s
    K = (N // 2 + 1) * N * N * s
    fused = s == 8
    rd = 5 * R + (70 if fused else 78) * H + 9 * K
    wr = 3 * R + (59 if fused else 65) * H
    return {"read": rd, "write": wr, "total": rd + wr, "per_voxel_step": (rd + wr) / N ** 3}


def run_c4(G, W, args):
    """Two-field precipitate with the Khachaturyan elastic term, 192^3, RK4, dt 0.01, 500 steps."""
    N, dt, steps = 192, 0.01, args.steps
    dims = [N, N, N]
    conc0, phase0 = W.precipitate_initial(N, 20.0, seed=4)
    res = {}
    fields = {}
    for prec in ("fp64", "fp32"):
        s = G.Solver(precipitate_model(G, W, conc0, phase0), dims, dt, stepper="rk4", precision=prec)
        s.Propagate(args.warmup)
        s.SetField("conc", conc0)
        s.SetField("phase", phase0)
        el = 0.0
        s.Synchronize()
        t0 = time.perf_counter()
        s.Propagate(steps)
        s.Synchronize()
        el = time.perf_counter() - t0
        tr = c4_traffic(N, 4 if prec == "fp32" else 8)
        res[prec] = {"value": N ** 3 * steps / el, "ms_per_step": 1e3 * el / steps,
                     "algorithmic_bytes_per_voxel_step": tr["per_voxel_step"],
                     "achieved_GBs": tr["total"] * steps / el / 1e9,
                     "frac_of_8TBs": tr["total"] * steps / el / 8e12}
        fields[prec] = (s.Field("conc").real, s.Field("phase").real)
        print(f"# c4 {prec}: {1e3 * el / steps:.3f} ms/step", file=sys.stderr, flush=True)
        del s
    err = {f: float(np.linalg.norm(fields["fp32"][i] - fields["fp64"][i]) / np.linalg.norm(fields["fp64"][i]))
           for i, f in enumerate(("conc", "phase"))}
    out = {"metric": "voxel-steps/s, 192^3 two-field elastic RK4", "value": res["fp32"]["value"],
           "unit": "voxel-steps/s", "ms_per_step": res["fp32"]["ms_per_step"], "dtype": "f32",
           "fp64": res["fp64"], "speedup_fp32_vs_fp64": res["fp64"]["ms_per_step"] / res["fp32"]["ms_per_step"],
           "algorithmic_bytes_per_voxel_step": res["fp32"]["algorithmic_bytes_per_voxel_step"],
           "achieved_GBs": res["fp32"]["achieved_GBs"], "frac_of_8TBs": res["fp32"]["frac_of_8TBs"],
           "err_vs_fp64_rel_l2": err,
           "config": {"workload": "precipitate_elastic_3d_192", "grid": dims, "dt": dt, "steps": steps,
                      "stepper": "RK4", "elastic": "cubic C11=110 C12=60 C44=30, misfit 0.01 I",
                      "init": "sphere r=20 + U(-0.01, 0.01), seed 4; lagged elastic field starts at zero"}}
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline_c4(W, 32, 2)
    return out


def run_c5(G, W, args):
    """Adjoint: 512^2 CH, 2000 steps, dt 0.1, checkpoints every 100 steps, target from a
    forward run at kappa = 1.3 (seed 6); gradients with respect to c0 and kappa."""
    N, dt, steps = 512, 0.1, args.steps
    dims = [N, N]
    n = N * N
    c0 = W.ch_initial(5, n)

    def solver(prec, scal, init):
        m = G.Model()
        m.AddField("conc", init)
        for k, v in scal.items():
            m.AddScalar(k, v)
        m.AddEquation(W.CH_EQUATION)
        m.Init()
        return G.Solver(m, dims, dt, precision=prec, autotune=True)

    ts = solver("fp64", W.ch_scalars(1.0, 1.3), W.ch_initial(6, n))
    ts.StepperPropagate(steps)
    tgt = ts.Field(0).real
    del ts
    res = {}
    grads = {}
    for prec in ("fp64", "fp32"):
        s = solver(prec, W.ch_scalars(1.0, 1.0), c0)
        s.StepperPropagate(args.warmup)
        s.SetField(0, c0)
        tf = time_steps(s, steps)
        s.SetField(0, c0)
        s.Synchronize()
        t0 = time.perf_counter()
        J, g0, gs = s.Adjoint(steps, tgt, ["mk"], checkpoint_every=100)
        ta = time.perf_counter() - t0
        res[prec] = {"adjoint_s": ta, "forward_s": tf, "adjoint_over_forward": ta / tf, "loss": J,
                     "dJ_dkappa": gs["mk"], "value": n * steps / ta}
        grads[prec] = g0
        print(f"# c5 {prec}: forward {tf:.3f} s, adjoint {ta:.3f} s", file=sys.stderr, flush=True)
        del s
    rel = float(np.linalg.norm(grads["fp32"] - grads["fp64"]) / np.linalg.norm(grads["fp64"]))
    out = {"metric": "voxel-steps/s of the adjoint (forward + checkpointed recompute + reverse), 512^2 CH",
           "value": res["fp64"]["value"], "unit": "voxel-steps/s", "dtype": "f64", "fp64": res["fp64"],
           "fp32": res["fp32"], "fp32_grad_rel_l2_vs_fp64": rel,
           "config": {"workload": "adjoint_ch_2d_512", "grid": dims, "dt": dt, "steps": steps,
                      "checkpoint_every": 100, "loss": "sum (c_T - c_target)^2",
                      "target": "forward run at kappa = 1.3 from seed 6", "init": "seed 5"}}
    return out


def run_ie(G, W, args):
    """ImplicitEuler on the 128^2 CH workload (dt 0.1, M = kappa = 1, seed 0) with the
    reference's default NewtonKrylov settings (Maxiter 50, StepSize 1e-3, Tol 1e-7, Stencil 6)."""
    N, dt, steps = 128, 0.1, args.steps
    dims = [N, N]
    n = N * N
    c0 = W.ch_initial(0, n)
    m = G.Model()
    m.AddField("conc", c0)
    for k, v in W.ch_scalars(1.0, 1.0).items():
        m.AddScalar(k, v)
    m.AddEquation(W.CH_EQUATION)
    m.Init()
    s = G.Solver(m, dims, dt, stepper="implicitEuler")
    s.Propagate(args.warmup)
    s.SetField(0, c0)
    s.Synchronize()
    newton, krylov, conv = [], [], 0
    t0 = time.perf_counter()
    for _ in range(steps):
        s.Propagate(1)
        st = s.NewtonStats()
        newton.append(st["newton_iters"])
        krylov.append(st["krylov_iters"])
        conv += int(st["converged"])
    s.Synchronize()
    el = time.perf_counter() - t0
    # residual evaluations: one per Newton check, Stencil per Arnoldi step
    evals = sum(newton) + steps + 6 * sum(krylov)
    print(f"# ie: {1e3 * el / steps:.2f} ms/step, {evals / steps:.1f} residual evaluations/step",
          file=sys.stderr, flush=True)
    out = {"metric": "steps/s, 128^2 CH FP64 ImplicitEuler (Newton-Krylov, GMRES on the device)",
           "value": steps / el, "unit": "steps/s", "ms_per_step": 1e3 * el / steps, "dtype": "f64",
           "newton_iters_per_step": float(np.mean(newton)), "krylov_iters_per_step": float(np.mean(krylov)),
           "residual_evals_per_step": evals / steps, "us_per_residual_eval": 1e6 * el / evals,
           "converged_steps": conv,
           "config": {"workload": "implicit_euler_ch_2d_128", "grid": dims, "dt": dt, "steps": steps,
                      "newton_krylov": {"Maxiter": 50, "StepSize": 1e-3, "Tol": 1e-7, "Stencil": 6, "Restart": 30},
                      "init": "0.5 + U(-0.05, 0.05), SplitMix64 seed 0"}}
    if not args.no_cpu:
        from oracle import engine as E
        from oracle import parser as P
        from oracle.implicit import ImplicitEulerOracle
        E.build()
        om = P.Model()
        om.add_field("conc", n)
        for k, v in W.ch_scalars(1.0, 1.0).items():
            om.add_scalar(k, v)
        om.add_equation(W.CH_EQUATION)
        om.init()
        o = ImplicitEulerOracle(om, dims, {"conc": c0.astype(complex)}, dt, f_tol=1e-7)
        k = 5
        t0 = time.perf_counter()
        for _ in range(k):
            o.step()
        oel = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": k / oel, "unit": "steps/s", "cores": 1, "kind": "port",
                               "sample": f"{k} ImplicitEuler steps of the same sim by the oracle (C oracle "
                                         f"bricks, numpy FFTs, SciPy Newton-Krylov to max|F| < 1e-7), {oel:.1f} s"}
    return out


def run_io(G, W, args):
    """Float64IO snapshots of the 256^3 CH FP32 sim while it steps: 1000 steps with a snapshot
    every 200 steps against 1000 plain steps, and one synchronous snapshot (Field() to the host,
    big-endian encoding, file write) for comparison.  The files go to /tmp (page cache), whose
    write rate bounds how often snapshots can be taken without holding the stepping back."""
    import shutil
    import tempfile
    N, dt, steps, every = 256, 0.05, 1000, 200
    dims = [N, N, N]
    n = N ** 3
    c0 = W.ch_initial(1, n)
    s = build_solver(G, W, n, dims, c0, "fp32", dt)
    s.StepperPropagate(args.warmup)
    d = tempfile.mkdtemp(prefix="gpf_io_", dir="/tmp")
    try:
        w = G.Float64IO(os.path.join(d, "snap"))
        for k in (98, 99):    # allocates both staging and pinned buffers, starts the writer
            w.SaveFields(s, k)
        s.FlushIO()
        plain = time_steps(s, steps)
        s.Synchronize()
        t0 = time.perf_counter()
        for k in range(steps // every):
            s.StepperPropagate(every)
            w.SaveFields(s, k)
        s.Synchronize()
        t_dev = time.perf_counter() - t0
        s.FlushIO()
        t_all = time.perf_counter() - t0
        nsnap = steps // every
        # the synchronous path a host callback takes
        t1 = time.perf_counter()
        f = s.Field("conc").real
        np.asarray(f, dtype=">f8").tofile(os.path.join(d, "sync_conc_0.bin"))
        t_sync = time.perf_counter() - t1
        ok = G.LoadFloat64(os.path.join(d, f"snap_conc_{nsnap - 1}.bin")).size == n
    finally:
        shutil.rmtree(d, ignore_errors=True)
    snap_bytes = n * 8
    overhead = (t_dev - plain) / nsnap
    print(f"# io: plain {1e3 * plain / steps:.4f} ms/step, with snapshots {1e3 * t_dev / steps:.4f} ms/step, "
          f"all files written after {t_all:.2f} s; sync snapshot {t_sync:.3f} s", file=sys.stderr, flush=True)
    out = {"metric": "stepping overhead per asynchronous Float64IO snapshot, 256^3 CH FP32",
           "value": 1e3 * overhead, "unit": "ms/snapshot", "higher_is_better": False, "dtype": "f32",
           "ms_per_step_plain": 1e3 * plain / steps, "ms_per_step_with_snapshots": 1e3 * t_dev / steps,
           "snapshots": nsnap, "snapshot_bytes": snap_bytes, "files_complete_s": t_all,
           "file_rate_GBs": nsnap * snap_bytes / t_all / 1e9, "sync_snapshot_s": t_sync, "files_ok": bool(ok),
           "config": {"workload": "float64io_snapshots_ch3d_256", "grid": dims, "dt": dt, "steps": steps,
                      "snapshot_every": every}}
    if not args.no_cpu:
        x = np.random.default_rng(0).random(n)
        d2 = tempfile.mkdtemp(prefix="gpf_io_cpu_", dir="/tmp")
        try:
            t1 = time.perf_counter()
            np.asarray(x, dtype=">f8").tofile(os.path.join(d2, "c.bin"))
            el = time.perf_counter() - t1
        finally:
            shutil.rmtree(d2, ignore_errors=True)
        out["cpu_baseline"] = {"value": 1e3 * el, "unit": "ms/snapshot", "cores": 1, "kind": "port",
                               "sample": "one 256^3 field encoded big-endian and written from host memory "
                                         "(SaveFloat64's work, fileIO.go:87-95), numpy"}
    return out


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
    ap.add_argument("--autotune", action="store_true",
                    help="run the autotuner instead of applying the stored launch layout")
    ap.add_argument("--save-tuning", default=None,
                    help="write the layouts this run used to this JSON file (stored-layout format)")
    ap.add_argument("--persistent", default="auto", choices=["auto", "on", "off"],
                    help="c1: the persistent step (every step in one cooperative launch) or the per-pass launches")
    ap.add_argument("--config", default="c2", choices=["c1", "c2", "c3", "c4", "c5", "ie", "io"],
                    help="c2 (default): 256^3 CH headline; c1 128^2 CH FP64 with energy series; c3 batched "
                         "sweep; c4 192^3 elastic RK4; c5 512^2 adjoint; ie 128^2 CH ImplicitEuler; io snapshots")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this library runs on one GPU: the multi-GPU slab decomposition along z "
                                   "is not implemented; run with --gpus 1"}))
        return 2
    import gopf_amd as G
    from gopf_amd import workloads as W

    if args.config != "c2":
        defaults = {"c1": 1000, "c3": 500, "c4": 500, "c5": 2000, "ie": 50, "io": 1000}
        if "--steps" not in sys.argv:
            args.steps = defaults[args.config]
        out = {"c1": run_c1, "c3": run_c3, "c4": run_c4, "c5": run_c5, "ie": run_ie,
               "io": run_io}[args.config](G, W, args)
        out.setdefault("higher_is_better", True)
        out.update({"n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "data": "synthetic",
                    "vs_baseline": None})
        print(json.dumps(out))
        return 0

    N = args.n
    dims = [N, N, N]
    n = N ** 3
    c0 = W.ch_initial(1, n)
    workload = f"cahn_hilliard_3d_{N}"
    probe_bytes = 1 << 30
    bw = {k: G.DeviceBandwidth(probe_bytes, k) for k in ("copy", "read", "write")}
    print(f"# device bandwidth over {probe_bytes} B: copy {bw['copy']:.0f}, read {bw['read']:.0f}, "
          f"write {bw['write']:.0f} GB/s", file=sys.stderr, flush=True)
    lay = {p: (None if args.autotune else stored_layout(workload, p)) for p in ("fp32", "fp64")}
    print(f"# building FP32-bulk solver {dims}", file=sys.stderr, flush=True)
    timer = StreamTimer()
    s32 = build_solver(G, W, n, dims, c0, "fp32", args.dt, lay["fp32"])
    prep32 = s32.Prepare(args.steps)    # autotuner and step-graph capture, outside the timed region
    s32.StepperPropagate(args.warmup)
    el32, dev32 = timer.time(s32, args.steps)
    # per-pass device times over a second timed region
    tsteps = min(args.steps, 200)
    s32.EnableTiming(True)
    s32.StepperPropagate(tsteps)
    passes = s32.PassTimes()
    s32.EnableTiming(False)
    tuning = s32.Tuning()
    print(f"# fp32: {1e3 * el32 / args.steps:.4f} ms/step (events {1e3 * dev32 / args.steps:.4f}); "
          f"prepare {prep32:.1f} ms; passes {passes}", file=sys.stderr, flush=True)
    res = {}
    if not args.no_fp64:
        s64 = build_solver(G, W, n, dims, c0, "fp64", args.dt, lay["fp64"])
        prep64 = s64.Prepare(args.steps)
        tuning64 = s64.Tuning()
        s64.StepperPropagate(args.warmup)
        el64, dev64 = timer.time(s64, args.steps)
        del s64
        print(f"# fp64: {1e3 * el64 / args.steps:.4f} ms/step (events {1e3 * dev64 / args.steps:.4f}); "
              f"prepare {prep64:.1f} ms", file=sys.stderr, flush=True)
        # accuracy at the stated step count, from the same initial field
        del s32
        e32 = build_solver(G, W, n, dims, c0, "fp32", args.dt, tuning)
        e32.StepperPropagate(args.err_steps)
        f32 = e32.Field(0).real
        _, m32 = e32.Energy(0, W.CH_BULK_POLY, 1.0)
        del e32
        e64 = build_solver(G, W, n, dims, c0, "fp64", args.dt, tuning64)
        e64.StepperPropagate(args.err_steps)
        f64 = e64.Field(0).real
        _, m64 = e64.Energy(0, W.CH_BULK_POLY, 1.0)
        del e64
        m0 = float(np.sum(c0))
        res = {
            "fp64": {"value": n * args.steps / el64, "ms_per_step": 1e3 * el64 / args.steps,
                     "ms_per_step_events": 1e3 * dev64 / args.steps, "prepare_ms": prep64},
            "speedup_fp32_vs_fp64": el64 / el32,
            "speedup_fp32_vs_fp64_events": dev64 / dev32,
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
    traffic, why = pmc_traffic(f"{workload}_fp32", name, tuning)
    roof = {"bound": "hbm", "kernel": name, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
            "bytes_per_launch": algo,
            "avg_launch_ms": ms / cnt, "lines_per_block": tuning,
            "layout": "stored (gopf_amd/tuning/gfx950.json)" if lay["fp32"] else "autotuned in this run",
            "passes_ms_per_launch": {k: v[0] / v[1] for k, v in passes.items()},
            "passes_GBs": {k: pass_bytes(k, N, N, N, 4) / (v[0] / v[1] * 1e-3) / 1e9 for k, v in passes.items()},
            "step_effective_GBs": step_bytes * args.steps / el32 / 1e9,
            "step_frac": step_bytes * args.steps / el32 / 1e9 / HBM_PEAK_GBS,
            "measured_GBs": bw, "frac_of_measured_copy": achieved / bw["copy"],
            "step_frac_of_measured_copy": step_bytes * args.steps / el32 / 1e9 / bw["copy"]}
    if why:
        roof["traffic_note"] = why
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
                   "precision": "fp32 arrays and transforms; fp64 k=0 mode (mean), fp64-built "
                                "implicit denominators and update multipliers (rounded to fp32), "
                                "fp64 reductions"},
        "roofline": roof, "cpu_baseline": cpu,
        "ms_per_step_events": 1e3 * dev32 / args.steps, "prepare_ms": prep32,
        "passes_sum_ms": sum(v[0] / v[1] for v in passes.values()),
    }
    out.update(res)
    if args.save_tuning and not args.no_fp64:
        with open(args.save_tuning, "w") as f:
            json.dump({workload: {"fp32": tuning, "fp64": tuning64}}, f, indent=1)
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
