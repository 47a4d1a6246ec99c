"""Benchmark: pressure-Poisson solves per second on one MI355X.

A step is one complete solve — rhs resident in HBM in Cubism's 8^3 block
layout, pressure written back in that layout — to the true relative residual
||b - A x|| / ||b|| <= 1e-7 (recomputed in FP64 on the device). The steps are
consecutive time steps of a moving-blob sequence, each warm-started from the
previous step's (mean-free) pressure, as PressureProjection feeds the solver.

Default workload: the 4-level AMR "school" grid (base 16x8x8 blocks, domain
[0,2]x[0,1]^2, walls, refined up to 3 levels near the surfaces of 50
ellipsoids: 59 012 blocks, 30.2 M cells), rhs = divergence of a moving
Gaussian-blob velocity with penalised ellipsoid interiors; the previous
pressure is the solution of the step before (blob centres 0.005 earlier).
--workload c2 selects the uniform 256x128x128 walls grid (256 blobs and a
penalised sphere), c3 the 3-level AMR sphere grid, c1 the 32^3 periodic
Fourier case. The forms timed on the same steps, each with the same
iteration cap (3000), each run the way the caller runs its time steps: every
solve is handed the form's own solution of the step before, its
volume-weighted mean removed (PressureProjection.cpp:325-431):
  * fp64: the FP64 drop-in (the reference algorithm, FP64 MFMA block-Jacobi
    GEMM); its steps that do not reach the tolerance are counted;
  * prev_guess: the chosen form — FP16 split-precision MFMA GEMM with FP32
    accumulation, FP32 Krylov vectors, FP64 iterative refinement — from the
    previous pressure;
  * learned_only: the same from the learned warm start (a 3-layer CNN
    trained, untimed, by the library's HIP kernels on the (rhs, p_prev, p)
    triples of the --train steps before the timed ones), previous pressure
    as the fallback;
  * extrapolation_linear: the same from the linear extrapolation in time of
    the previous two pressures (no net);
  * extrapolation_only: the same from the least-squares combination of the
    differences of the last --history (6) pressures (no net): the
    extrapolations in time up to that degree are in its span, the depth
    taken is scored against the rounding of the kept pressures;
  * the headline `value`: the learned guess and that extrapolation
    together, weighted by FP64 least squares on their residuals (fallback:
    the previous pressure); the handle's history is primed with the steps
    before the warm-up, so every timed solve sees its full history;
  * tight: the headline form to ||b - A x|| <= 1e-8 ||b||, the residual at
    which the pressure meets the 1e-6 relative-L2 bar.
Every solve's pressure is compared with the reference solution of its step
(FP64 true residual <= 1e-10, by the FP64-refined form: on the 625-block
school grid a 1e-10 solve is 8e-9 from the discrete solution, a 1e-9 solve
2.4e-7).

The line carries the roofline of the dominant kernel (HIP events on the
library's stream inside the timed solves) and of the other four, and a CPU
baseline (the oracle's restatement of the reference CPU solver on a bounded
sample of the same workload).

  python bench.py [--steps K] [--warmup W] [--form mixed_fp16|mixed_bf16|fp64]
                  [--workload c1|c2|c3|c4|c5] [--tol 1e-7]
  python bench.py --workload c4        the learned-warm-start protocol (C4)
  python bench.py --workload diff2     implicit diffusion steps on the uniform grid
                                       (diff3: the 3-level AMR sphere grid)
  python bench.py --cpu-full           complete CPU-baseline solves on C1, C2, C3
"""
from __future__ import annotations

import os

# The CPU baseline's OpenMP threads: one per core, packed (launch/launchDaint.sh
# binds the reference this way). Set before any library starts its runtime.
os.environ.setdefault("OMP_PROC_BIND", "close")
os.environ.setdefault("OMP_PLACES", "cores")

import argparse  # noqa: E402
import ctypes  # noqa: E402
import json  # noqa: E402
import sys  # noqa: E402
import time  # noqa: E402

import numpy as np  # noqa: E402

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Poisson solves/sec at rel. residual 1e-7 (mixed+learned vs FP64); HBM GB/s"
HBM_PEAK = 8000.0       # GB/s, MI355X
FP64_MFMA_PEAK = 78.6   # TFLOP/s, dense FP64 matrix
F16_MFMA_PEAK = 2500.0  # TFLOP/s, dense FP16/BF16 matrix
MAX_ITER = 3000         # iteration cap of every timed form
REF_TOL = 1e-10         # true relative residual of the untimed reference solves
CPU_MAX_ITER = 1000     # the reference CPU solver's cap (PoissonSolverAMR.cpp:127)
KERNELS = {0: "spmv", 1: "precond_gemm", 2: "update_K5", 3: "precond_update_K1",
           4: "spmv_dot_K2"}
# Algorithmic FP32 words per cell of the mixed form's five kernels: vectors read
# + written (K1 r, nu, p -> p, z1; K2 z1, rhat -> nu; K3 r, nu -> r, z2;
# K4 z2, r -> t; K5 x, z1, z2, t, r, rhat -> x, r).
WORDS = {"K1": 5, "K2": 3, "K3": 4, "K4": 3, "K5": 8}
COPY_BYTES = 256 << 20  # calibration copy for the FETCH_SIZE / WRITE_SIZE counters
# HBM bytes per launch from two rocprofv3 counter passes (FETCH_SIZE and
# WRITE_SIZE, each scaled by what it reads for the 256 MiB copy of
# --kernels-only); the summaries are kept in profiles/ (file named per entry).
TRAFFIC = {
    ("amr4_school50_blobs64_sequence", "mixed_fp16", "K1"):
        (611.8e6, "profiles/mixed_fp16_amr4_school_traffic.md"),
    ("amr4_school50_blobs64_sequence", "mixed_fp16", "K2"):
        (392.9e6, "profiles/mixed_fp16_amr4_school_traffic.md"),
    ("amr4_school50_blobs64_sequence", "mixed_fp16", "K5"):
        (971.88e6, "profiles/mixed_fp16_amr4_school_traffic.md"),
    ("uniform_256x128x128_walls_blobs256_sphere", "mixed_fp16", "K1"):
        (88.25e6, "profiles/mixed_fp16_uniform_256x128x128_traffic.md"),
    ("uniform_256x128x128_walls_blobs256_sphere", "mixed_fp16", "K2"):
        (53.10e6, "profiles/mixed_fp16_uniform_256x128x128_traffic.md"),
    ("uniform_256x128x128_walls_blobs256_sphere", "mixed_fp16", "K5"):
        (134.3e6, "profiles/mixed_fp16_uniform_256x128x128_traffic.md"),
}


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=16)
    p.add_argument("--warmup", type=int, default=1)
    p.add_argument("--form", default="mixed_fp16", choices=["fp64", "mixed_fp16", "mixed_bf16"])
    p.add_argument("--workload", default="c5",
                   choices=["c1", "c2", "c3", "c4", "c5", "diff2", "diff3"],
                   help="c5: 4-level AMR school grid (default); c2: uniform 256x128x128 walls; "
                        "c3: 3-level AMR sphere; c1: 32^3 periodic; c4: the learned-warm-start "
                        "protocol on the uniform blob sequence")
    p.add_argument("--bmean", type=int, default=0)
    p.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    p.add_argument("--cpu-iters", type=int, default=4,
                   help="iterations per timed batch of the CPU baseline (5 batches)")
    p.add_argument("--cpu-full", action="store_true",
                   help="complete CPU-baseline solves (median of 5 after 1 warm-up) on C1, "
                        "C2 and C3 beside the GPU forms on the same inputs, then exit")
    p.add_argument("--tol", type=float, default=1e-7,
                   help="true relative residual every timed solve must reach")
    p.add_argument("--tight-tol", type=float, default=1e-8,
                   help="tolerance of the `tight` leg (0 = skip it)")
    p.add_argument("--train", type=int, default=24,
                   help="sequence steps before the timed ones whose (rhs, p_prev, p) triples "
                        "train the warm-start net")
    p.add_argument("--train-steps", type=int, default=4000, help="Adam steps of the net")
    p.add_argument("--no-learned", action="store_true",
                   help="headline form without the learned warm start")
    p.add_argument("--c4-steps", type=int, default=400)
    p.add_argument("--c4-train", type=int, default=300)
    p.add_argument("--c4-fresh", type=int, default=40)
    p.add_argument("--history", type=int, default=6,
                   help="previous pressures the warm start extrapolates from (1 = linear, "
                        "at most 8)")
    p.add_argument("--no-dst", action="store_true",
                   help="skip the leg with the sine-transform block inverse")
    p.add_argument("--fp64-dst-steps", type=int, default=4,
                   help="timed steps of the FP64 form with the sine-transform block inverse")
    p.add_argument("--nu", type=float, default=1e-3,
                   help="diffusion workloads: kinematic viscosity")
    p.add_argument("--dt", type=float, default=0.01, help="diffusion workloads: time step")
    p.add_argument("--kernels-only", action="store_true",
                   help="launch the timed kernels a few times and a 256 MiB copy (for "
                        "hardware-counter runs), print their names, and exit")
    return p.parse_args()


def make_workload(name):
    from cup3d_amd import workloads as W
    return {"c1": W.c1_periodic_32, "c2": W.c2_uniform_walls, "c3": W.c3_amr_sphere,
            "c5": W.c5_amr_school}[name]()


def in_solve_profile(S, enable=None):
    """HIP-event durations (us) of the main kernels of K1 .. K5 inside the
    mixed form's solves since the last call (cup3d_debug_profile*), with the
    sample count."""
    from cup3d_amd.poisson import lib
    L = lib()
    if enable is not None:
        L.cup3d_debug_profile.argtypes = [ctypes.c_void_p, ctypes.c_int]
        if L.cup3d_debug_profile(S.h, int(enable)) != 0:
            raise RuntimeError(L.cup3d_poisson_last_error(S.h).decode())
        return None
    L.cup3d_debug_profile_read.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_double),
                                           ctypes.POINTER(ctypes.c_longlong)]
    us = (ctypes.c_double * 5)()
    n = ctypes.c_longlong()
    if L.cup3d_debug_profile_read(S.h, us, ctypes.byref(n)) != 0:
        raise RuntimeError("cup3d_debug_profile_read failed")
    out = {f"K{k + 1}": us[k] for k in range(5)}
    out["samples"] = n.value
    return out


def time_form(w, form, bmean, seq, warmup, err_fn, weights=None, profile=False, tol=1e-7,
              extrapolate=False, log=None, precond="gemm"):
    """Solve every step of the sequence with `form` the way the caller runs
    its time steps (PressureProjection.cpp:325-431): each solve is handed the
    form's own solution of the step before, its volume-weighted mean removed
    (the first from the reference pressure of the step before the sequence);
    given net weights, the learned guess with that pressure as the fallback;
    with `extrapolate`, also the extrapolation in time from the pressures
    handed to the solves before. Time the steps after the first `warmup`.
    Returns the handle and, per timed step, (ms, report, error against the
    step's reference solution). With `profile`, the library also times K1 ..
    K5 with HIP events on its stream during the timed solves (one iteration
    in eight) and the handle carries them as `prof`."""
    import torch
    from cup3d_amd.poisson import solver_for
    S = solver_for(w, tol_rel=0.0, tol_rhs=tol, precision=form, bmean=bmean, max_iter=MAX_ITER,
                   guess="given" if weights is None else "cnn", extrapolate=extrapolate,
                   precond=precond)
    if weights is not None:
        S.cnn_set_weights(weights)
    if profile:
        in_solve_profile(S, True)
    x = seq[0]["p_prev"].clone()
    out = []
    for k, st in enumerate(seq):
        if profile and k == warmup:
            in_solve_profile(S)  # drop the warm-up samples
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        rep = S.solve_device(st["rhs"], x)
        dt = time.perf_counter() - t0
        S.remove_mean(x)  # the caller's projection after the solve; x is the next guess
        if k >= warmup:
            out.append((dt * 1e3, rep, err_fn(x, st["x_ref"])))
        if log:
            print(f"  {log} step {k}: {rep['iterations']} iterations, {dt * 1e3:.1f} ms",
                  file=sys.stderr, flush=True)
    if profile:
        S.prof = in_solve_profile(S)
        in_solve_profile(S, False)
    return S, out


def converged(rep, tol):
    return rep["status"] == 0 and rep["true_rel_resid"] <= tol


def summary(res, tol):
    ms = [r[0] for r in res]
    reps = [r[1] for r in res]
    errs = [r[2] for r in res]
    conv = [converged(r, tol) for r in reps]
    return {"solves_per_s": round(1e3 / np.mean(ms), 4), "ms_per_step": round(float(np.mean(ms)), 3),
            "iterations_mean": float(np.mean([r["iterations"] for r in reps])),
            "iterations": [r["iterations"] for r in reps],
            "converged": f"{sum(conv)}/{len(conv)}",
            "true_rel_resid_max": max(r["true_rel_resid"] for r in reps),
            "err_vs_reference_max": max(errs),
            "err_vs_reference_median": float(np.median(errs))}


def kernel_times(S, which_list, reps=20):
    from cup3d_amd.poisson import lib
    L = lib()
    f = L.cup3d_debug_time_kernel
    f.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int] + [ctypes.POINTER(ctypes.c_double)] * 3
    out = {}
    for which in which_list:
        a, b, c = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        if f(S.h, which, reps, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)) != 0:
            raise RuntimeError(L.cup3d_poisson_last_error(S.h).decode())
        sec = a.value * 1e-3
        out[which] = {"avg_us": a.value * 1e3, "bytes": b.value, "flops": c.value,
                      "GB/s": b.value / sec / 1e9, "TFLOP/s": c.value / sec / 1e12}
    return out


def copy_ceiling(dev, nbytes=1 << 30, reps=10):
    """Device-to-device copy rate (read + write bytes / s) over 1 GiB, the
    practical HBM ceiling next to the 8 TB/s datasheet peak: the faster of
    torch's copy and the library's nontemporal streaming copy
    (cup3d_debug_copy_rate). Returns (ceiling, torch rate, streaming rate)
    in GB/s."""
    import torch
    from cup3d_amd.poisson import lib
    a = torch.ones(nbytes // 8, dtype=torch.float64, device=dev)
    b = torch.empty_like(a)
    b.copy_(a)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        b.copy_(a)
    e1.record()
    e1.synchronize()
    t = round(2 * nbytes * reps / (e0.elapsed_time(e1) * 1e-3) / 1e9, 1)
    del a, b
    f = lib().cup3d_debug_copy_rate
    f.argtypes = [ctypes.c_int, ctypes.c_int64, ctypes.c_int]
    f.restype = ctypes.c_double
    sr = f(dev.index or 0, nbytes, reps)
    st = round(sr / 1e9, 1) if sr > 0 else None
    return max(t, st or 0.0), t, st


def host_solve_ms(w, form, bmean, tol, rhs, p_prev):
    """One solve through the host entry point (H2D of rhs and guess, D2H of the
    pressure over PCIe included) and the same solve device-resident, on a
    fresh handle (no history, no net), so the two differ only by the host
    path."""
    import torch
    from cup3d_amd.poisson import solver_for
    S = solver_for(w, tol_rel=0.0, tol_rhs=tol, precision=form, bmean=bmean, max_iter=MAX_ITER)
    x = p_prev.clone()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    S.solve_device(rhs, x)
    torch.cuda.synchronize()
    dev_ms = (time.perf_counter() - t0) * 1e3
    r = rhs.cpu().numpy()
    xh = p_prev.cpu().numpy().copy()
    t0 = time.perf_counter()
    rep = S.solve_host(r, xh)
    host_ms = (time.perf_counter() - t0) * 1e3
    return round(host_ms, 3), round(dev_ms, 3), rep["status"]


_HOST_CPUS = None


def physical_cores():
    """The value of host_cpus() taken when bench.py started, before an OpenMP
    runtime bound this thread to one place."""
    return _HOST_CPUS


def host_cpus():
    """Physical cores (distinct (socket, core) pairs of `lscpu -p`, or of
    /proc/cpuinfo) among the CPUs this process may run on, with the count of
    those CPUs: {"physical": P, "logical_allowed": L}."""
    import subprocess
    try:
        allowed = os.sched_getaffinity(0)
    except OSError:
        return None
    cores = set()
    try:
        out = subprocess.run(["lscpu", "-p=CPU,CORE,SOCKET"], capture_output=True, text=True,
                             timeout=10).stdout
        for line in out.splitlines():
            if line.startswith("#"):
                continue
            cpu, core, sock = (int(v) if v else -1 for v in line.split(","))
            if cpu in allowed:
                cores.add((sock, core))
    except (OSError, ValueError, subprocess.SubprocessError):
        cores = set()
    return {"physical": len(cores) or None, "logical_allowed": len(allowed)}


def cpu_threads():
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    return pyoracle.num_threads(fast=True)


def cpu_baseline(w, bmean, rhs, x0, scale_iters, scale_what, iters):
    """Reference CPU algorithm (pipelined BiCGSTAB + per-block CG), restated in
    oracle/oracle.cpp and built with the reference's release flags, on the
    full workload: one warm-up iteration, then 5 batches of `iters`
    iterations; the median seconds per iteration times `scale_iters`."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    O = pyoracle.Oracle(w.blocks, w.bpd, w.h0, w.bc, bmean, fast=True)
    O.pipelined(rhs, x0, tol_rel=0.0, max_iter=1)  # first touch of every vector
    per = []
    for _ in range(5):
        t0 = time.perf_counter()
        O.pipelined(rhs, x0, tol_rel=0.0, max_iter=iters)
        per.append((time.perf_counter() - t0) / iters)
    dt = float(np.median(per))
    return {
        "value": round(1.0 / (dt * scale_iters), 5),
        "unit": "solves/s",
        "cores": pyoracle.num_threads(fast=True),
        "host_cpus": physical_cores(),
        "omp": {k: os.environ.get(k) for k in ("OMP_NUM_THREADS", "OMP_PROC_BIND", "OMP_PLACES")},
        "kind": "port",
        "sample": (f"pipelined BiCGSTAB + block-CG CPU solver on the full {w.ncells}-cell "
                   f"workload from the last timed step's guess: 1 warm-up iteration, then the "
                   f"median of 5 batches of {iters} iterations ({dt:.4f} s/iteration), times "
                   f"{scale_iters} iterations ({scale_what})"),
        "seconds_per_iteration": round(dt, 5),
    }


def reference_sequence(w, bmean, nsteps, dev, seed=None):
    """Untimed setup: the reference solution of steps 0..nsteps, each from the
    previous step's pressure, ||b - Ax|| <= REF_TOL ||b|| on the FP64 true
    residual by the FP64-refined solve; the caller removes the
    volume-weighted mean after every solve (PressureProjection.cpp:404-431),
    so every guess is mean-free. Returns the steps 1..nsteps as dicts (rhs,
    p_prev, x_ref), the reference residuals and the count that converged."""
    import torch
    from cup3d_amd import workloads as W
    from cup3d_amd.poisson import solver_for
    R = solver_for(w, tol_rel=0.0, tol_rhs=REF_TOL, bmean=bmean, precision="mixed_fp16",
                   max_iter=2000, max_refine=12)
    p_prev = torch.zeros(w.ncells, dtype=torch.float64, device=dev)
    seq, ref_res, ref_ok = [], [], 0
    for t in range(nsteps + 1):
        b = torch.from_numpy(w.rhs(step=t, seed=W.SEED if seed is None else seed)
                             .reshape(-1)).to(dev)
        x = p_prev.clone()
        r = R.solve_device(b, x)
        R.remove_mean(x)
        if t > 0:
            seq.append({"rhs": b, "p_prev": p_prev, "x_ref": x})
            ref_res.append(r["true_rel_resid"])
            ref_ok += r["status"] == 0
        p_prev = x
        print(f"setup: reference solve of step {t}/{nsteps} done", file=sys.stderr, flush=True)
    return seq, ref_res, ref_ok


def rel_err_fn(w, dev):
    import torch
    hw = torch.from_numpy(np.repeat((w.h0 / 2.0 ** w.blocks[:, 0]) ** 3, 512)).to(dev)

    def rel(a, b):
        """Relative L2 distance after removing the volume-weighted means (the
        caller's post-solve projection)."""
        a = a - (hw * a).sum() / hw.sum()
        b = b - (hw * b).sum() / hw.sum()
        return float(torch.linalg.norm(a - b) / torch.linalg.norm(b))
    return rel


def bench_cpu_full(args):
    """Complete solves of the CPU baseline (pipelined BiCGSTAB + block CG,
    reference release flags) to the same true relative residual as the GPU
    forms, on step 1 of each workload from step 0's pressure (C1: zero guess):
    median of 5 after 1 warm-up; the GPU forms on the same inputs beside it."""
    import torch
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    from cup3d_amd.poisson import solver_for
    dev = torch.device("cuda:0")
    out = {"metric": "CPU baseline: complete solves at ||b-Ax||/||b|| <= %g" % args.tol,
           "cores": pyoracle.num_threads(fast=True),
           "host_cpus": physical_cores(),
           "omp": {k: os.environ.get(k) for k in ("OMP_NUM_THREADS", "OMP_PROC_BIND",
                                                  "OMP_PLACES")},
           "kind": "port", "workloads": {}}
    for name in ("c1", "c2", "c3"):
        w = make_workload(name)
        if name == "c1":
            b = torch.from_numpy(w.rhs(step=1).reshape(-1)).to(dev)
            st = {"rhs": b, "p_prev": torch.zeros_like(b), "x_ref": None}
        else:
            seq, _, _ = reference_sequence(w, args.bmean, 1, dev)
            st = seq[0]
        rhs, x0 = st["rhs"].cpu().numpy(), st["p_prev"].cpu().numpy()
        bn = float(np.linalg.norm(rhs))
        O = pyoracle.Oracle(w.blocks, w.bpd, w.h0, w.bc, args.bmean, fast=True)
        times, its = [], []
        for k in range(6):
            t0 = time.perf_counter()
            x, info = O.pipelined(rhs, x0, tol_abs=args.tol * bn, tol_rel=0.0,
                                  max_iter=MAX_ITER)
            dt = time.perf_counter() - t0
            if k > 0:
                times.append(dt)
                its.append(info["iters"])
            print(f"cpu {name} solve {k}: {info['iters']} iterations, {dt:.2f} s",
                  file=sys.stderr, flush=True)
        Os = pyoracle.Oracle(w.blocks, w.bpd, w.h0, w.bc, args.bmean)
        tr = float(np.linalg.norm(rhs - Os.spmv(x)) / bn)
        gpu = {}
        for form in ("fp64", "mixed_fp16"):
            S = solver_for(w, tol_rel=0.0, tol_rhs=args.tol, precision=form, bmean=args.bmean,
                           max_iter=MAX_ITER)
            xs = st["p_prev"].clone()
            S.solve_device(st["rhs"], xs)  # warm-up
            ms = []
            for _ in range(5):
                xs.copy_(st["p_prev"])
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                rep = S.solve_device(st["rhs"], xs)
                ms.append((time.perf_counter() - t0) * 1e3)
            gpu[form] = {"ms_median": round(float(np.median(ms)), 3),
                         "iterations": rep["iterations"], "status": rep["status"],
                         "true_rel_resid": rep["true_rel_resid"]}
        out["workloads"][w.name] = {
            "cells": w.ncells, "guess": "zero" if name == "c1" else "previous step's pressure",
            "cpu_seconds_median": round(float(np.median(times)), 3),
            "cpu_solves_per_s": round(1.0 / float(np.median(times)), 5),
            "cpu_iterations": its, "cpu_true_rel_resid": tr, "gpu": gpu,
            "gpu_mixed_vs_cpu": round(float(np.median(times)) * 1e3 /
                                      gpu["mixed_fp16"]["ms_median"], 1)}
    print(json.dumps(out))


def bench_c4(args):
    """The learned-warm-start protocol on the uniform 256x128x128 blob
    sequence (cup3d_amd.learned): train on the first --c4-train triples,
    evaluate the mixed form's iterations on the remaining steps and on a
    fresh-seed sequence against the previous-pressure guess."""
    from cup3d_amd import learned
    r = learned.run(steps=args.c4_steps, train=args.c4_train, fresh=args.c4_fresh,
                    forms=("mixed_fp16",), history=args.history)
    held, fresh = r["held_out"]["mixed_fp16"], r["fresh_seed"]["mixed_fp16"]
    out = {"metric": "learned warm start: iterations saved vs the previous-pressure guess",
           "value": held["iterations_saved_pct"], "unit": "% iterations (held-out steps)",
           "higher_is_better": True, "n_gpus": 1, "data": "synthetic",
           "config": {"workload": r["workload"], "cells": r["cells"], "steps": r["steps"],
                      "train_triples": r["train_triples"], "fresh_steps": args.c4_fresh,
                      "form": "mixed_fp16", "tolerance": "||b-Ax||/||b|| <= 1e-7"},
           **r}
    print(json.dumps(out))


def diffusion_kernel_times(S, nu, dt, reps=20):
    from cup3d_amd.poisson import lib
    L = lib()
    f = L.cup3d_debug_time_diffusion_kernel
    f.argtypes = ([ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_int]
                  + [ctypes.POINTER(ctypes.c_double)] * 3)
    out = {}
    for which, name in ((0, "helmholtz_spmv"), (1, "dst_precond_update_K1"),
                        (2, "dst_precond")):
        a, b, c = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        if f(S.h, which, nu, dt, reps, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)) != 0:
            raise RuntimeError(L.cup3d_poisson_last_error(S.h).decode())
        sec = a.value * 1e-3
        out[name] = {"avg_us": round(a.value * 1e3, 2), "alg_bytes_per_launch": b.value,
                     "GB/s": round(b.value / sec / 1e9, 1),
                     "frac": round(b.value / sec / 1e9 / HBM_PEAK, 4)}
    return out


def bench_diffusion(args):
    """Implicit diffusion (AdvectionDiffusionImplicit::euler after its
    advection kernel): one step = the rhs of the three velocity components,
    three Helmholtz solves (FP64, DST block-inverse preconditioner) to the
    reference's default tolerances (||r|| < 1e-6 or ||r||/||r0|| < 1e-4) and
    the update u += x, on [n][3] velocities resident in HBM. u^n and u* are
    the blob velocity of two consecutive steps of the moving-blob sequence;
    every timed step starts from the same u*."""
    import torch
    from cup3d_amd import workloads as W
    from cup3d_amd.poisson import solver_for
    dev = torch.device("cuda:0")
    w = W.c2_uniform_walls() if args.workload == "diff2" else W.c3_amr_sphere()
    vb = dict(sphere=W.SPHERE)
    u_old = torch.from_numpy(W.vel_blobs(w.blocks, w.h0, w.bpd, w.bc, 256, (0.02, 0.08),
                                         **vb)).to(dev)
    u_star = torch.from_numpy(W.vel_blobs(w.blocks, w.h0, w.bpd, w.bc, 256, (0.02, 0.08),
                                          step=1, **vb)).to(dev)
    S = solver_for(w, bmean=0)
    v = torch.empty_like(u_star)
    tol_abs, tol_rel = 1e-6, 1e-4  # SimulationData.cpp:92-93
    ms, its, reps_all = [], [], []
    for k in range(args.warmup + args.steps):
        v.copy_(u_star)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        reps = S.diffusion_step(u_old, v, args.nu, args.dt, tol_abs, tol_rel)
        dt_ = time.perf_counter() - t0
        if k >= args.warmup:
            ms.append(dt_ * 1e3)
            its.append([r["iterations"] for r in reps])
            reps_all.append(reps)
    v_gpu = v.cpu().numpy()
    kern = diffusion_kernel_times(S, args.nu, args.dt)
    dom = max(kern, key=lambda k: kern[k]["avg_us"] if k != "dst_precond" else 0.0)
    kd = kern[dom]
    N = float(w.ncells)
    msm = float(np.mean(ms))
    it_sum = float(np.mean([sum(i) for i in its]))
    out = {
        "metric": "implicit diffusion steps/s (3 Helmholtz solves per step, FP64)",
        "value": round(1e3 / msm, 3), "unit": "steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(msm, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"{w.name}_implicit_diffusion", "cells": w.ncells,
                   "blocks": int(len(w.blocks)), "nu": args.nu, "dt": args.dt,
                   "h2_over_nu_dt_finest": float((w.h0 / 2 ** w.level_max) ** 2 /
                                                 (args.nu * args.dt)),
                   "tolerance": "||r|| < 1e-6 or ||r||/||r0|| < 1e-4 (reference defaults)",
                   "parallelism": "single GPU"},
        "iterations_per_component": its[-1],
        "cell_iterations_per_s": round(N * it_sum / (msm * 1e-3), 1),
        "solve_ms_per_component": [round(r["solve_ms"], 3) for r in reps_all[-1]],
        "roofline": {"bound": "hbm", "achieved": kd["GB/s"], "peak": HBM_PEAK, "unit": "GB/s",
                     "frac": kd["frac"], "traffic": None, "kernel": dom,
                     "avg_us": kd["avg_us"], "alg_bytes_per_launch": kd["alg_bytes_per_launch"],
                     "timing": "HIP events around back-to-back launches", "kernels": kern},
    }
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import pyoracle
        O = pyoracle.Oracle(w.blocks, w.bpd, w.h0, w.bc, 0, fast=True)
        uo, us = u_old.cpu().numpy(), u_star.cpu().numpy()
        times = []
        for k in range(3):
            t0 = time.perf_counter()
            v_cpu, info = O.diffusion_step(uo, us, args.nu, args.dt, tol_abs, tol_rel)
            times.append(time.perf_counter() - t0)
        dcpu = float(np.median(times))
        out["cpu_baseline"] = {
            "value": round(1.0 / dcpu, 5), "unit": "steps/s",
            "cores": pyoracle.num_threads(fast=True), "kind": "port",
            "host_cpus": physical_cores(),
            "sample": (f"the whole step on the same {w.ncells}-cell inputs by the restated "
                       f"reference algorithm (pipelined BiCGSTAB + per-block CG, "
                       f"DiffusionSolver::solve), median of 3 ({dcpu:.2f} s per step), "
                       f"iterations {info['iters']}"),
            "gpu_vs_cpu_result_rel_l2": float(np.linalg.norm(v_gpu - v_cpu) /
                                              np.linalg.norm(v_cpu - us))}
    print(json.dumps(out))


def main():
    global _HOST_CPUS
    _HOST_CPUS = host_cpus()
    args = parse()
    if args.gpus != 1:
        print("bench.py measures one GPU (one MI355X per process); the solver has no "
              "multi-GPU path: how it would shard over GPUs is a design note, not code.",
              file=sys.stderr)
        sys.exit(2)
    import torch

    if args.cpu_full:
        return bench_cpu_full(args)
    if args.workload == "c4":
        return bench_c4(args)
    if args.workload in ("diff2", "diff3"):
        return bench_diffusion(args)
    from cup3d_amd.poisson import solver_for

    dev = torch.device("cuda:0")
    w = make_workload(args.workload)
    if args.kernels_only:
        S = solver_for(w, precision=args.form, bmean=args.bmean)
        k = kernel_times(S, [2, 3, 4], reps=5)
        a = torch.empty(COPY_BYTES // 8, dtype=torch.float64, device=dev).fill_(1.0)
        b = torch.empty_like(a)
        for _ in range(5):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"workload": w.name, "form": args.form, "copy_bytes": COPY_BYTES,
                          "kernels": {KERNELS[i]: v["bytes"] for i, v in k.items()}}))
        return
    learned = args.form != "fp64" and not args.no_learned
    ntrain = args.train if learned else 0
    seq, ref_res, ref_ok = reference_sequence(w, args.bmean, ntrain + args.warmup + args.steps,
                                              dev)
    rel = rel_err_fn(w, dev)

    # the warm-start net learns from the triples of the steps before the
    # timed ones (untimed setup, the library's own HIP training kernels)
    weights, learn = None, {}
    if learned:
        T = solver_for(w, precision=args.form, bmean=args.bmean, guess="cnn")
        for st in seq[:ntrain]:
            T.cnn_add_sample(st["rhs"], st["p_prev"], st["x_ref"])
        t0 = time.perf_counter()
        l0, _ = T.cnn_train(1, batch=256, lr=1e-3, seed=1, res_weight=100.0)
        _, l1 = T.cnn_train(args.train_steps - 1, batch=256, lr=1e-3, seed=2, res_weight=100.0)
        torch.cuda.synchronize()
        learn = {"train_triples": ntrain, "train_steps": args.train_steps, "batch_blocks": 256,
                 "train_seconds": round(time.perf_counter() - t0, 2), "loss_first": l0,
                 "loss_last": l1}
        weights = T.cnn_get_weights()
        del T
        print(f"setup: net trained on {ntrain} triples, loss {l0:.3e} -> {l1:.3e}",
              file=sys.stderr, flush=True)
    # the legs with the extrapolation keep the pressures of their last solves:
    # the history is primed with the steps before the warm-up (training
    # steps, untimed), so every timed solve has its full history
    prime = min(max(0, args.history - args.warmup), ntrain)
    seq_all = seq
    seq = seq_all[ntrain - prime:]
    warm = args.warmup + prime
    timed = seq[warm:]
    tol = args.tol

    def leg(name, form, **kw):
        S, res = time_form(w, form, args.bmean, seq, warm, rel, tol=kw.pop("tol", tol),
                           log=name, **kw)
        print(f"{name}: {np.mean([r[0] for r in res]):.2f} ms per step", file=sys.stderr,
              flush=True)
        return S, res

    # the FP64 drop-in, the same iteration cap as every other form; and the
    # same with its exact block inverse applied by sine transforms
    S64, res64 = leg("fp64", "fp64")
    res64_dst = None
    if args.fp64_dst_steps > 0:
        _, res64_dst = time_form(w, "fp64", args.bmean, seq[:warm + args.fp64_dst_steps],
                                 warm, rel, tol=tol, log="fp64_dst", precond="dst")
    legs = {}
    if args.form == "fp64":
        S, res = S64, res64
    elif not learned:
        S, res = leg(args.form, args.form, profile=True)
    else:
        _, legs["prev_guess"] = leg("prev_guess", args.form)
        _, legs["learned_only"] = leg("learned_only", args.form, weights=weights)
        _, legs["extrapolation_linear"] = leg("extrapolation_linear", args.form, extrapolate=1)
        _, legs["extrapolation_only"] = leg("extrapolation_only", args.form,
                                            extrapolate=args.history)
        S, res = leg("headline", args.form, weights=weights, extrapolate=args.history,
                     profile=True)
    # the headline form with the exact block inverse applied by sine
    # transforms in FP32 instead of the FP16/BF16 GEMM (cfg.precond = DST)
    res_dst = None
    if args.form != "fp64" and not args.no_dst:
        _, res_dst = leg("headline_dst", args.form, weights=weights,
                         extrapolate=args.history if learned else 0, precond="dst")
    tight = None
    if args.tight_tol > 0 and args.form != "fp64":
        _, rt = leg("tight", args.form, weights=weights, extrapolate=args.history if learned else 0,
                    tol=args.tight_tol)
        tight = {"tol": args.tight_tol, **summary(rt, args.tight_tol),
                 "pressure_within_1e-6": bool(max(r[2] for r in rt) <= 1e-6)}
    sm, sm64 = summary(res, tol), summary(res64, tol)
    reps = [r[1] for r in res]
    if not all(converged(r, tol) for r in reps):
        raise SystemExit(f"{args.form} did not reach the tolerance on every step: {sm}")
    ms = float(np.mean([r[0] for r in res]))
    # speed-up over the steps on which both the FP64 drop-in and the headline
    # form reached the tolerance
    both = [k for k in range(len(res)) if converged(res64[k][1], tol) and converged(reps[k], tol)]
    speedup = (round(sum(res64[k][0] for k in both) / sum(res[k][0] for k in both), 3)
               if both else None)
    fp64_conv_ms = [res64[k][0] for k in range(len(res64)) if converged(res64[k][1], tol)]
    if legs:
        base = legs["prev_guess"]

        def gains(res_):
            return {
                "initial_resid_ratio_median": float(np.median(
                    [r[1]["guess_resid"] / rp[1]["err_init"] for r, rp in zip(res_, base)])),
                "iterations_saved_vs_prev_guess_pct": round(100.0 * (
                    1.0 - sum(r[1]["iterations"] for r in res_) /
                    sum(r[1]["iterations"] for r in base)), 2),
                "speedup_vs_prev_guess": round(float(np.mean([r[0] for r in base])) /
                                               float(np.mean([r[0] for r in res_])), 3)}
        lo, eo = legs["learned_only"], legs["extrapolation_only"]
        learn.update({
            "net_weight_median": float(np.median([r["guess_weight_net"] for r in reps])),
            "time_weight_median": float(np.median([r["guess_weight_time"] for r in reps])),
            "net_guess_used": f"{sum(r['guess_weight_net'] != 0 for r in reps)}/{len(reps)}",
            **gains(res),
            # what the net adds on top of the extrapolation in time
            "net_iterations_saved_vs_extrapolation_only_pct": round(100.0 * (
                1.0 - sum(r["iterations"] for r in reps) /
                sum(r[1]["iterations"] for r in eo)), 2)})
        legs_out = {
            "prev_guess": summary(base, tol),
            "learned_only": {**summary(lo, tol), **gains(lo),
                             "net_guess_used": f"{sum(r[1]['guess_used'] == 2 for r in lo)}/"
                                               f"{len(lo)}"},
            "extrapolation_only": {**summary(eo, tol), **gains(eo),
                                   "history": args.history},
            "extrapolation_linear": {**summary(legs["extrapolation_linear"], tol),
                                     **gains(legs["extrapolation_linear"])},
        }
    rep, rhs = reps[-1], timed[-1]["rhs"]
    p_last = timed[-1]["p_prev"]

    host_ms, host_dev_ms, host_status = host_solve_ms(w, args.form, args.bmean, tol, rhs, p_last)
    # mesh change: set_topology again (host face tables, upload, workspace)
    t0 = time.perf_counter()
    S.set_topology(w.blocks)
    mesh_ms = round((time.perf_counter() - t0) * 1e3, 2)
    ceiling, copy_torch, copy_stream = copy_ceiling(dev)
    prof = getattr(S, "prof", None)
    N = float(w.ncells)
    if args.form != "fp64" and prof and prof["samples"] > 0:
        # every kernel of the iteration timed inside the timed solves (HIP
        # events on the library's stream around its main kernel)
        kr = {}
        for k, words in WORDS.items():
            by = 4.0 * words * N
            gbs = by / (prof[k] * 1e-6) / 1e9
            tr = TRAFFIC.get((w.name, args.form, k))
            kr[k] = {"avg_us": round(prof[k], 2), "alg_bytes_per_launch": by,
                     "GB/s": round(gbs, 1), "frac": round(gbs / HBM_PEAK, 4),
                     "frac_of_copy_ceiling": round(gbs / ceiling, 4),
                     "traffic": tr[0] if tr else None}
        # the dominant kernel: the largest share of an iteration's time
        dom = max(kr, key=lambda k: kr[k]["avg_us"])
        kd = kr[dom]
        tr = TRAFFIC.get((w.name, args.form, dom))
        roof = {"bound": "hbm", "achieved": kd["GB/s"], "peak": HBM_PEAK, "unit": "GB/s",
                "frac": kd["frac"], "traffic": tr[0] if tr else None,
                "traffic_source": tr[1] if tr else None,
                "kernel": dom, "avg_us": kd["avg_us"],
                "timing": f"HIP events around {dom}'s kernel inside the timed solves "
                          f"({prof['samples']} launches)",
                "alg_bytes_per_launch": kd["alg_bytes_per_launch"],
                "copy_ceiling_GB/s": ceiling, "frac_of_copy_ceiling": kd["frac_of_copy_ceiling"],
                "kernels": kr}
    else:
        kern = kernel_times(S, [2, 3, 4])
        dom = max(kern, key=lambda k: kern[k]["avg_us"])
        kd = kern[dom]
        gbs = kd["bytes"] / (kd["avg_us"] * 1e-6) / 1e9
        roof = {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK, "unit": "GB/s",
                "frac": round(gbs / HBM_PEAK, 4), "traffic": None, "kernel": KERNELS[dom],
                "avg_us": round(kd["avg_us"], 2),
                "timing": "HIP events around back-to-back launches",
                "alg_bytes_per_launch": kd["bytes"], "copy_ceiling_GB/s": ceiling,
                "frac_of_copy_ceiling": round(gbs / ceiling, 4)}
    kern_mixed = kernel_times(S, [1]) if args.form != "fp64" else {}
    k64 = kernel_times(S64, [1, 2, 3, 4])
    g64 = k64[1]
    out = {
        "metric": METRIC,
        "value": round(1e3 / ms, 3),
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"fp64": "f64", "mixed_fp16": "f32+f16mfma(f64 refine)",
                  "mixed_bf16": "f32+bf16mfma(f64 refine)"}[args.form],
        "data": "synthetic",
        "config": {"workload": w.name, "form": args.form, "cells": w.ncells,
                   "blocks": int(len(w.blocks)), "tolerance": f"||b-Ax||/||b|| <= {tol:g}",
                   "steps": "consecutive time steps t = warmup+1 .. warmup+steps, each solve "
                            "handed the form's own previous solution, mean removed",
                   "warm_start": (f"learned CNN guess + extrapolation in time from up to "
                                  f"{args.history} previous pressures, FP64 least-squares "
                                  "weights, depth scored against the history's rounding; "
                                  "previous-step pressure as fallback"
                                  if learned else "previous-step pressure"),
                   "bMeanConstraint": args.bmean, "max_iter": MAX_ITER,
                   "parallelism": "single GPU"},
        "iterations": sm["iterations_mean"],
        "refinements": rep["refinements"],
        "cell_iterations_per_s": round(N * sm["iterations_mean"] / (ms * 1e-3), 1),
        "solves": sm,
        "pressure_err_vs_reference_max": sm["err_vs_reference_max"],
        "pressure_within_1e-6": bool(sm["err_vs_reference_max"] <= 1e-6),
        "reference": {"form": f"mixed_fp16 refined to ||b-Ax|| <= {REF_TOL:g} ||b|| (FP64 "
                              "true residual): the discrete solution to within that residual",
                      "converged": f"{ref_ok}/{len(ref_res)}",
                      "true_rel_resid_max": max(ref_res)},
        "tight": tight,
        "fp64": {**sm64, "max_iter": MAX_ITER,
                 "ms_per_iteration": round(float(np.mean([r[0] / r[1]["iterations"]
                                                          for r in res64])), 4),
                 "solves_per_s_converged_steps": (round(1e3 / float(np.mean(fp64_conv_ms)), 4)
                                                  if fp64_conv_ms else None),
                 "gemm_TFLOP/s": round(g64["TFLOP/s"], 2),
                 "gemm_mfma_frac": round(g64["TFLOP/s"] / FP64_MFMA_PEAK, 4),
                 "kernels_us": {KERNELS[k]: round(v["avg_us"], 2) for k, v in k64.items()}},
        "fp64_dst": ({**summary(res64_dst, tol),
                      "ms_per_iteration": round(float(np.mean([r[0] / r[1]["iterations"]
                                                               for r in res64_dst])), 4),
                      "note": "FP64 drop-in with precond = DST: the same exact block inverse "
                              "by sine transforms instead of the P_inv GEMM"}
                     if res64_dst else None),
        "precond_dst": ({**summary(res_dst, tol),
                         "speedup_vs_gemm_form": round(ms / float(np.mean([r[0] for r in res_dst])),
                                                       3),
                         "note": "the headline form with cfg.precond = DST: the exact block "
                                 "inverse by sine transforms in FP32 instead of the "
                                 f"{args.form} GEMM"}
                        if res_dst else None),
        "speedup_vs_fp64": speedup,
        "speedup_vs_fp64_steps": f"{len(both)}/{len(res)} (both forms converged)",
        "per_iteration_speedup_vs_fp64": round((float(np.mean([r[0] for r in res64])) /
                                                sm64["iterations_mean"]) /
                                               (ms / sm["iterations_mean"]), 3),
        "mesh_change_ms": mesh_ms,
        "host_to_host": {"ms_per_solve": host_ms, "device_resident_ms": host_dev_ms,
                         "status": host_status,
                         "note": "the last step solved from the previous pressure on a fresh "
                                 "handle: cup3d_poisson_solve from host arrays (PCIe copies "
                                 "and the gather/scatter included) against the same solve "
                                 "device-resident"},
        "hbm_copy_GB/s": {"ceiling": ceiling, "torch_copy": copy_torch,
                          "streaming_copy": copy_stream},
        "precond_gemm_mfma": ({"TFLOP/s": round(kern_mixed[1]["TFLOP/s"], 1),
                               "frac_of_f16_peak": round(kern_mixed[1]["TFLOP/s"] /
                                                         F16_MFMA_PEAK, 4)}
                              if kern_mixed else None),
        "roofline": roof,
    }
    if learned:
        out["learned"] = learn
        out.update(legs_out)
    if not args.no_cpu:
        # the reference CPU solver stops after 1000 iterations
        # (PoissonSolverAMR.cpp:127), returning its best iterate
        conv_its = [r[1]["iterations"] for r in res64 if converged(r[1], tol)]
        scale = min(int(round(np.mean(conv_its))), CPU_MAX_ITER) if conv_its else CPU_MAX_ITER
        what = (f"the FP64 GPU drop-in's mean over its {len(conv_its)} converged steps (the "
                f"same algorithm), capped at the CPU solver's own 1000" if conv_its else
                "the CPU solver's own cap of 1000 iterations (PoissonSolverAMR.cpp:127): "
                "the FP64 GPU drop-in reached the tolerance on no step within 3000")
        out["cpu_baseline"] = cpu_baseline(w, args.bmean, rhs.cpu().numpy(),
                                           p_last.cpu().numpy(), scale, what, args.cpu_iters)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
