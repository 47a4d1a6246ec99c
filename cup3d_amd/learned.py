"""Learned warm start on a sequence of pressure solves: generate, train, evaluate.

The sequence is the uniform 256x128x128 walls workload with 64 Gaussian
blobs (sigma 0.04) moving 0.005 per step (``workloads.c4_uniform_sequence``).
Each step is solved to ||b - A x|| / ||b|| <= 1e-7 (FP64 true residual) from
the previous step's pressure by the FP64-refined solve (the FP64 drop-in's
recurrence stagnates above 1e-7 on some steps of this sequence); the first
``--train`` steps give the training triples (rhs_t, p_{t-1}, p_t), which the
library's own HIP kernels learn (FP32 forward/backward, Adam). The net is then evaluated on the remaining steps
and on a sequence from a fresh seed by what it saves: the number of
iterations, for the FP64 and the mixed-precision solve, against starting from
the previous pressure; also, once the handle holds `history` previous
pressures, the extrapolation in time from them (least squares over their
differences; history 1 = the linear extrapolation) alone and combined with
the net (least-squares weights on their FP64 residuals). The solve
always finishes to the same tolerance and falls back to the previous
pressure whenever the net's guess has the larger residual, so a worse net
costs one inference and two residuals, never accuracy.

  python -m cup3d_amd.learned [--steps 400] [--train 300] [--fresh 40]
                              [--train-steps 20000] [--batch 256] [--res-weight 100]
                              [--bpd 32 16 16] [--workload c4|c5] [--fp64-max-iter N]

With --workload c5 the sequence is the moving-blob rhs on the 4-level AMR
school grid (30.2 M cells) instead.

Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np


def run(bpd=(32, 16, 16), steps=400, train=300, fresh=40, train_steps=20000, batch=256,
        lr=1e-3, res_weight=100.0, seed=1, tol=1e-7, log=sys.stderr, workload="c4",
        fp64_max_iter=3000, forms=("fp64", "mixed_fp16"), history=1):
    import torch

    from . import workloads as W
    from .poisson import solver_for

    dev = torch.device("cuda:0")
    w = W.c4_uniform_sequence(tuple(bpd)) if workload == "c4" else W.c5_amr_school()
    kw = dict(tol_rel=0.0, tol_rhs=tol, bmean=0, max_iter=3000)
    ref = solver_for(w, precision="mixed_fp16", max_refine=40, **kw)

    hw = torch.from_numpy(np.repeat((w.h0 / 2.0 ** w.blocks[:, 0]) ** 3, 512)).to(dev)

    def sequence(n, s):
        """Solutions p_0 .. p_n of steps 0..n, each from the previous pressure
        and, as the caller does after every solve (PressureProjection.cpp:
        404-431), with its volume-weighted mean removed."""
        rhs, sols, its = [], [], []
        p = torch.zeros(w.ncells, dtype=torch.float64, device=dev)
        for t in range(n + 1):
            b = torch.from_numpy(w.rhs(step=t, seed=s).reshape(-1)).to(dev)
            x = p.clone()
            rep = ref.solve_device(b, x)
            if rep["status"] != 0:
                raise RuntimeError(f"step {t}: solve did not converge: {rep}")
            ref.remove_mean(x)
            rhs.append(b)
            sols.append(x)
            its.append(rep["iterations"])
            p = x
            if t % 25 == 0:
                print(f"  step {t}/{n}: {rep['iterations']} iterations", file=log, flush=True)
        return rhs, sols, its

    t0 = time.time()
    rhs, sols, its = sequence(steps, W.SEED)
    print(f"sequence of {steps} solves: {time.time() - t0:.1f} s", file=log, flush=True)
    for t in range(1, train + 1):
        ref.cnn_add_sample(rhs[t], sols[t - 1], sols[t])
    t0 = time.time()
    l0, _ = ref.cnn_train(1, batch=batch, lr=lr, seed=seed, res_weight=res_weight)
    _, l1 = ref.cnn_train(train_steps - 1, batch=batch, lr=lr, seed=seed + 1,
                          res_weight=res_weight)
    train_s = time.time() - t0
    print(f"training: {train_steps} steps of {batch} blocks, loss {l0:.3e} -> {l1:.3e}, "
          f"{train_s:.1f} s", file=log, flush=True)
    weights = ref.cnn_get_weights()

    def evaluate(rhs_, sols_, ts):
        out = {}
        for form in forms:
            kf = dict(kw, max_iter=fp64_max_iter) if form == "fp64" else kw
            base = solver_for(w, precision=form, guess="given", **kf)
            net = solver_for(w, precision=form, guess="cnn", **kf)
            net.cnn_set_weights(weights)
            # with the extrapolation in time from the last `history` previous
            # pressures: alone, and together with the net (least-squares
            # weights)
            ext = solver_for(w, precision=form, guess="given", extrapolate=history, **kf)
            both = solver_for(w, precision=form, guess="cnn", extrapolate=history, **kf)
            both.cnn_set_weights(weights)
            it_e, it_x, w_net, fail_e, fail_x = [], [], [], 0, 0
            it_b, it_n, used, ratio, err, raw, fail_b, fail_n = [], [], 0, [], [], [], 0, 0
            xp = torch.empty(w.ncells, dtype=torch.float64, device=dev)
            ax = torch.empty_like(xp)
            for t in ts:
                # the net's own residual, whether or not the solve takes it
                net.cnn_predict(rhs_[t], sols_[t - 1], xp)
                net.apply_operator(xp, ax)
                rn_ = torch.linalg.norm(rhs_[t] - ax)
                net.apply_operator(sols_[t - 1], ax)
                raw.append(float(rn_ / torch.linalg.norm(rhs_[t] - ax)))
                xb = sols_[t - 1].clone()
                rb = base.solve_device(rhs_[t], xb)
                xn = sols_[t - 1].clone()
                rn = net.solve_device(rhs_[t], xn)
                xe = sols_[t - 1].clone()
                re_ = ext.solve_device(rhs_[t], xe)
                xx = sols_[t - 1].clone()
                rx = both.solve_device(rhs_[t], xx)
                if t >= ts[0] + history:  # the history is full from here on
                    it_e.append(re_["iterations"])
                    it_x.append(rx["iterations"])
                    w_net.append(rx["guess_weight_net"])
                    fail_e += int(re_["status"] != 0)
                    fail_x += int(rx["status"] != 0)
                # a solve that stops at max_iter counts with its iterations
                fail_b += int(rb["status"] != 0)
                fail_n += int(rn["status"] != 0)
                if t % 25 == 0:
                    print(f"  {form} step {t}: {rb['iterations']} vs {rn['iterations']} iterations",
                          file=log, flush=True)
                it_b.append(rb["iterations"])
                it_n.append(rn["iterations"])
                used += int(rn["guess_used"] == 2)
                ratio.append(rn["guess_resid"] / rb["err_init"])
                # pressure is defined up to a constant here (walls, bMean = 0):
                # compare after removing the mean, as the caller does
                dx = (xn - (hw * xn).sum() / hw.sum()) - sols_[t]
                err.append(float(torch.linalg.norm(dx) / torch.linalg.norm(sols_[t])))
            out[form] = {
                "iterations_prev_guess": float(np.mean(it_b)),
                "iterations_net_guess": float(np.mean(it_n)),
                "iterations_prev_guess_median": float(np.median(it_b)),
                "iterations_net_guess_median": float(np.median(it_n)),
                "iterations_saved_pct": round(100.0 * (1 - np.sum(it_n) / np.sum(it_b)), 2),
                "net_guess_used": f"{used}/{len(ts)}",
                "not_converged_prev_guess": fail_b,
                "not_converged_net_guess": fail_n,
                "initial_resid_ratio_median": float(np.median(ratio)),
                "net_resid_ratio_median": float(np.median(raw)),
                "err_vs_fp64_max": float(np.max(err)),
                # the steps with a full history: extrapolation alone and with
                # the net, against the previous pressure on the same steps
                "extrap_history": history,
                "iterations_prev_guess_same_steps": float(np.mean(it_b[history:])),
                "iterations_extrap": float(np.mean(it_e)),
                "iterations_net_extrap": float(np.mean(it_x)),
                "iterations_saved_extrap_pct": round(
                    100.0 * (1 - np.sum(it_e) / np.sum(it_b[history:])), 2),
                "iterations_saved_net_extrap_pct": round(
                    100.0 * (1 - np.sum(it_x) / np.sum(it_b[history:])), 2),
                "net_weight_with_extrap_median": float(np.median(w_net)),
                "not_converged_extrap": fail_e,
                "not_converged_net_extrap": fail_x,
            }
        return out

    held = evaluate(rhs, sols, range(train + 1, steps + 1))
    frhs, fsols, _ = sequence(fresh, W.FRESH_SEED)
    fr = evaluate(frhs, fsols, range(1, fresh + 1))
    return {
        "workload": w.name, "cells": w.ncells, "steps": steps, "train_triples": train,
        "train_steps": train_steps, "batch_blocks": batch, "lr": lr, "res_weight": res_weight,
        "loss_first": l0, "loss_last": l1, "train_seconds": round(train_s, 1),
        "sequence_solver": "mixed_fp16 refined to the FP64 true residual",
        "sequence_iterations_mean_train": float(np.mean(its[1:train + 1])),
        "held_out": held, "fresh_seed": fr,
    }


def main():
    p = argparse.ArgumentParser()
    p.add_argument("--steps", type=int, default=400)
    p.add_argument("--train", type=int, default=300)
    p.add_argument("--fresh", type=int, default=40)
    p.add_argument("--train-steps", type=int, default=20000)
    p.add_argument("--batch", type=int, default=256)
    p.add_argument("--lr", type=float, default=1e-3)
    p.add_argument("--res-weight", type=float, default=100.0,
                   help="weight of the local-residual term of the training loss")
    p.add_argument("--bpd", type=int, nargs=3, default=[32, 16, 16],
                   help="blocks per dimension of the uniform sequence")
    p.add_argument("--workload", default="c4", choices=["c4", "c5"],
                   help="c4: uniform blob sequence; c5: the 4-level AMR school grid")
    p.add_argument("--fp64-max-iter", type=int, default=3000)
    p.add_argument("--history", type=int, default=1,
                   help="previous pressures of the extrapolation legs (1 = linear)")
    a = p.parse_args()
    if not 0 < a.train < a.steps:
        raise SystemExit("need 0 < --train < --steps")
    print(json.dumps(run(a.bpd, a.steps, a.train, a.fresh, a.train_steps, a.batch, a.lr,
                         a.res_weight, workload=a.workload, fp64_max_iter=a.fp64_max_iter,
                         history=a.history)))


if __name__ == "__main__":
    main()
