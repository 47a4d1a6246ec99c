"""Writes c5_fp64_fixture.json: the oracle's classical FP64 BiCGSTAB
(BiCGSTAB.cu:416-628, restated in oracle/oracle.cpp) on the full 4-level AMR
school grid (59 012 blocks, 30.2 M cells), bMeanConstraint 0, rhs of step 0
of the moving-blob sequence, zero guess: the recurrence residual ||r_k|| of
the first ITERS iterations, and the best iterate after 10 iterations at 64
fixed cells (with its norm). The same recurrence on the rhs times
(1 + eps xi), xi standard normal per cell (seed PERTURB_SEED), for eps = 1e-15
and 1e-11: how far the oracle moves from itself under perturbations at the
rounding level, which the recurrence amplifies as it runs. The GPU test
compares the FP64 drop-in with these numbers at the same iteration counts, on
a grid too large for the oracle to run inside the test (about a minute of CPU
per 10 iterations; the whole script takes about ten minutes).

  python tests/golden/make_c5_fp64_fixture.py     (from the repository root)
"""
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import pyoracle  # noqa: E402
from cup3d_amd import workloads as W  # noqa: E402

ITERS = 30
SAMPLE_SEED = 7
PERTURB_SEED = 11
PERTURB_EPS = (1e-15, 1e-11)


def main():
    w = W.c5_amr_school()
    b = w.rhs(step=0).reshape(-1)
    t0 = time.time()
    O = pyoracle.Oracle(w.blocks, w.bpd, w.h0, w.bc, 0)
    print(f"assembled {w.ncells} cells in {time.time() - t0:.1f} s", flush=True)
    t0 = time.time()
    _, info = O.bicgstab(b, tol_rel=0.0, max_iter=ITERS, history=True)
    print(f"{ITERS} iterations in {time.time() - t0:.1f} s", flush=True)
    x10, info10 = O.bicgstab(b, tol_rel=0.0, max_iter=10)
    idx = np.sort(np.random.default_rng(SAMPLE_SEED).choice(w.ncells, 64, replace=False))
    xi = np.random.default_rng(PERTURB_SEED).standard_normal(b.size)
    perturbed = {}
    for eps in PERTURB_EPS:
        _, pinfo = O.bicgstab(b * (1 + eps * xi), tol_rel=0.0, max_iter=ITERS, history=True)
        perturbed[f"{eps:g}"] = [float(v) for v in pinfo["history"]]
        print(f"perturbed by {eps:g}: done", flush=True)
    out = {
        "workload": w.name, "cells": w.ncells, "bmean": 0, "rhs_step": 0, "rhs_seed": W.SEED,
        "guess": "zero", "err_init": info["err_init"],
        "history": [float(v) for v in info["history"]],
        "x10_norm": float(np.linalg.norm(x10)), "x10_err_final": info10["err_final"],
        "x10_index": [int(i) for i in idx], "x10_values": [float(x10[i]) for i in idx],
        "perturb_seed": PERTURB_SEED, "history_perturbed": perturbed,
    }
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "c5_fp64_fixture.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
    print("wrote", path)


if __name__ == "__main__":
    main()
