"""Writes c5_fp64_fixture.json: the oracle's classical FP64 BiCGSTAB
(BiCGSTAB.cu:416-628, restated in oracle/oracle.cpp) on the full 4-level AMR
school grid (59 012 blocks, 30.2 M cells), bMeanConstraint 0, rhs of step 0
of the moving-blob sequence, zero guess: the recurrence residual ||r_k|| of
the first ITERS iterations, and the best iterate after 10 iterations at 64
fixed cells (with its norm). The GPU test compares the FP64 drop-in with
these numbers at the same iteration counts, on a grid too large for the
oracle to run inside the test (about a minute of CPU per 10 iterations).

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
    out = {
        "workload": w.name, "cells": w.ncells, "bmean": 0, "rhs_step": 0, "rhs_seed": W.SEED,
        "guess": "zero", "err_init": info["err_init"],
        "history": [float(v) for v in info["history"]],
        "x10_norm": float(np.linalg.norm(x10)), "x10_err_final": info10["err_final"],
        "x10_index": [int(i) for i in idx], "x10_values": [float(x10[i]) for i in idx],
    }
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "c5_fp64_fixture.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
    print("wrote", path)


if __name__ == "__main__":
    main()
