"""Pins the oracle's block-Jacobi inverse (PoissonSolverExp.cpp:17-107)."""
import json
import os

import numpy as np
import pyoracle

GOLD = os.path.join(os.path.dirname(__file__), "golden", "pinv_fixture.json")


def K_local():
    """6 I - face adjacency of the 8^3 block with zero Dirichlet ghosts."""
    n = 512
    K = 6.0 * np.eye(n)
    for i in range(n):
        x, y, z = i % 8, (i // 8) % 8, i // 64
        for dx, dy, dz in ((1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)):
            a, b, c = x + dx, y + dy, z + dz
            if 0 <= a < 8 and 0 <= b < 8 and 0 <= c < 8:
                K[i, a + 8 * b + 64 * c] = -1.0
    return K


def dst_inverse():
    """-K^-1 from the 3-D discrete sine transform eigendecomposition."""
    a = np.arange(1, 9)
    S = np.sqrt(2 / 9) * np.sin(np.pi * np.outer(a, a) / 9)  # S[a-1, i]
    lam1 = 2 - 2 * np.cos(np.pi * a / 9)
    lam = lam1[None, None, :] + lam1[None, :, None] + lam1[:, None, None]  # [c][b][a]
    V = np.einsum("ck,bj,ai->cbakji", S, S, S).reshape(512, 512)  # rows (c,b,a), cols (k,j,i)
    return -(V.T * (1.0 / lam.reshape(-1))) @ V


def test_pinv_inverts_local_laplacian():
    P = pyoracle.pinv()
    K = K_local()
    assert np.abs(P @ K + np.eye(512)).max() < 1e-12


def test_pinv_matches_dst_closed_form():
    P = pyoracle.pinv()
    assert np.abs(P - dst_inverse()).max() < 1e-14


def test_pinv_symmetric_and_negative_definite():
    P = pyoracle.pinv()
    assert np.array_equal(P, P.T)
    assert np.linalg.eigvalsh(P).max() < 0


def test_pinv_golden_fixture():
    g = json.load(open(GOLD))
    P = pyoracle.pinv()
    assert abs(P[0, 0] - g["p00"]) < 1e-14
    assert abs(P[255, 256] - g["p255_256"]) < 1e-14
    assert np.allclose(P[0, :8], g["row0_first8"], rtol=0, atol=1e-15)
    assert abs(np.trace(P) - g["trace"]) < 1e-12
    assert abs(np.linalg.norm(P) - g["fro"]) < 1e-12
    assert abs(P.min() - g["min"]) < 1e-15 and abs(P.max() - g["max"]) < 1e-15
