"""Writes pinv_fixture.json: summary values of -K^-1 for the 8^3 Dirichlet
block Laplacian K (6 on the diagonal, -1 per face neighbour), computed with
numpy.linalg.inv — independent of the oracle's Cholesky/Gauss-Jordan code.
Run: python tests/golden/make_pinv_fixture.py
"""
import json
import os

import numpy as np


def K_local():
    K = 6.0 * np.eye(512)
    for i in range(512):
        x, y, z = i % 8, (i // 8) % 8, i // 64
        for dx, dy, dz in ((1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)):
            a, b, c = x + dx, y + dy, z + dz
            if 0 <= a < 8 and 0 <= b < 8 and 0 <= c < 8:
                K[i, a + 8 * b + 64 * c] = -1.0
    return K


P = -np.linalg.inv(K_local())
out = {
    "p00": P[0, 0], "p255_256": P[255, 256], "row0_first8": P[0, :8].tolist(),
    "trace": float(np.trace(P)), "fro": float(np.linalg.norm(P)),
    "min": float(P.min()), "max": float(P.max()),
}
path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "pinv_fixture.json")
json.dump(out, open(path, "w"), indent=1)
print(path)
