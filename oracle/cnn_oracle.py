"""FP64 CPU oracle of the learned warm start — TEST INFRASTRUCTURE ONLY.

The network has no counterpart in the reference solver: a 3-layer 3-D CNN
applied per 8^3 block with a one-cell halo, in the increment form of the
pressure solve (PressureProjection.cpp:368-385 solves A (p - p_old) =
b - A p_old):
  input  2 x 10^3  (res / (h_b s_r), p_prev / s_p),  res = b - A p_prev
  conv 3^3 valid, 2 -> 16, GELU
  conv 3^3 same (zero padded), 16 -> 16, GELU
  conv 3^3 same (zero padded), 16 -> 1
  p_net = p_prev + s_r * out,
  s_r = RMS(res / h_b), s_p = RMS(p_prev) over the grid (1 where zero).
res / h_b carries the units of the 7-point stencil applied to the increment
(A = h^3 lap), so the input is commensurate on every AMR level.

Halo rule: a face halo cell takes the adjacent cell of the same-level
neighbour block when the face is shared with one; every other halo cell
(domain wall, coarse/fine face, edges and corners) repeats the nearest cell
of the block itself.

Weights: flat float vector, W1[16][2][3][3][3], b1[16], W2[16][16][3][3][3],
b2[16], W3[1][16][3][3][3], b3[1] (8241 values); convolution is correlation
(out[o, z, y, x] = sum W[o, c, dz, dy, dx] in[c, z+dz, y+dy, x+dx]) as in
torch.nn.functional.conv3d. GELU is the exact erf form.

Pinned in tests/test_cnn_oracle.py by torch.nn.functional.conv3d/gelu on the
CPU and by central finite differences of the loss.
"""
from __future__ import annotations

import math

import numpy as np
from scipy.special import erf

SHAPES = [(16, 2, 3, 3, 3), (16,), (16, 16, 3, 3, 3), (16,), (1, 16, 3, 3, 3), (1,)]
NW = sum(int(np.prod(s)) for s in SHAPES)  # 8241


def unpack(w):
    out, o = [], 0
    for s in SHAPES:
        n = int(np.prod(s))
        out.append(np.asarray(w[o:o + n], np.float64).reshape(s))
        o += n
    return out


def pack(ws):
    return np.concatenate([np.asarray(a, np.float64).reshape(-1) for a in ws])


def init_weights(seed=0, scale=1.0):
    rng = np.random.default_rng(seed)
    ws = []
    for s in SHAPES:
        if len(s) == 5:
            fan_in = s[1] * 27
            ws.append(rng.standard_normal(s) * scale * math.sqrt(2.0 / fan_in))
        else:
            ws.append(np.zeros(s))
    return pack(ws)


def gelu(x):
    return 0.5 * x * (1.0 + erf(x / math.sqrt(2.0)))


def gelu_grad(x):
    return 0.5 * (1.0 + erf(x / math.sqrt(2.0))) + x * np.exp(-0.5 * x * x) / math.sqrt(2 * math.pi)


def halo(field, nbr):
    """field [nb,8,8,8] ([z][y][x]) -> [nb,10,10,10] with the halo rule above.
    nbr [nb,6]: same-level neighbour per face (x-, x+, y-, y+, z-, z+) or -1."""
    f = np.asarray(field, np.float64)
    nb = f.shape[0]
    out = np.zeros((nb, 10, 10, 10))
    out[:, 1:9, 1:9, 1:9] = f
    # clamp everything first (edges, corners, walls, coarse/fine faces)
    idx = np.clip(np.arange(10) - 1, 0, 7)
    out[:] = f[:, idx][:, :, idx][:, :, :, idx]
    for b in range(nb):
        n = nbr[b]
        if n[0] >= 0: out[b, 1:9, 1:9, 0] = f[n[0], :, :, 7]
        if n[1] >= 0: out[b, 1:9, 1:9, 9] = f[n[1], :, :, 0]
        if n[2] >= 0: out[b, 1:9, 0, 1:9] = f[n[2], :, 7, :]
        if n[3] >= 0: out[b, 1:9, 9, 1:9] = f[n[3], :, 0, :]
        if n[4] >= 0: out[b, 0, 1:9, 1:9] = f[n[4], 7, :, :]
        if n[5] >= 0: out[b, 9, 1:9, 1:9] = f[n[5], 0, :, :]
    return out


def rms(f):
    s = math.sqrt(float(np.mean(np.asarray(f, np.float64) ** 2)))
    return s if s > 0 else 1.0


def inputs(res, prev, h, nbr):
    """Network input [nb,2,10,10,10] and the scales (s_r, s_p)."""
    prev = np.asarray(prev, np.float64).reshape(-1, 8, 8, 8)
    rh = np.asarray(res, np.float64).reshape(-1, 8, 8, 8) / np.asarray(h)[:, None, None, None]
    s_r, s_p = rms(rh), rms(prev)
    x = np.stack([halo(rh / s_r, nbr), halo(prev / s_p, nbr)], 1)
    return x, s_r, s_p


def conv_valid(x, W):
    """x [n,C,Z+2,Y+2,X+2], W [O,C,3,3,3] -> [n,O,Z,Y,X]."""
    n, C, Zp, Yp, Xp = x.shape
    Z, Y, X = Zp - 2, Yp - 2, Xp - 2
    out = np.zeros((n, W.shape[0], Z, Y, X))
    for dz in range(3):
        for dy in range(3):
            for dx in range(3):
                out += np.einsum("nczyx,oc->nozyx", x[:, :, dz:dz + Z, dy:dy + Y, dx:dx + X],
                                 W[:, :, dz, dy, dx])
    return out


def pad1(x):
    return np.pad(x, ((0, 0), (0, 0), (1, 1), (1, 1), (1, 1)))


def forward(w, x):
    W1, b1, W2, b2, W3, b3 = unpack(w)
    z1 = conv_valid(x, W1) + b1[None, :, None, None, None]
    h1 = gelu(z1)
    z2 = conv_valid(pad1(h1), W2) + b2[None, :, None, None, None]
    h2 = gelu(z2)
    out = conv_valid(pad1(h2), W3) + b3[None, :, None, None, None]
    return out[:, 0], (x, z1, h1, z2, h2)


def local_residual(out, q, target_halo):
    """rho = (7-point Laplacian of out) - q in every cell of every block, the
    stencil closed outside the block by the target's halo [nb,10,10,10]. With
    out = delta / s_r and q = res / (h s_r), rho = 0 is A delta = res."""
    o = np.array(target_halo, np.float64)
    o[:, 1:9, 1:9, 1:9] = out
    lap = (o[:, :8, 1:9, 1:9] + o[:, 2:, 1:9, 1:9] + o[:, 1:9, :8, 1:9] + o[:, 1:9, 2:, 1:9] +
           o[:, 1:9, 1:9, :8] + o[:, 1:9, 1:9, 2:] - 6.0 * out)
    return lap - q


def lap_transpose_in_block(rho):
    """Adjoint of the in-block part of the Laplacian: sum of rho over the
    in-block neighbours of each cell minus 6 rho."""
    g = -6.0 * rho
    g[:, 1:] += rho[:, :-1]
    g[:, :-1] += rho[:, 1:]
    g[:, :, 1:] += rho[:, :, :-1]
    g[:, :, :-1] += rho[:, :, 1:]
    g[:, :, :, 1:] += rho[:, :, :, :-1]
    g[:, :, :, :-1] += rho[:, :, :, 1:]
    return g


def loss_and_grad(w, x, target, lam=0.0, target_halo=None):
    """Loss 0.5 mean((out - target)^2) + lam 0.5 mean(rho^2), rho the local
    residual of the prediction (local_residual; needs target_halo when
    lam != 0), and its gradient."""
    W1, b1, W2, b2, W3, b3 = unpack(w)
    out, (x, z1, h1, z2, h2) = forward(w, x)
    n = out.size
    d = (out - target) / n
    loss = 0.5 * float(np.sum((out - target) ** 2)) / n
    if lam != 0.0:
        rho = local_residual(out, x[:, 0, 1:9, 1:9, 1:9], target_halo)
        loss += lam * 0.5 * float(np.sum(rho ** 2)) / n
        d = d + lam * lap_transpose_in_block(rho) / n
    dout = d[:, None]

    def wgrad(inp_padded, dz_):
        Z, Y, X = dz_.shape[2:]
        g = np.zeros((dz_.shape[1], inp_padded.shape[1], 3, 3, 3))
        for a in range(3):
            for b_ in range(3):
                for c in range(3):
                    g[:, :, a, b_, c] = np.einsum("nozyx,nczyx->oc", dz_,
                                                  inp_padded[:, :, a:a + Z, b_:b_ + Y, c:c + X])
        return g

    def back_input(dz_, W):
        # gradient w.r.t. the padded input of conv_valid
        n_, O, Z, Y, X = dz_.shape
        gi = np.zeros((n_, W.shape[1], Z + 2, Y + 2, X + 2))
        for a in range(3):
            for b_ in range(3):
                for c in range(3):
                    gi[:, :, a:a + Z, b_:b_ + Y, c:c + X] += np.einsum("nozyx,oc->nczyx", dz_,
                                                                       W[:, :, a, b_, c])
        return gi

    gW3 = wgrad(pad1(h2), dout)
    gb3 = dout.sum(axis=(0, 2, 3, 4))
    dh2 = back_input(dout, W3)[:, :, 1:-1, 1:-1, 1:-1]
    dz2 = dh2 * gelu_grad(z2)
    gW2 = wgrad(pad1(h1), dz2)
    gb2 = dz2.sum(axis=(0, 2, 3, 4))
    dh1 = back_input(dz2, W2)[:, :, 1:-1, 1:-1, 1:-1]
    dz1 = dh1 * gelu_grad(z1)
    gW1 = wgrad(x, dz1)
    gb1 = dz1.sum(axis=(0, 2, 3, 4))
    return loss, pack([gW1, gb1, gW2, gb2, gW3, gb3])


def adam(w, g, m, v, t, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam step (Kingma & Ba 2015), t counted from 1."""
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** t)
    vh = v / (1 - b2 ** t)
    return w - lr * mh / (np.sqrt(vh) + eps), m, v


def predict(w, res, prev, h, nbr):
    """p_net for res = b - A p_prev."""
    x, s_r, _ = inputs(res, prev, h, nbr)
    out, _ = forward(w, x)
    return np.asarray(prev, np.float64).reshape(-1, 8, 8, 8) + s_r * out
