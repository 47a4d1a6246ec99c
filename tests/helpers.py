"""Small grids and fields shared by the tests."""
from __future__ import annotations

import numpy as np

from cup3d_amd import workloads as W

PER, WALL, FREE = W.BC_PERIODIC, W.BC_WALL, W.BC_FREESPACE


def uniform(bpd, bc, h0=None):
    h0 = h0 if h0 is not None else 1.0 / (8 * max(bpd))
    return W.uniform_blocks(*bpd), tuple(bpd), h0, tuple(bc)


def amr_sphere(bpd=(4, 4, 4), lmax=1, thr=(0.15,), centre=(0.5, 0.5, 0.5), r=0.2, bc=(2, 2, 2),
               h0=None):
    h0 = h0 if h0 is not None else 1.0 / (8 * max(bpd))
    blocks = W.amr_sphere_blocks(bpd, h0, lmax, thr, centre, r)
    return blocks, tuple(bpd), h0, tuple(bc)


# A family of grids that exercise same-level, coarse and fine faces in all six
# directions, periodic wrap across coarse-fine faces, and 3 levels.
def grid_cases():
    return {
        "uniform_periodic_2x3x2": uniform((2, 3, 2), (PER, PER, PER)),
        "uniform_walls_3x2x2": uniform((3, 2, 2), (WALL, WALL, WALL)),
        "uniform_mixed_bc": uniform((2, 2, 3), (PER, WALL, FREE)),
        "amr2_walls": amr_sphere((4, 4, 4), 1, (0.15,), (0.5, 0.5, 0.5), 0.2, (WALL,) * 3),
        "amr2_periodic_offcentre": amr_sphere((4, 4, 4), 1, (0.12,), (0.1, 0.9, 0.15), 0.15,
                                              (PER,) * 3),
        "amr3_walls": amr_sphere((2, 2, 2), 2, (0.3, 0.12), (0.45, 0.55, 0.5), 0.15, (WALL,) * 3,
                                 h0=1.0 / 16),
    }


def random_field(nb, seed=0):
    return np.random.default_rng(seed).standard_normal((nb, 8, 8, 8))


def compatible_rhs(blocks, h0, seed):
    """A random rhs with zero volume-weighted sum (solvable with walls)."""
    h3 = np.repeat((h0 / 2.0 ** np.asarray(blocks)[:, 0]) ** 3, 512)
    b = random_field(len(blocks), seed).reshape(-1)
    return b - h3 * b.sum() / h3.sum()


def cell_centres(blocks, h0):
    """Cell-centre coordinates [nb, 8, 8, 8, 3] and spacing h[nb]."""
    blocks = np.asarray(blocks)
    h = h0 / (2.0 ** blocks[:, 0])
    i = np.arange(8) + 0.5
    iz, iy, ix = np.meshgrid(i, i, i, indexing="ij")
    X = (blocks[:, 1, None, None, None] * 8 + ix[None]) * h[:, None, None, None]
    Y = (blocks[:, 2, None, None, None] * 8 + iy[None]) * h[:, None, None, None]
    Z = (blocks[:, 3, None, None, None] * 8 + iz[None]) * h[:, None, None, None]
    return np.stack([X, Y, Z], -1), h


def boundary_cells(blocks, bpd, bc):
    """Mask of cells that sit on a non-periodic domain face."""
    blocks = np.asarray(blocks)
    m = np.zeros((len(blocks), 8, 8, 8), bool)
    for b, (l, i, j, k) in enumerate(blocks):
        I = (i, j, k)
        for d in range(3):
            if bc[d] == PER:
                continue
            n = bpd[d] << l
            sl_lo = [slice(None)] * 3
            sl_hi = [slice(None)] * 3
            ax = 2 - d  # arrays are [iz][iy][ix]
            sl_lo[ax] = 0
            sl_hi[ax] = 7
            if I[d] == 0:
                m[b][tuple(sl_lo)] = True
            if I[d] == n - 1:
                m[b][tuple(sl_hi)] = True
    return m


def tiny_amr(periodic=False, three_levels=False):
    """2x2x2 base blocks with block (0,0,0) refined (15 blocks); with
    three_levels (walls only) its child at the origin is refined again."""
    h0 = 1.0 / 16
    fine = [[1, i, j, k] for k in range(2) for j in range(2) for i in range(2)]
    if three_levels:
        fine = fine[1:] + [[2, i, j, k] for k in range(2) for j in range(2) for i in range(2)]
    coarse = [[0, 1, 0, 0], [0, 0, 1, 0], [0, 1, 1, 0], [0, 0, 0, 1], [0, 1, 0, 1],
              [0, 0, 1, 1], [0, 1, 1, 1]]
    blocks = np.array(fine + coarse, np.int32)
    bc = (PER,) * 3 if periodic else (WALL,) * 3
    return blocks, (2, 2, 2), h0, bc


def school4_small():
    """Four levels around 6 seeded ellipsoids on 4x2x2 base blocks (625
    blocks): every coarse-fine configuration of the school grid at a size the
    oracle handles in seconds."""
    from cup3d_amd import workloads as W
    bpd, h0 = (4, 2, 2), 1.0 / 16
    blocks = W.amr_school_blocks(bpd, h0, 3, (0.18, 0.09, 0.045), 6, W.SCHOOL_AXES, W.SEED)
    return blocks, bpd, h0, (WALL,) * 3


def global_to_blocks(g, blocks):
    return np.stack([g[8 * k:8 * k + 8, 8 * j:8 * j + 8, 8 * i:8 * i + 8]
                     for (_, i, j, k) in blocks])


def blocks_to_global(f, blocks, bpd):
    nx, ny, nz = (8 * b for b in bpd)
    g = np.zeros((nz, ny, nx))
    for b, (_, i, j, k) in enumerate(blocks):
        g[8 * k:8 * k + 8, 8 * j:8 * j + 8, 8 * i:8 * i + 8] = f[b]
    return g


def fft_solution_periodic(rhs, blocks, bpd, h0, row):
    """Exact solution of the bMean = 1 system on a uniform periodic grid:
    A x = b on every row but the origin row, sum x = 0. Uses the Fourier
    eigenvalues h (2 cos(2 pi m_x/n_x) - 2 + ...)."""
    b = np.asarray(rhs, np.float64).reshape(len(blocks), 8, 8, 8).copy()
    flat = b.reshape(-1)
    flat[row] = 0.0
    flat[row] = -flat.sum()
    g = blocks_to_global(b, blocks, bpd)
    n = g.shape
    lam = sum(np.reshape(2 * np.cos(2 * np.pi * np.arange(n[a]) / n[a]) - 2,
                         [-1 if d == a else 1 for d in range(3)]) for a in range(3)) * h0
    gh = np.fft.fftn(g)
    lam[0, 0, 0] = 1.0
    xh = gh / lam
    xh[0, 0, 0] = 0.0
    x = np.real(np.fft.ifftn(xh))
    return global_to_blocks(x, blocks).reshape(-1)


def index_encoded_case(blocks, bpd, h0, bc):
    """Known answer for the block-neighbour tables on a uniform level-0 grid
    (the idea of the reference's tests/cubismup3d/TestBoundaries.cpp:19-82:
    an index-encoded field whose shifted values are known exactly).

    x = i + Nx j + Nx Ny k on global cell (i, j, k). The 7-point Laplacian of
    a linear field is 0 away from the domain faces, so h (sum of the six
    neighbours - 6 x) is nonzero only on the faces: a periodic wrap brings
    the value from the far side, +/- (stride of the direction) x (cells in
    it); a wall or freespace ghost equals the centre, +/- stride. All values
    are integers times a power of two: the answer is exact in FP64 and FP32.
    Returns (x, A x) in block layout, flat."""
    blocks = np.asarray(blocks)
    assert (blocks[:, 0] == 0).all()
    n = [8 * b for b in bpd]
    stride = [1, n[0], n[0] * n[1]]
    r = np.arange(8)
    iz, iy, ix = np.meshgrid(r, r, r, indexing="ij")
    g = [blocks[:, 1, None, None, None] * 8 + ix[None],
         blocks[:, 2, None, None, None] * 8 + iy[None],
         blocks[:, 3, None, None, None] * 8 + iz[None]]
    x = sum(stride[d] * g[d] for d in range(3)).astype(np.float64)
    y = np.zeros_like(x)
    for d in range(3):
        jump = stride[d] * (n[d] if bc[d] == PER else 1)
        y += jump * (g[d] == 0) - jump * (g[d] == n[d] - 1)
    return x.reshape(-1), (h0 * y).reshape(-1)


def nbr_table(blocks, bpd, bc):
    """Same-level face neighbours (x-, x+, y-, y+, z-, z+), -1 otherwise."""
    idx = {tuple(b): i for i, b in enumerate(np.asarray(blocks).tolist())}
    out = -np.ones((len(blocks), 6), np.int64)
    for i, (l, a, b, c) in enumerate(np.asarray(blocks).tolist()):
        I = [a, b, c]
        for f in range(6):
            d, s = f // 2, (1 if f % 2 else -1)
            n = bpd[d] << l
            J = list(I)
            J[d] += s
            if J[d] < 0 or J[d] >= n:
                if bc[d] != 1:
                    continue
                J[d] %= n
            out[i, f] = idx.get((l, *J), -1)
    return out


def face_tables(blocks, bpd, bc):
    """Per block face (x-, x+, y-, y+, z-, z+): kind (0 none, 1 same level,
    2 coarser, 3 finer), the same-level neighbour, and the partners (the
    coarser block, or the four finer blocks by quadrant q = t1 half + 2 t2
    half, tangents (y, z), (z, x), (x, y)); the index parities per block.
    A dictionary walk of the block list, for checking the library's builds."""
    blocks = np.asarray(blocks).tolist()
    idx = {tuple(b): i for i, b in enumerate(blocks)}
    nb = len(blocks)
    kind = np.zeros((nb, 6), np.int8)
    nbr = -np.ones((nb, 6), np.int32)
    fnb = -np.ones((nb, 6, 4), np.int32)
    bpar = np.zeros(nb, np.int8)
    tang = {0: (1, 2), 1: (2, 0), 2: (0, 1)}
    for i, (l, a, b, c) in enumerate(blocks):
        bpar[i] = (a & 1) | ((b & 1) << 1) | ((c & 1) << 2)
        for f in range(6):
            d, s = f // 2, (1 if f % 2 else -1)
            n = bpd[d] << l
            J = [a, b, c]
            J[d] += s
            if J[d] < 0 or J[d] >= n:
                if bc[d] != PER:
                    continue
                J[d] %= n
            if (l, *J) in idx:
                kind[i, f], nbr[i, f] = 1, idx[(l, *J)]
                continue
            par = (l - 1, J[0] // 2, J[1] // 2, J[2] // 2)
            if l > 0 and par in idx:
                kind[i, f], fnb[i, f, 0] = 2, idx[par]
                continue
            t1, t2 = tang[d]
            for q in range(4):
                C = [2 * J[0], 2 * J[1], 2 * J[2]]
                C[d] += 1 if s < 0 else 0
                C[t1] += q & 1
                C[t2] += q >> 1
                fnb[i, f, q] = idx[(l + 1, *C)]
            kind[i, f] = 3
    return kind, nbr, fnb.reshape(nb, 24), bpar
