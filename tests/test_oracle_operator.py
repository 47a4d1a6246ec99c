"""Pins the oracle's operator (PoissonSolverExp::getMat, .cpp:221-373).

Uniform grids: exact eigenpairs (periodic Fourier modes, Neumann cosine
modes) and a scipy Kronecker-product brute force. AMR grids: zero row sums,
conservation 1^T A = 0, exactness on quadratics (the coarse-fine ghost is a
quadratic Lagrange/Taylor interpolant), and hand-derived rows.
"""
import numpy as np
import pyoracle
import pytest
import scipy.sparse as sp

from helpers import (PER, WALL, amr_sphere, boundary_cells, cell_centres, grid_cases,
                     index_encoded_case, random_field, uniform)


def make(case, bmean=0):
    blocks, bpd, h0, bc = case
    return pyoracle.Oracle(blocks, bpd, h0, bc, bmean)


def to_global(field, blocks, bpd):
    """Block layout -> global [z][y][x] array (uniform level-0 grids)."""
    nx, ny, nz = (8 * b for b in bpd)
    g = np.zeros((nz, ny, nx))
    for b, (_, i, j, k) in enumerate(blocks):
        g[8 * k:8 * k + 8, 8 * j:8 * j + 8, 8 * i:8 * i + 8] = field[b]
    return g


@pytest.mark.parametrize("bpd", [(2, 3, 2), (4, 4, 4)])
def test_periodic_fourier_eigenpairs(bpd):
    case = uniform(bpd, (PER, PER, PER))
    O = make(case)
    blocks, _, h0, _ = case
    X, h = cell_centres(blocks, h0)
    L = [8 * b * h0 for b in bpd]
    for m in [(1, 0, 0), (2, 1, 3), (0, 3, 1)]:
        k = [2 * np.pi * m[d] / L[d] for d in range(3)]
        phase = sum(k[d] * X[..., d] for d in range(3))
        lam = h0 * sum(2 * np.cos(k[d] * h0) - 2 for d in range(3))
        for f in (np.cos(phase), np.sin(phase)):
            y = O.spmv_raw(f).reshape(f.shape)
            assert np.abs(y - lam * f).max() < 1e-13 * max(1.0, abs(lam))


def test_neumann_cosine_eigenpairs():
    bpd = (3, 2, 2)
    case = uniform(bpd, (WALL, WALL, WALL))
    O = make(case)
    blocks, _, h0, _ = case
    n = [8 * b for b in bpd]
    gi = [np.arange(nn) for nn in n]
    for m in [(1, 0, 0), (2, 3, 1), (5, 1, 7)]:
        g = (np.cos(np.pi * m[2] * (gi[2][:, None, None] + 0.5) / n[2]) *
             np.cos(np.pi * m[1] * (gi[1][None, :, None] + 0.5) / n[1]) *
             np.cos(np.pi * m[0] * (gi[0][None, None, :] + 0.5) / n[0]))
        f = np.stack([g[8 * k:8 * k + 8, 8 * j:8 * j + 8, 8 * i:8 * i + 8]
                      for (_, i, j, k) in blocks])
        lam = h0 * sum(2 * np.cos(np.pi * m[d] / n[d]) - 2 for d in range(3))
        y = O.spmv_raw(f).reshape(f.shape)
        assert np.abs(y - lam * f).max() < 1e-13


@pytest.mark.parametrize("bc", [(PER, PER, PER), (WALL, WALL, WALL), (PER, WALL, 0)])
def test_uniform_matches_kronecker_bruteforce(bc):
    bpd = (2, 3, 2)
    case = uniform(bpd, bc)
    blocks, _, h0, _ = case
    O = make(case)

    def lap1d(n, periodic):
        T = sp.diags([np.ones(n - 1), -2 * np.ones(n), np.ones(n - 1)], [-1, 0, 1]).tolil()
        if periodic:
            T[0, n - 1] = T[n - 1, 0] = 1
        else:  # zero-flux: ghost = interior
            T[0, 0] = T[n - 1, n - 1] = -1
        return T.tocsr()

    n = [8 * b for b in bpd]
    I = [sp.identity(k) for k in n]
    Lx, Ly, Lz = (lap1d(n[d], bc[d] == PER) for d in range(3))
    G = h0 * (sp.kron(sp.kron(I[2], I[1]), Lx) + sp.kron(sp.kron(I[2], Ly), I[0]) +
              sp.kron(sp.kron(Lz, I[1]), I[0]))
    # permutation global (z,y,x) -> block layout
    perm = np.zeros(O.n, np.int64)
    for b, (_, i, j, k) in enumerate(blocks):
        zz, yy, xx = np.meshgrid(np.arange(8) + 8 * k, np.arange(8) + 8 * j, np.arange(8) + 8 * i,
                                 indexing="ij")
        perm[b * 512:(b + 1) * 512] = ((zz * n[1] + yy) * n[0] + xx).reshape(-1)
    A_ref = G.tocsr()[perm][:, perm]
    A = O.scipy_matrix()
    assert abs(A - A_ref).max() < 1e-15


INDEX_CASES = [((2, 3, 2), (PER, PER, PER)), ((3, 2, 2), (WALL, WALL, WALL)),
               ((2, 2, 3), (PER, WALL, 0))]


@pytest.mark.parametrize("bpd,bc", INDEX_CASES)
def test_index_encoded_field_known_answer(bpd, bc):
    """Neighbour tables, block ordering and periodic wrap: A applied to the
    index-encoded field x = i + Nx j + Nx Ny k equals the closed form of
    helpers.index_encoded_case exactly (nonzero only on the domain faces)."""
    case = uniform(bpd, bc, h0=1.0 / 32)
    x, y = index_encoded_case(*case)
    assert np.array_equal(make(case).spmv(x), y)
    assert np.count_nonzero(y) > 0


@pytest.mark.parametrize("name", list(grid_cases()))
def test_row_sums_zero(name):
    O = make(grid_cases()[name])
    y = O.spmv_raw(np.ones(O.n))
    assert np.abs(y).max() < 1e-14


@pytest.mark.parametrize("name", list(grid_cases()))
def test_conservation_column_sums(name):
    case = grid_cases()[name]
    O = make(case)
    A = O.scipy_matrix()
    colsum = np.asarray(A.sum(axis=0)).ravel()
    assert np.abs(colsum).max() < 1e-14


@pytest.mark.parametrize("name", ["amr2_walls", "amr3_walls", "uniform_walls_3x2x2"])
def test_exact_on_quadratics(name):
    blocks, bpd, h0, bc = case = grid_cases()[name]
    O = make(case)
    X, h = cell_centres(blocks, h0)
    rng = np.random.default_rng(1)
    c = rng.standard_normal(10)
    x, y, z = X[..., 0], X[..., 1], X[..., 2]
    q = (c[0] * x * x + c[1] * y * y + c[2] * z * z + c[3] * x * y + c[4] * y * z + c[5] * x * z
         + c[6] * x + c[7] * y + c[8] * z + c[9])
    lap = 2 * (c[0] + c[1] + c[2])
    Aq = O.spmv_raw(q).reshape(q.shape)
    expect = (h ** 3)[:, None, None, None] * lap * np.ones_like(q)
    mask = ~boundary_cells(blocks, bpd, bc)
    err = np.abs(Aq - expect)[mask].max() / np.abs(expect).max()
    assert err < 1e-10
    if name.startswith("amr"):  # the test has teeth: coarse-fine rows are included
        A = O.scipy_matrix()
        assert np.diff(A.indptr).max() > 12


def test_coarse_fine_rows_known_answer():
    """A fine cell in the middle of a coarse face, hand-derived from
    PoissonSolverExp.cpp:112-190: h*(ghost - self) with
    ghost = 2/3 self - 1/5 inward + 8/15 T, T central in both tangents."""
    # base 2x2x2 blocks, refine block (0,0,0): its x+ face touches coarse block (1,0,0)
    h0 = 1.0 / 16
    blocks = [[1, i, j, k] for k in range(2) for j in range(2) for i in range(2)]
    blocks += [[0, 1, 0, 0], [0, 0, 1, 0], [0, 1, 1, 0], [0, 0, 0, 1], [0, 1, 0, 1],
               [0, 0, 1, 1], [0, 1, 1, 1]]
    blocks = np.array(blocks, np.int32)
    O = pyoracle.Oracle(blocks, (2, 2, 2), h0, (WALL,) * 3, 0)
    A = O.scipy_matrix().tolil()
    hf = h0 / 2
    # fine block (1,1,0,0) (index 1 in the list): x+ face cell (7, iy=2, iz=2)
    fb = 1
    row = fb * 512 + 2 * 64 + 2 * 8 + 7
    coarse = 8  # block (0,1,0,0)
    # coarse cell: ix_c = 0, iy_c = (iy/2) + 4*(index_y odd) = 1, iz_c = 1
    cc = coarse * 512 + 1 * 64 + 1 * 8 + 0
    r = dict(zip(A.rows[row], A.data[row]))
    # in-block neighbours: x-, y+-, z+- each +hf, self -5 hf; face: (2/3-1) hf self, -hf/5 inward
    assert abs(r[row] - (-5 * hf + (2 / 3 - 1) * hf)) < 1e-15
    assert abs(r[row - 1] - (hf - hf / 5)) < 1e-15
    tf = hf * 8 / 15
    # Taylor: iy = 2, iz = 2 are even -> taylorSign -1, -1; coarse (iy_c, iz_c) = (1,1): central
    # T = c + (-1)(D1y) + (-1)(D1z) + D2y + D2z + (+1) Dm
    #   c: 1 - 1/16 - 1/16 ; y+-1: -(+-1/8) + 1/32 ; mixed diagonals +-1/64
    assert abs(r[cc] - tf * (1 - 1 / 16 - 1 / 16)) < 1e-15
    assert abs(r[cc + 8] - tf * (-1 / 8 + 1 / 32)) < 1e-15   # iy_c + 1
    assert abs(r[cc - 8] - tf * (1 / 8 + 1 / 32)) < 1e-15    # iy_c - 1
    assert abs(r[cc + 64] - tf * (-1 / 8 + 1 / 32)) < 1e-15  # iz_c + 1
    assert abs(r[cc + 72] - tf * (1 / 64)) < 1e-15           # (+1, +1)
    assert abs(r[cc + 56] - tf * (-1 / 64)) < 1e-15          # (-1 in y, +1 in z)
    assert len(r) == 6 + 9


@pytest.mark.parametrize("name", ["uniform_periodic_2x3x2", "amr2_walls", "amr3_walls"])
def test_bmean2_rank_one_term(name):
    """bMeanConstraint = 2 adds (sum_b h_b^3 sum x) h^3 to every row
    (BiCGSTAB.cu:392-411). Rows sum to zero, so a constant c gives
    y = c V h^3 (V the domain volume), and a field f of zero h^3-weighted
    mean gives A f unchanged: y(f + c) = A f + c V h^3."""
    blocks, bpd, h0, bc = case = grid_cases()[name]
    O = make(case, bmean=2)
    _, h = cell_centres(blocks, h0)
    h3 = np.repeat(h ** 3, 512)
    V = np.prod([8 * b * h0 for b in bpd])
    assert abs(h3.sum() - V) < 1e-12 * V  # the blocks tile the domain
    f = random_field(len(blocks), 21).reshape(-1)
    f -= h3 @ f / V
    for c in (0.0, 1.0, -2.5):
        y = O.spmv(f + c)
        expect = O.spmv_raw(f) + c * V * h3
        assert np.abs(y - expect).max() <= 1e-12 * np.abs(expect).max()
