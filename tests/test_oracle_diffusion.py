"""Pins the oracle's implicit-diffusion restatement (AdvectionDiffusionImplicit.h
and .cpp, DiffusionSolverAMRKernels.cpp).

Operator (KernelLHSDiffusion, h (sum lab - 6 x) - h^3 x / (nu dt), one
velocity component): exact eigenpairs on uniform grids for every boundary
class of BlockLabBC (periodic Fourier modes; Dirichlet sine modes for walls
and for the normal component at freespace faces; Neumann cosine modes for
the tangential components at freespace faces). On AMR grids the lab-form
operator (ghosts + flux correction) equals the pressure operator's assembled
rows (already pinned) plus the Dirichlet faces' -2 h x and the shift: two
independent code paths. Preconditioner: the exact per-level block inverse
against a numpy dense solve of the Kronecker-sum block matrix; the
reference's block CG against it. Solvers: the FFT closed form on a periodic
grid, the residual of the pinned operator, the pipelined (reference)
algorithm against the classical one. Step: a single Fourier mode decays by
the backward-Euler factor 1 / (1 - nu dt lambda / h^2).
"""
import numpy as np
import pyoracle
import pytest

from helpers import (FREE, PER, WALL, blocks_to_global, global_to_blocks, grid_cases,
                     random_field, school4_small, tiny_amr, uniform)


def make(case):
    blocks, bpd, h0, bc = case
    return pyoracle.Oracle(blocks, bpd, h0, bc, 0)


def global_index(blocks):
    r = np.arange(8)
    iz, iy, ix = np.meshgrid(r, r, r, indexing="ij")
    return [blocks[:, 1, None, None, None] * 8 + ix[None],
            blocks[:, 2, None, None, None] * 8 + iy[None],
            blocks[:, 3, None, None, None] * 8 + iz[None]]


def mode_1d(kind, m, i, n):
    """(values, eigenvalue of the 1-D second difference) for one axis."""
    if kind == "periodic":
        return np.cos(2 * np.pi * m * i / n + 0.3), 2 * np.cos(2 * np.pi * m / n) - 2
    if kind == "dirichlet":  # ghost = -x: zero on the face half a cell out
        return np.sin(np.pi * m * (i + 0.5) / n), 2 * np.cos(np.pi * m / n) - 2
    return np.cos(np.pi * m * (i + 0.5) / n), 2 * np.cos(np.pi * m / n) - 2  # neumann


def axis_kind(bc, d, comp):
    if bc[d] == PER:
        return "periodic"
    if bc[d] == WALL or d == comp:
        return "dirichlet"
    return "neumann"


@pytest.mark.parametrize("bc", [(PER, PER, PER), (WALL, WALL, WALL), (FREE, FREE, FREE),
                                (PER, WALL, FREE)])
@pytest.mark.parametrize("comp", [0, 1, 2])
def test_uniform_eigenpairs(bc, comp):
    bpd = (2, 3, 2)
    case = uniform(bpd, bc)
    blocks, _, h0, _ = case
    O = make(case)
    g = global_index(blocks)
    n = [8 * b for b in bpd]
    sigma = 37.0
    for m in [(1, 2, 1), (3, 1, 4), (5, 7, 2)]:
        f = np.ones((len(blocks), 8, 8, 8))
        lam = 0.0
        for d in range(3):
            v, l1 = mode_1d(axis_kind(bc, d, comp), m[d], g[d], n[d])
            f = f * v
            lam += l1
        y = O.diffusion_lhs(f, comp, sigma).reshape(f.shape)
        want = (h0 * lam - h0 ** 3 * sigma) * f
        assert np.abs(y - want).max() <= 1e-13 * np.abs(want).max()


def dirichlet_faces(blocks, bpd, bc, comp):
    """Per cell: the number of its faces on a domain boundary whose velocity
    ghost is -u for component `comp`."""
    blocks = np.asarray(blocks)
    g = global_index(blocks)
    nd = np.zeros((len(blocks), 8, 8, 8))
    for d in range(3):
        if bc[d] == PER or not (bc[d] == WALL or d == comp):
            continue
        n = (np.asarray(bpd[d]) << blocks[:, 0])[:, None, None, None] * 8
        nd += (g[d] == 0) + (g[d] == n - 1)
    return nd


CASES = dict(grid_cases())
CASES["tiny_amr_free"] = tiny_amr()[:3] + ((FREE, FREE, FREE),)
CASES["school4_small"] = school4_small()


@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("comp", [0, 2])
def test_lab_form_equals_assembled_rows(name, comp):
    """Ghosts + flux correction (the Cubism form of KernelLHSDiffusion) give
    the pressure operator's rows (getMat) plus the Dirichlet faces and the
    shift, cell for cell."""
    blocks, bpd, h0, bc = CASES[name]
    O = make(CASES[name])
    x = random_field(len(blocks), seed=3).reshape(-1)
    sigma = 5.0
    h = (h0 / 2.0 ** np.asarray(blocks)[:, 0])[:, None, None, None]
    diag = (-2 * h * dirichlet_faces(blocks, bpd, bc, comp) - h ** 3 * sigma).reshape(-1)
    want = O.spmv_raw(x) + diag * x
    got = O.diffusion_lhs(x, comp, sigma)
    assert np.abs(got - want).max() <= 1e-12 * np.abs(want).max()


def test_lap_vec_is_componentwise():
    blocks, bpd, h0, bc = CASES["amr2_walls"]
    O = make(CASES["amr2_walls"])
    vel = np.random.default_rng(4).standard_normal((len(blocks) * 512, 3))
    for c in range(3):
        assert np.array_equal(O.diffusion_lap_vec(vel, c),
                              O.diffusion_lhs(np.ascontiguousarray(vel[:, c]), c, 0.0))


def block_matrix(coef):
    """A_loc: `coef` on the diagonal, 1 for face neighbours inside an 8^3
    block (zero padding), as a Kronecker sum."""
    T = np.diag(np.ones(7), 1) + np.diag(np.ones(7), -1)
    I = np.eye(8)
    adj = np.kron(np.kron(I, I), T) + np.kron(np.kron(I, T), I) + np.kron(np.kron(T, I), I)
    return adj + coef * np.eye(512)


def test_precond_exact_and_block_cg():
    blocks, bpd, h0, bc = CASES["amr3_walls"]
    O = make(CASES["amr3_walls"])
    sigma = 900.0
    p = random_field(len(blocks), seed=5)
    z = O.diffusion_precond(p, sigma).reshape(p.shape)
    z_cg = O.diffusion_block_cg(p, sigma).reshape(p.shape)
    h = h0 / 2.0 ** np.asarray(blocks)[:, 0]
    for lev in np.unique(blocks[:, 0]):
        A = block_matrix(-6.0 - h[blocks[:, 0] == lev][0] ** 2 * sigma)
        sel = np.nonzero(blocks[:, 0] == lev)[0]
        for b in sel[:4]:
            want = np.linalg.solve(A, p[b].reshape(-1) / h[b])
            assert np.abs(z[b].reshape(-1) - want).max() <= 1e-12 * np.abs(want).max()
            # the reference's CG stops at 1e-7 of its initial residual
            assert np.linalg.norm(z_cg[b].reshape(-1) - want) <= 1e-6 * np.linalg.norm(want)


def fft_solve_periodic(b, blocks, bpd, h0, sigma):
    g = blocks_to_global(b.reshape(len(blocks), 8, 8, 8), blocks, bpd)
    n = g.shape
    lam = sum(np.reshape(2 * np.cos(2 * np.pi * np.arange(n[a]) / n[a]) - 2,
                         [-1 if d == a else 1 for d in range(3)]) for a in range(3))
    x = np.real(np.fft.ifftn(np.fft.fftn(g) / (h0 * lam - h0 ** 3 * sigma)))
    return global_to_blocks(x, blocks).reshape(-1)


def test_classical_and_pipelined_match_fft_solution():
    case = uniform((2, 3, 2), (PER, PER, PER))
    blocks, bpd, h0, _ = case
    O = make(case)
    sigma = 2.0 / h0 ** 2  # h^2 / (nu dt) = 2
    b = random_field(len(blocks), seed=6).reshape(-1)
    want = fft_solve_periodic(b, blocks, bpd, h0, sigma)
    x, info = O.diffusion_bicgstab(b, 0, sigma, tol_rel=1e-13, max_iter=200)
    assert info["status"] == 0
    assert np.linalg.norm(x - want) <= 1e-10 * np.linalg.norm(want)
    xp, infop = O.diffusion_pipelined(b, 0, sigma, tol_rel=1e-12, max_iter=200)
    assert infop["status"] == 0
    assert np.linalg.norm(xp - want) <= 1e-9 * np.linalg.norm(want)


@pytest.mark.parametrize("name", ["amr3_walls", "tiny_amr_free", "school4_small"])
def test_solvers_reach_the_residual_on_amr(name):
    blocks, bpd, h0, bc = CASES[name]
    O = make(CASES[name])
    sigma = 1.0 / (1e-3 * 0.05)  # nu = 1e-3, dt = 0.05
    b = random_field(len(blocks), seed=7).reshape(-1)
    bn = np.linalg.norm(b)
    for comp in (0, 1):
        x, info = O.diffusion_bicgstab(b, comp, sigma, tol_rel=1e-12, max_iter=500)
        assert info["status"] == 0
        r = b - O.diffusion_lhs(x, comp, sigma)
        assert np.linalg.norm(r) <= 1e-11 * bn
        xp, infop = O.diffusion_pipelined(b, comp, sigma, tol_rel=1e-11, max_iter=500)
        assert infop["status"] == 0
        assert np.linalg.norm(b - O.diffusion_lhs(xp, comp, sigma)) <= 1e-10 * bn
        assert np.linalg.norm(xp - x) <= 1e-8 * np.linalg.norm(x)


def test_step_decays_a_fourier_mode_by_the_backward_euler_factor():
    case = uniform((2, 2, 3), (PER, PER, PER))
    blocks, bpd, h0, _ = case
    O = make(case)
    g = global_index(blocks)
    n = [8 * b for b in bpd]
    nu, dt = 2e-3, 0.1
    m = (1, 3, 2)
    lam = sum(2 * np.cos(2 * np.pi * m[d] / n[d]) - 2 for d in range(3))
    phase = sum(2 * np.pi * m[d] * g[d] / n[d] for d in range(3))
    u_old = np.stack([np.cos(phase), 0.5 * np.sin(phase), -np.cos(phase + 1.0)], -1).reshape(-1, 3)
    u_new, info = O.diffusion_step(u_old, u_old, nu, dt, tol_abs=0.0, tol_rel=1e-12)
    want = u_old / (1.0 - nu * dt * lam / h0 ** 2)
    assert np.abs(u_new - want).max() <= 1e-10 * np.abs(want).max()
    assert all(i >= 1 for i in info["iters"])


@pytest.mark.parametrize("name", ["amr2_walls", "uniform_mixed_bc"])
def test_step_solves_backward_euler(name):
    """u^{n+1} - u* solves A_diff x = -A_lap u* + h^3 (u* - u^n) / (nu dt):
    equivalently A_lap u^{n+1} = h^3 (u^{n+1} - u^n) / (nu dt)."""
    blocks, bpd, h0, bc = CASES[name]
    O = make(CASES[name])
    n = len(blocks) * 512
    rng = np.random.default_rng(8)
    u_old = rng.standard_normal((n, 3))
    u_star = u_old + 0.1 * rng.standard_normal((n, 3))
    nu, dt = 1e-3, 0.02
    u_new, _ = O.diffusion_step(u_old, u_star, nu, dt, tol_abs=0.0, tol_rel=1e-12)
    h3 = np.repeat((h0 / 2.0 ** np.asarray(blocks)[:, 0]) ** 3, 512)
    s = h3 / (nu * dt)
    for c in range(3):
        lap = O.diffusion_lap_vec(u_new, c)
        res = lap - s * (u_new[:, c] - u_old[:, c])
        b = -O.diffusion_lap_vec(u_star, c) + s * (u_star[:, c] - u_old[:, c])
        assert np.linalg.norm(res) <= 1e-10 * np.linalg.norm(b)
