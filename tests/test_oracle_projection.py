"""Pins the oracle's restatement of the caller around the solve:
KernelPressureRHS (PressureProjection.cpp:96-199) and KernelGradP with the
velocity update (:244-323, 449-466), flux correction at coarse-fine faces
included. What the mathematics fixes:

* on a linear velocity field the rhs is h^3 div u exactly, and on a linear
  pressure the velocity changes by -dt grad p exactly, at every cell off the
  domain walls: same-level, fine-side and coarse-side (flux-corrected) cells
  alike (the interface interpolation is exact for quadratics, and the four
  fine face terms of a coarse face sum to the coarse one for linear fields);
* the rhs is conservative: summed over all cells it is the net flux through
  the domain boundary, which is zero for periodic axes and for walls (the
  mirrored normal velocity carries no flux), for any velocity field;
* the pressure-gradient terms telescope the same way: sum_i h_i^3 dv_i is the
  boundary term, zero on a periodic grid;
* the obstacle term is -chi fac div(udef): with chi = 1 everywhere it is the
  rhs of u - udef.
"""
import numpy as np
import pyoracle
import pytest

from helpers import PER, boundary_cells, cell_centres, grid_cases, school4_small, tiny_amr

CASES = dict(grid_cases())
CASES["tiny_amr3_walls"] = tiny_amr(three_levels=True)
CASES["tiny_amr_periodic"] = tiny_amr(periodic=True)
CASES["school4_small"] = school4_small()
WALLED = [k for k, v in CASES.items() if all(b != PER for b in v[3])]


def linear_field(X, seed):
    """u_a = c_a + G_a . x: a random linear vector field and its divergence."""
    rng = np.random.default_rng(seed)
    G = rng.standard_normal((3, 3))
    c = rng.standard_normal(3)
    u = c + np.einsum("...b,ab->...a", X, G)
    return u, np.trace(G), G


def off_walls(blocks, bpd, bc):
    """Cells not on a non-periodic domain face."""
    return ~boundary_cells(blocks, bpd, bc)


@pytest.mark.parametrize("name", WALLED)
def test_rhs_exact_on_linear_velocity(name):
    blocks, bpd, h0, bc = CASES[name]
    O = pyoracle.Oracle(blocks, bpd, h0, bc, 0)
    X, h = cell_centres(blocks, h0)
    u, div, _ = linear_field(X, 1)
    dt = 0.37
    b = O.pressure_rhs(u.reshape(-1, 3), dt).reshape(len(blocks), 8, 8, 8)
    want = np.broadcast_to((h ** 3)[:, None, None, None] * div / dt, b.shape)
    m = off_walls(blocks, bpd, bc)
    assert np.abs(b[m] - want[m]).max() <= 1e-12 * np.abs(want).max()


@pytest.mark.parametrize("name", list(CASES))
def test_rhs_conservative(name):
    blocks, bpd, h0, bc = CASES[name]
    O = pyoracle.Oracle(blocks, bpd, h0, bc, 0)
    u = np.random.default_rng(2).standard_normal((len(blocks) * 512, 3))
    b = O.pressure_rhs(u, 1.0)
    scale = (0.5 * (h0 / 2.0 ** np.asarray(blocks)[:, 0]) ** 2).max() * np.abs(u).max()
    assert abs(b.sum()) <= 1e-13 * scale * b.size
    assert np.abs(b).max() > 0.1 * scale


def test_rhs_obstacle_term():
    """chi = 1 everywhere: b(u, udef) = b(u - udef); chi = 0: b(u, udef) = b(u)."""
    blocks, bpd, h0, bc = CASES["amr3_walls"]
    O = pyoracle.Oracle(blocks, bpd, h0, bc, 0)
    rng = np.random.default_rng(3)
    u = rng.standard_normal((len(blocks) * 512, 3))
    ud = rng.standard_normal((len(blocks) * 512, 3))
    one, zero = np.ones(len(blocks) * 512), np.zeros(len(blocks) * 512)
    b1 = O.pressure_rhs(u, 0.5, udef=ud, chi=one)
    assert np.abs(b1 - O.pressure_rhs(u - ud, 0.5)).max() <= 1e-13 * np.abs(b1).max()
    b0 = O.pressure_rhs(u, 0.5, udef=ud, chi=zero)
    assert np.array_equal(b0, O.pressure_rhs(u, 0.5))


@pytest.mark.parametrize("name", WALLED)
def test_grad_p_exact_on_linear_pressure(name):
    blocks, bpd, h0, bc = CASES[name]
    O = pyoracle.Oracle(blocks, bpd, h0, bc, 0)
    X, h = cell_centres(blocks, h0)
    rng = np.random.default_rng(4)
    g, c = rng.standard_normal(3), rng.standard_normal()
    p = c + X @ g
    v0 = rng.standard_normal((len(blocks) * 512, 3))
    dt = 0.8
    v = O.grad_p_update(p.reshape(-1), v0, dt).reshape(len(blocks), 8, 8, 8, 3)
    dv = v - v0.reshape(v.shape)
    m = off_walls(blocks, bpd, bc)
    assert np.abs(dv[m] + dt * g).max() <= 1e-11 * dt * np.abs(g).max()


@pytest.mark.parametrize("name", ["uniform_periodic_2x3x2", "amr2_periodic_offcentre",
                                  "tiny_amr_periodic"])
def test_grad_p_telescopes_on_periodic_grids(name):
    blocks, bpd, h0, bc = CASES[name]
    O = pyoracle.Oracle(blocks, bpd, h0, bc, 0)
    rng = np.random.default_rng(5)
    p = rng.standard_normal(len(blocks) * 512)
    v0 = np.zeros((len(blocks) * 512, 3))
    dv = O.grad_p_update(p, v0, 1.0)
    h3 = np.repeat((h0 / 2.0 ** np.asarray(blocks)[:, 0]) ** 3, 512)
    tot = (h3[:, None] * dv).sum(0)
    scale = (h3[:, None] * np.abs(dv)).sum(0)
    assert np.all(np.abs(tot) <= 1e-13 * scale)


def test_grad_p_walls_boundary_term():
    """Walls: the zero-gradient ghost makes each boundary face contribute
    fac (p_b + p_b); sum_i h^3 dv_a = -dt h^2 / 2 (sum over +a faces of 2 p_b
    - sum over -a faces of 2 p_b)."""
    blocks, bpd, h0, bc = CASES["uniform_walls_3x2x2"]
    O = pyoracle.Oracle(blocks, bpd, h0, bc, 0)
    p = np.random.default_rng(6).standard_normal(len(blocks) * 512)
    dv = O.grad_p_update(p, np.zeros((len(blocks) * 512, 3)), 1.0)
    from helpers import blocks_to_global
    P = blocks_to_global(p.reshape(-1, 8, 8, 8), blocks, bpd)  # [z][y][x]
    h = h0
    want = -0.5 * h * h * np.array([
        2 * (P[:, :, -1].sum() - P[:, :, 0].sum()),
        2 * (P[:, -1, :].sum() - P[:, 0, :].sum()),
        2 * (P[-1, :, :].sum() - P[0, :, :].sum())])
    got = (h ** 3 * dv).sum(0)
    assert np.allclose(got, want, rtol=1e-12, atol=1e-14 * np.abs(want).max())
