"""FP64 drop-in on the GPU vs the CPU oracle, through the C ABI.

Bars: operator and preconditioner <= 1e-13 relative (max norm); solver
<= 1e-10 relative L2 in pressure at the same iteration count wherever the
recurrence does not amplify rounding (bMean = 0 consistent systems; the first
iterations of bMean = 1, 2 and 3), and for converged bMean = 1 and 2 solves
within the tolerance-implied distance of the exact solution.
"""
import numpy as np
import pyoracle
import pytest
import scipy.sparse.linalg as sla
import torch

from cup3d_amd import workloads as W
from cup3d_amd.poisson import PoissonSolver, solver_for
from helpers import (PER, WALL, compatible_rhs, fft_solution_periodic, grid_cases,
                     index_encoded_case, random_field, school4_small, tiny_amr, uniform)

pytestmark = pytest.mark.gpu

CASES = dict(grid_cases())
CASES["tiny_amr3_walls"] = tiny_amr(three_levels=True)
CASES["school4_small"] = school4_small()
CASES["tiny_amr_periodic"] = tiny_amr(periodic=True)


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a, np.float64).reshape(-1)).cuda()


def pair(name, bmean, **kw):
    blocks, bpd, h0, bc = CASES[name]
    S = PoissonSolver(bpd, h0, bc, bmean=bmean, **kw).set_topology(blocks)
    O = pyoracle.Oracle(blocks, bpd, h0, bc, bmean)
    return S, O, blocks, h0


@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("bmean", [0, 1, 2, 3])
def test_operator_matches_oracle(gpu, name, bmean):
    S, O, blocks, h0 = pair(name, bmean)
    x = random_field(len(blocks), 1).reshape(-1)
    y = S.apply_operator(dev(x), torch.zeros(O.n, dtype=torch.float64, device=gpu))
    yo = O.spmv(x)
    assert torch.isfinite(y).all()
    assert np.abs(y.cpu().numpy() - yo).max() <= 1e-13 * np.abs(yo).max()


@pytest.mark.parametrize("bpd,bc", [((2, 3, 2), (PER, PER, PER)), ((3, 2, 2), (WALL,) * 3),
                                    ((2, 2, 3), (PER, WALL, 0))])
def test_operator_index_encoded_field(gpu, bpd, bc):
    """The GPU neighbour tables and block order against a closed form, not the
    oracle: A x for x = i + Nx j + Nx Ny k is exact (integers times a power of
    two) in the FP64 operator and in the inner solve's FP32 operator."""
    import ctypes
    from cup3d_amd.poisson import lib
    blocks, bpd, h0, bc = uniform(bpd, bc, h0=1.0 / 32)
    x, y = index_encoded_case(blocks, bpd, h0, bc)
    S = PoissonSolver(bpd, h0, bc, bmean=0, precision="mixed_fp16").set_topology(blocks)
    y64 = S.apply_operator(dev(x), torch.zeros(len(x), dtype=torch.float64, device=gpu))
    assert np.array_equal(y64.cpu().numpy(), y)
    f = lib().cup3d_debug_apply_operator32
    f.argtypes = [ctypes.c_void_p] * 3
    xd, y32 = dev(x), torch.zeros(len(x), dtype=torch.float64, device=gpu)
    assert f(S.h, xd.data_ptr(), y32.data_ptr()) == 0
    torch.cuda.synchronize()
    assert np.array_equal(y32.cpu().numpy(), y)


@pytest.mark.parametrize("name", ["uniform_walls_3x2x2", "amr3_walls", "amr2_periodic_offcentre"])
def test_preconditioner_matches_oracle(gpu, name):
    S, O, blocks, h0 = pair(name, 1)
    p = random_field(len(blocks), 2).reshape(-1)
    z = S.apply_preconditioner(dev(p), torch.zeros(O.n, dtype=torch.float64, device=gpu))
    zo = O.precond(p)
    assert np.abs(z.cpu().numpy() - zo).max() <= 1e-13 * np.abs(zo).max()


@pytest.mark.parametrize("name", ["uniform_walls_3x2x2", "uniform_periodic_2x3x2", "amr2_walls",
                                  "amr3_walls", "tiny_amr_periodic"])
def test_solve_bmean0_same_iterations(gpu, name):
    S, O, blocks, h0 = pair(name, 0, tol_rel=1e-7)
    b = compatible_rhs(blocks, h0, 3)
    x = torch.zeros(O.n, dtype=torch.float64, device=gpu)
    rep = S.solve_device(dev(b), x)
    xo, info = O.bicgstab(b, tol_rel=1e-7)
    assert rep["status"] == info["status"] == 0
    assert rep["iterations"] == info["iters"] and rep["restarts"] == info["restarts"]
    xg = x.cpu().numpy()
    assert np.linalg.norm(xg - xo) / np.linalg.norm(xo) <= 1e-10
    assert abs(rep["err_init"] - info["err_init"]) <= 1e-13 * info["err_init"]


@pytest.mark.parametrize("name", ["uniform_walls_3x2x2", "amr3_walls", "amr2_periodic_offcentre",
                                  "school4_small"])
def test_dst_preconditioner_matches_oracle(gpu, name):
    """precond = DST: the exact block inverse by sine transforms equals the
    oracle's Cholesky P_inv product (PoissonSolverExp.cpp:17-111, BiCGSTAB.cu:
    526-528) to rounding."""
    blocks, bpd, h0, bc = CASES[name]
    S = PoissonSolver(bpd, h0, bc, bmean=1, precond="dst").set_topology(blocks)
    O = pyoracle.Oracle(blocks, bpd, h0, bc, 1)
    p = random_field(len(blocks), 2).reshape(-1)
    z = S.apply_preconditioner(dev(p), torch.zeros(O.n, dtype=torch.float64, device=gpu))
    zo = O.precond(p)
    assert np.abs(z.cpu().numpy() - zo).max() <= 1e-13 * np.abs(zo).max()


@pytest.mark.parametrize("name", ["uniform_walls_3x2x2", "uniform_periodic_2x3x2", "amr2_walls",
                                  "amr3_walls", "tiny_amr_periodic"])
def test_dst_solve_bmean0_same_iterations(gpu, name):
    """The FP64 drop-in with precond = DST: the reference's iteration, same
    count and iterate as the oracle's (which multiplies by P_inv)."""
    blocks, bpd, h0, bc = CASES[name]
    S = PoissonSolver(bpd, h0, bc, bmean=0, tol_rel=1e-7, precond="dst").set_topology(blocks)
    O = pyoracle.Oracle(blocks, bpd, h0, bc, 0)
    b = compatible_rhs(blocks, h0, 3)
    x = torch.zeros(O.n, dtype=torch.float64, device=gpu)
    rep = S.solve_device(dev(b), x)
    xo, info = O.bicgstab(b, tol_rel=1e-7)
    assert rep["status"] == info["status"] == 0
    assert rep["iterations"] == info["iters"] and rep["restarts"] == info["restarts"]
    assert np.linalg.norm(x.cpu().numpy() - xo) / np.linalg.norm(xo) <= 1e-10


def test_solve_bmean0_school4_small(gpu):
    """On the 4-level school grid the recurrence amplifies rounding even at
    bMean = 0: the oracle against itself under 1e-15 rhs perturbations needs
    107-115 iterations instead of 108 and moves the solution by up to 4e-6
    (test_oracle_solver.py::test_school4_bmean0_rounding_sensitivity). So the
    first 10 iterates agree to rounding, and the converged solve agrees
    within the spread the oracle shows against itself."""
    name = "school4_small"
    S, O, blocks, h0 = pair(name, 0, tol_rel=0.0, max_iter=10)
    b = compatible_rhs(blocks, h0, 3)
    x = torch.zeros(O.n, dtype=torch.float64, device=gpu)
    rep = S.solve_device(dev(b), x)
    xo, info = O.bicgstab(b, tol_rel=0.0, max_iter=10)
    assert rep["iterations"] == info["iters"] == 10
    assert np.linalg.norm(x.cpu().numpy() - xo) <= 1e-10 * np.linalg.norm(xo)
    S, O, blocks, h0 = pair(name, 0, tol_rel=1e-7)
    x = torch.zeros(O.n, dtype=torch.float64, device=gpu)
    rep = S.solve_device(dev(b), x)
    xo, info = O.bicgstab(b, tol_rel=1e-7)
    assert rep["status"] == info["status"] == 0
    assert abs(rep["iterations"] - info["iters"]) <= 0.15 * info["iters"], (rep, info)
    assert np.linalg.norm(x.cpu().numpy() - xo) <= 1e-5 * np.linalg.norm(xo)
    assert rep["true_rel_resid"] <= 1.5e-7 * rep["err_init"] / rep["rhs_norm"]


@pytest.mark.parametrize("name", list(CASES))
def test_solve_bmean1_first_iterations(gpu, name):
    """bMean = 1 amplifies rounding as iterations accumulate (the oracle
    against itself under 1e-15 rhs perturbations stays within 1e-11 up to
    iteration 12 and drifts past 1e-8 by iteration 20,
    test_oracle_solver.py::test_bmean1_recurrence_amplifies_rounding); the
    first 12 iterates agree."""
    S, O, blocks, h0 = pair(name, 1, tol_rel=0.0, max_iter=12)
    b = compatible_rhs(blocks, h0, 4)
    x0 = random_field(len(blocks), 5).reshape(-1) * 1e-3
    x = dev(x0)
    rep = S.solve_device(dev(b), x)
    xo, info = O.bicgstab(b, x0=x0, tol_rel=0.0, max_iter=12)
    assert rep["status"] == info["status"] == 1 and rep["iterations"] == info["iters"] == 12
    xg = x.cpu().numpy()
    assert np.linalg.norm(xg - xo) / np.linalg.norm(xo) <= 1e-10
    assert abs(rep["err_final"] - info["err_final"]) <= 1e-10 * info["err_final"]


@pytest.mark.parametrize("bmean", [2, 3])
@pytest.mark.parametrize("name", ["uniform_walls_3x2x2", "amr3_walls", "school4_small"])
def test_solve_bmean23_first_iterations(gpu, name, bmean):
    """bMeanConstraint = 2 (every row plus (sum h^3 x) h^3) and 3 (the origin
    row replaced by h x, getVec zeroing b and x there): first iterates of the
    GPU solve against the oracle."""
    S, O, blocks, h0 = pair(name, bmean, tol_rel=0.0, max_iter=6)
    b = compatible_rhs(blocks, h0, 6)
    x0 = random_field(len(blocks), 7).reshape(-1) * 1e-3
    x = dev(x0)
    rep = S.solve_device(dev(b), x)
    xo, info = O.bicgstab(b, x0=x0, tol_rel=0.0, max_iter=6)
    assert rep["status"] == info["status"] == 1 and rep["iterations"] == info["iters"] == 6
    assert np.linalg.norm(x.cpu().numpy() - xo) / np.linalg.norm(xo) <= 1e-10


@pytest.mark.parametrize("name", ["uniform_walls_3x2x2", "tiny_amr3_walls"])
def test_solve_bmean2_converged(gpu, name):
    """bMeanConstraint = 2 makes the walls system regular; for a compatible
    rhs the solution is the one with zero h^3-weighted mean. Known answer:
    a direct sparse solve with one cell pinned, shifted to that mean."""
    S, O, blocks, h0 = pair(name, 2, tol_rel=1e-11)
    b = compatible_rhs(blocks, h0, 8)
    x = torch.zeros(O.n, dtype=torch.float64, device=gpu)
    rep = S.solve_device(dev(b), x)
    assert rep["status"] == 0, rep
    A = O.scipy_matrix().tocsr()
    keep = np.ones(O.n, bool)
    keep[0] = False
    xe = np.zeros(O.n)
    xe[keep] = sla.spsolve(A[keep][:, keep].tocsc(), b[keep])
    h3 = np.repeat(O.h ** 3, 512)
    xe -= (h3 @ xe) / h3.sum()
    xg = x.cpu().numpy()
    assert abs(h3 @ xg) <= 1e-7 * np.abs(h3 * xg).sum()
    # residual 1e-11 times a condition number below 1e4
    assert np.linalg.norm(xg - xe) / np.linalg.norm(xe) < 1e-7


def test_solve_bmean1_converged_c1(gpu):
    """32^3 periodic Fourier workload, reference default settings: both
    solvers converge, and both land within the tolerance-implied distance of
    the exact (FFT) solution."""
    w = W.c1_periodic_32()
    S = solver_for(w, tol_rel=1e-7)
    O = pyoracle.Oracle(w.blocks, w.bpd, w.h0, w.bc, 1)
    b = w.rhs().reshape(-1)
    x = torch.zeros(O.n, dtype=torch.float64, device=gpu)
    rep = S.solve_device(dev(b), x)
    xo, info = O.bicgstab(b, tol_rel=1e-7)
    assert rep["status"] == 0 and info["status"] == 0
    assert abs(rep["iterations"] - info["iters"]) <= max(3, info["iters"] // 5)
    xe = fft_solution_periodic(b, w.blocks, w.bpd, w.h0, O.bmean_row)
    for sol in (x.cpu().numpy(), xo):
        assert np.linalg.norm(sol - xe) / np.linalg.norm(xe) < 1e-5
    assert rep["true_rel_resid"] < 2e-7 * rep["err_init"] / rep["rhs_norm"]


def test_host_paths_match_device_path(gpu):
    name = "amr2_walls"
    S, O, blocks, h0 = pair(name, 1, tol_rel=1e-7)
    b = compatible_rhs(blocks, h0, 6)
    x = torch.zeros(O.n, dtype=torch.float64, device=gpu)
    rep = S.solve_device(dev(b), x)
    xh = np.zeros(O.n)
    reph = S.solve_host(b, xh)
    rb = [b[i * 512:(i + 1) * 512].copy() for i in range(len(blocks))]
    xb = [np.zeros(512) for _ in range(len(blocks))]
    repb = S.solve_blocks(rb, xb)
    assert np.array_equal(x.cpu().numpy(), xh)
    assert np.array_equal(xh, np.concatenate(xb))
    assert rep["iterations"] == reph["iterations"] == repb["iterations"]


def test_host_paths_match_device_path_threaded(gpu):
    """8192 blocks (the full uniform walls grid): the host entry point
    gathers and scatters the blocks on several host threads, each over its
    own contiguous range of blocks; host arrays, block pointers and device
    vectors give the same pressure bitwise."""
    w = W.c2_uniform_walls()
    S = solver_for(w, tol_rel=0.0, tol_rhs=1e-6, bmean=0, max_iter=3000)
    b = w.rhs(step=1).reshape(-1)
    x = torch.zeros(w.ncells, dtype=torch.float64, device=gpu)
    rep = S.solve_device(dev(b), x)
    xh = np.zeros(w.ncells)
    reph = S.solve_host(b, xh)
    nb = len(w.blocks)
    rb = [b[i * 512:(i + 1) * 512].copy() for i in range(nb)]
    xb = [np.zeros(512) for _ in range(nb)]
    repb = S.solve_blocks(rb, xb)
    assert rep["status"] == reph["status"] == repb["status"] == 0
    assert np.array_equal(x.cpu().numpy(), xh)
    assert np.array_equal(xh, np.concatenate(xb))
    assert rep["iterations"] == reph["iterations"] == repb["iterations"]


def test_host_staging_follows_topology_change(gpu):
    """The host entry point keeps its staging buffers between solves and
    replaces them when a new topology changes the cell count: host and
    device paths agree bitwise on every topology of the handle."""
    amr = CASES["amr2_walls"]
    uni = uniform((4, 4, 4), (WALL, WALL, WALL))
    assert amr[1:] == uni[1:]
    S = PoissonSolver(amr[1], amr[2], amr[3], bmean=0, tol_rel=1e-9)
    for blocks in (amr[0], uni[0], amr[0]):
        S.set_topology(blocks)
        b = compatible_rhs(blocks, amr[2], len(blocks))
        x = torch.zeros(len(blocks) * 512, dtype=torch.float64, device=gpu)
        rep = S.solve_device(dev(b), x)
        xh = np.zeros(len(blocks) * 512)
        reph = S.solve_host(b, xh)
        assert rep["status"] == reph["status"] == 0
        assert np.array_equal(x.cpu().numpy(), xh)


@pytest.mark.parametrize("name", ["amr2_periodic_offcentre", "amr3_walls", "school4_small"])
def test_coarse_fine_face_count(gpu, name):
    """The face tables find every coarse-fine face: counted here from the block
    list alone (a face whose neighbour across it is not a leaf of the same
    level but exists one level up or down), both sides counted."""
    blocks, bpd, h0, bc = CASES[name]
    S = PoissonSolver(bpd, h0, bc, bmean=0).set_topology(blocks)
    leaves = {tuple(b) for b in np.asarray(blocks).tolist()}
    count = 0
    for (l, i, j, k) in leaves:
        for d in range(3):
            n = bpd[d] << l
            for s in (-1, 1):
                c = [i, j, k]
                c[d] += s
                if not 0 <= c[d] < n:
                    if bc[d] != 1:
                        continue
                    c[d] %= n
                if (l, *c) not in leaves:
                    count += 1
    assert S.coarse_fine_faces() == count > 0


@pytest.mark.parametrize("name", ["uniform_walls_3x2x2", "amr3_walls", "school4_small"])
def test_remove_mean(gpu, name):
    """The caller's projection (PressureProjection.cpp:404-431) on the device:
    the volume-weighted mean it returns is numpy's, and afterwards the field
    has zero weighted mean and differs from the input by that constant."""
    blocks, bpd, h0, bc = CASES[name]
    S = PoissonSolver(bpd, h0, bc, bmean=0).set_topology(blocks)
    h3 = np.repeat((h0 / 2.0 ** np.asarray(blocks)[:, 0]) ** 3, 512)
    x0 = random_field(len(blocks), 71).reshape(-1) + 3.0
    x = dev(x0)
    m = S.remove_mean(x)
    me = (h3 @ x0) / h3.sum()
    assert abs(m - me) <= 1e-14 * abs(me)
    xr = x.cpu().numpy()
    assert np.abs(xr - (x0 - m)).max() == 0.0
    assert abs(h3 @ xr) <= 1e-14 * (h3 @ np.abs(xr))


def test_host_checked_errors(gpu):
    """Bad calls are refused on the host with a message, before any kernel:
    solving before set_topology, null or misaligned vectors, a block list
    that is not 2:1 balanced (which leaves the previous topology in place);
    the handle stays usable afterwards."""
    from cup3d_amd.poisson import lib
    L = lib()
    blocks, bpd, h0, bc = CASES["uniform_walls_3x2x2"]
    S = PoissonSolver(bpd, h0, bc, bmean=0, tol_rel=1e-8)
    n = len(blocks) * 512
    x = torch.zeros(n + 2, dtype=torch.float64, device=gpu)
    b = torch.ones(n + 2, dtype=torch.float64, device=gpu)
    assert L.cup3d_poisson_solve_device(S.h, b.data_ptr(), x.data_ptr(), None) < 0
    assert b"set_topology" in L.cup3d_poisson_last_error(S.h)
    S.set_topology(blocks)
    assert L.cup3d_poisson_solve_device(S.h, None, x.data_ptr(), None) < 0
    assert L.cup3d_poisson_solve_device(S.h, b.data_ptr() + 8, x.data_ptr(), None) < 0
    assert b"aligned" in L.cup3d_poisson_last_error(S.h)
    # every entry point that hands device vectors to a kernel checks them first
    for fn in (L.cup3d_poisson_apply_operator, L.cup3d_poisson_apply_preconditioner,
               L.cup3d_debug_apply_operator32):
        assert fn(S.h, None, x.data_ptr()) < 0
        assert b"null" in L.cup3d_poisson_last_error(S.h)
        assert fn(S.h, b.data_ptr(), x.data_ptr() + 8) < 0
        assert b"aligned" in L.cup3d_poisson_last_error(S.h)
    for fn in (L.cup3d_poisson_cnn_predict, L.cup3d_poisson_cnn_add_sample):
        assert fn(S.h, b.data_ptr(), None, x.data_ptr()) < 0
        assert b"null" in L.cup3d_poisson_last_error(S.h)
        assert fn(S.h, b.data_ptr() + 8, x.data_ptr(), x.data_ptr()) < 0
        assert b"aligned" in L.cup3d_poisson_last_error(S.h)
    assert L.cup3d_poisson_remove_mean(S.h, None, None) < 0
    with pytest.raises(RuntimeError):
        S.set_topology(np.array([[2, 0, 0, 0], [0, 1, 0, 0]], np.int32))
    assert b"balance" in L.cup3d_poisson_last_error(S.h)
    # the rejected block list left the previous topology in place
    rhs = compatible_rhs(blocks, h0, 81)
    xs = torch.zeros(n, dtype=torch.float64, device=gpu)
    assert S.solve_device(dev(rhs), xs)["status"] == 0


def test_deterministic(gpu):
    S, O, blocks, h0 = pair("amr3_walls", 1, tol_rel=1e-7)
    b = dev(compatible_rhs(blocks, h0, 7))
    xs = []
    for _ in range(3):
        x = torch.zeros(O.n, dtype=torch.float64, device=gpu)
        S.solve_device(b, x)
        xs.append(x.cpu().numpy())
    assert np.array_equal(xs[0], xs[1]) and np.array_equal(xs[0], xs[2])


def test_zero_rhs_matches_reference_semantics(gpu):
    """r0 = 0: the reference never sees an improvement (error < error_opt is
    false at 0) and runs to max_iter returning x_opt = x0 (BiCGSTAB.cu:590)."""
    S, O, blocks, h0 = pair("uniform_walls_3x2x2", 1, tol_rel=1e-7, max_iter=20)
    x = torch.zeros(O.n, dtype=torch.float64, device=gpu)
    rep = S.solve_device(torch.zeros(O.n, dtype=torch.float64, device=gpu), x)
    xo, info = O.bicgstab(np.zeros(O.n), tol_rel=1e-7, max_iter=20)
    assert rep["status"] == info["status"] == 1
    assert rep["iterations"] == info["iters"] == 20
    assert (x == 0).all() and (xo == 0).all()


def test_nonfinite_rhs_reports_status(gpu):
    S, O, blocks, h0 = pair("uniform_walls_3x2x2", 1, tol_rel=1e-7)
    b = compatible_rhs(blocks, h0, 8)
    b[100] = np.nan
    x = torch.zeros(O.n, dtype=torch.float64, device=gpu)
    rep = S.solve_device(dev(b), x)
    assert rep["status"] == 3


def test_c2_full_size_properties(gpu):
    """256x128x128 walls workload at full size: converges to the true relative
    residual asked for, deterministically, with the mean constraint met. (With
    bMean = 1 the recurrence can stagnate near 1e-7 on this grid, and whether
    it gets below depends on rounding — the oracle shows the same on smaller
    grids — so the target here is 1e-6.)"""
    w = W.c2_uniform_walls()
    S = solver_for(w, tol_rel=0.0, tol_rhs=1e-6, max_iter=3000)
    b = dev(w.rhs(step=1))
    runs = []
    for _ in range(2):
        x = torch.zeros(w.ncells, dtype=torch.float64, device=gpu)
        rep = S.solve_device(b, x)
        runs.append((rep, x.clone()))
    rep, x = runs[0]
    assert rep["status"] == 0
    assert rep["true_rel_resid"] <= 1.05e-6
    assert torch.equal(runs[0][1], runs[1][1])
    # the mean-constraint row is part of the residual: |sum h^3 x| <= ||r||
    assert abs(float(x.sum()) * w.h0 ** 3) <= 1.05e-6 * rep["rhs_norm"]


def test_c5_full_size_first_iterations_match_oracle(gpu):
    """The FP64 drop-in on the full 4-level AMR school grid (59 012 blocks,
    30.2 M cells, bMean = 0, step-0 rhs, zero guess) against the oracle's
    classical BiCGSTAB on the same system (tests/golden/c5_fp64_fixture.json,
    made by tests/golden/make_c5_fp64_fixture.py, about ten minutes of CPU).
    The oracle sums its norms and dots over the 30.2 M cells one after the
    other, whose rounding is bounded by n eps = 3.4e-9 (measured on the
    initial residual: 8e-11), the GPU in a fixed tree: the initial residual
    and, after 10 iterations, the best residual agree to 5e-9 and the best
    iterate at 64 cells to 1e-8.

    Later the recurrence amplifies rounding. The fixture holds the oracle's
    own residual history on the rhs perturbed by 1e-15 and 1e-11 relative:
    the larger of the two moves the best residual by s_k = 2.0e-12, 4.6e-8
    and 1.1e-3 after k = 10, 20 and 30 iterations. The GPU's deviation d_k
    must grow no faster than that: d_20 <= 10 (d_10 / s_10) s_20 (measured:
    d_10 / s_10 = 80 and d_20 / s_20 = 70, the GPU rounding like a larger
    perturbation of the same recurrence), and d_30 <= 10 s_30, where both
    have saturated (measured: d_30 = 1.9e-3)."""
    import json
    import os
    fx = json.load(open(os.path.join(os.path.dirname(__file__), "golden",
                                     "c5_fp64_fixture.json")))
    w = W.c5_amr_school()
    assert w.ncells == fx["cells"]
    b = dev(w.rhs(step=0).reshape(-1))
    hist = fx["history"]

    def spread(k):  # the oracle against itself under the perturbations
        best = min(hist[:k])
        return max(abs(min(h[:k]) - best) / best for h in fx["history_perturbed"].values())

    dev10 = None
    for k in (10, 20, 30):
        S = solver_for(w, tol_rel=0.0, bmean=0, precision="fp64", max_iter=k)
        x = torch.zeros(w.ncells, dtype=torch.float64, device=gpu)
        rep = S.solve_device(b, x)
        best = min(hist[:k])
        d = abs(rep["err_final"] - best) / best
        if k == 10:
            bar = 5e-9
            dev10 = d
        elif k == 20:
            bar = max(1e-6, 10 * max(dev10, spread(10)) / spread(10) * spread(20))
        else:
            bar = 10 * spread(30)
        print(f"k={k}: GPU best residual {rep['err_final']:.12e}, oracle {best:.12e}, "
              f"deviation {d:.2e}, oracle spread {spread(k):.2e}, bar {bar:.2e}")
        assert rep["iterations"] == k
        assert abs(rep["err_init"] - fx["err_init"]) <= 5e-9 * fx["err_init"]
        assert d <= bar
        if k == 10:
            got = x[torch.tensor(fx["x10_index"], device=gpu)].cpu().numpy()
            want = np.asarray(fx["x10_values"])
            assert np.abs(got - want).max() <= 1e-8 * np.abs(want).max()
        del S, x
