"""Warm start from the previous steps on the GPU (cfg.extrapolate): the
linear extrapolation in time 2 p_{t-1} - p_{t-2} (K = 1) and the
least-squares combination of the differences of the last K = 3, 5, 6 and 8
pressures, checked against what the mathematics fixes — exactness on
sequences that are polynomials of degree K in time, the least-squares optimum and
the residual it reaches computed independently with the oracle's operator,
and never a worse residual than the given guess."""
import numpy as np
import pyoracle
import pytest
import torch

from cup3d_amd.poisson import PoissonSolver
from helpers import grid_cases, random_field, tiny_amr

pytestmark = pytest.mark.gpu

CASES = dict(grid_cases())
CASES["tiny_amr3_walls"] = tiny_amr(three_levels=True)


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a, np.float64).reshape(-1)).cuda()


def mean_free(f, blocks, h0):
    h3 = np.repeat((h0 / 2.0 ** np.asarray(blocks)[:, 0]) ** 3, 512)
    f = np.asarray(f, np.float64).reshape(-1)
    return f - (h3 @ f) / h3.sum()


@pytest.mark.parametrize("prec", ["fp64", "mixed_fp16"])
@pytest.mark.parametrize("name", ["uniform_walls_3x2x2", "tiny_amr3_walls"])
def test_extrapolation_exact_on_linear_sequence(gpu, name, prec):
    """p_t = p0 + t d and b_t = A p_t: from the third step on, the extrapolated
    guess is the solution up to the previous solves' error, with weight 1."""
    blocks, bpd, h0, bc = CASES[name]
    O = pyoracle.Oracle(blocks, bpd, h0, bc, 0)
    p0 = mean_free(random_field(len(blocks), 31), blocks, h0)
    d = 0.1 * mean_free(random_field(len(blocks), 32), blocks, h0)
    S = PoissonSolver(bpd, h0, bc, bmean=0, tol_rel=0.0, tol_rhs=1e-10, precision=prec,
                      max_iter=3000, extrapolate=True).set_topology(blocks)
    x = torch.zeros(O.n, dtype=torch.float64, device=gpu)
    for t in range(5):
        b = O.spmv(p0 + t * d)
        prev = x.cpu().numpy()
        rep = S.solve_device(dev(b), x)
        assert rep["status"] == 0, (t, rep)
        r_prev = np.linalg.norm(b - O.spmv(prev))
        if t == 0:
            assert rep["guess_weight_time"] == 0.0  # no history yet
        elif t >= 2:  # (t = 1 extrapolates from the zero guess of t = 0)
            assert abs(rep["guess_weight_time"] - 1.0) < 1e-6, rep
            assert rep["guess_resid"] < 1e-7 * r_prev, (rep["guess_resid"], r_prev)
        xm = mean_free(x.cpu().numpy(), blocks, h0)
        x.copy_(dev(xm))  # the caller's projection (PressureProjection.cpp:404-431)


@pytest.mark.parametrize("name", ["amr3_walls", "tiny_amr3_walls"])
def test_extrapolation_weight_is_the_least_squares_step(gpu, name):
    """On a sequence that is not linear in time the weight is the minimiser of
    |b - A (x + g (x - x_prev))| and the reported guess residual its value,
    both recomputed on the CPU with the oracle's operator (or 0 and the given
    guess's residual when the step lowers the squared residual by less than
    1 %); the guess is never worse than the given one."""
    blocks, bpd, h0, bc = CASES[name]
    O = pyoracle.Oracle(blocks, bpd, h0, bc, 0)
    S = PoissonSolver(bpd, h0, bc, bmean=0, tol_rel=0.0, tol_rhs=1e-9, max_iter=3000,
                      extrapolate=True).set_topology(blocks)
    rng = np.random.default_rng(5)
    base = [mean_free(random_field(len(blocks), 40 + k), blocks, h0) for k in range(3)]
    x = torch.zeros(O.n, dtype=torch.float64, device=gpu)
    hist = None
    taken = 0
    for t in range(4):
        c = [np.cos(0.7 * t + k) for k in range(3)] + rng.uniform(-0.05, 0.05, 3)
        b = O.spmv(sum(ck * bk for ck, bk in zip(c, base)))
        xin = x.cpu().numpy()
        rep = S.solve_device(dev(b), x)
        assert rep["status"] == 0
        if hist is not None:
            r0 = b - O.spmv(xin)
            u = O.spmv(xin - hist)
            g = (r0 @ u) / (u @ u)
            resid = np.linalg.norm(r0 - g * u)
            if resid ** 2 >= 0.99 * (r0 @ r0):  # less than 1 %: the given guess is kept
                g, resid = 0.0, np.linalg.norm(r0)
            assert abs(rep["guess_weight_time"] - g) <= 1e-9 * max(1.0, abs(g)), (rep, g)
            assert abs(rep["guess_resid"] - resid) <= 1e-9 * resid
            assert rep["guess_resid"] <= np.linalg.norm(r0)
            taken += g != 0.0
        hist = xin
        x.copy_(dev(mean_free(x.cpu().numpy(), blocks, h0)))
    assert taken >= 2  # the test has teeth: the extrapolation was used


def test_set_topology_clears_the_history(gpu):
    blocks, bpd, h0, bc = CASES["uniform_walls_3x2x2"]
    O = pyoracle.Oracle(blocks, bpd, h0, bc, 0)
    S = PoissonSolver(bpd, h0, bc, bmean=0, tol_rel=0.0, tol_rhs=1e-10,
                      extrapolate=True).set_topology(blocks)
    p0 = mean_free(random_field(len(blocks), 51), blocks, h0)
    d = mean_free(random_field(len(blocks), 52), blocks, h0)
    x = torch.zeros(O.n, dtype=torch.float64, device=gpu)
    weights = []
    for t in range(4):
        if t == 2:
            S.set_topology(blocks)
        rep = S.solve_device(dev(O.spmv(p0 + t * d)), x)
        assert rep["status"] == 0
        weights.append(rep["guess_weight_time"])
    # t = 0: no history; t = 2: history cleared (t = 1 extrapolates from the
    # zero guess of t = 0, a direction that does not reach the 1 % gain here)
    assert weights[0] == 0.0 and weights[2] == 0.0
    assert abs(weights[3] - 1.0) < 1e-6


def test_extrapolation_off_keeps_the_reference_contract(gpu):
    """Without cfg.extrapolate the given x is the guess (no history kept)."""
    blocks, bpd, h0, bc = CASES["uniform_walls_3x2x2"]
    O = pyoracle.Oracle(blocks, bpd, h0, bc, 0)
    S = PoissonSolver(bpd, h0, bc, bmean=0, tol_rel=0.0, tol_rhs=1e-10).set_topology(blocks)
    p0 = mean_free(random_field(len(blocks), 61), blocks, h0)
    d = mean_free(random_field(len(blocks), 62), blocks, h0)
    x = torch.zeros(O.n, dtype=torch.float64, device=gpu)
    for t in range(3):
        b = O.spmv(p0 + t * d)
        r_in = np.linalg.norm(b - O.spmv(x.cpu().numpy()))
        rep = S.solve_device(dev(b), x)
        assert rep["guess_weight_time"] == 0.0
        assert abs(rep["err_init"] - r_in) <= 1e-12 * r_in


@pytest.mark.parametrize("K", [3, 5])
@pytest.mark.parametrize("name", ["uniform_walls_3x2x2", "tiny_amr3_walls"])
def test_history_exact_on_polynomial_sequence(gpu, name, K):
    """cfg.extrapolate = K on p_t = p0 + sum_{j=1..K} t^j d_j: once the handle
    holds K previous pressures (from solve K + 1 on; the first solve is handed
    the zero guess), the extrapolation of degree K, sum_{j=1..K+1} (-1)^(j+1)
    C(K+1, j) p_{t-j}, lies in the span of the K differences, and the
    least-squares guess reaches it: weight C(K, 1) = K on x - h1 (K = 3: the
    weights (3, -3, 1) of the cubic extrapolation 4 x - 6 h1 + 4 h2 - h3),
    residual below 1e-6 of the given guess's."""
    blocks, bpd, h0, bc = CASES[name]
    O = pyoracle.Oracle(blocks, bpd, h0, bc, 0)
    p0 = mean_free(random_field(len(blocks), 71), blocks, h0)
    d = [0.1 * 0.2 ** k * mean_free(random_field(len(blocks), 72 + k), blocks, h0)
         for k in range(K)]
    S = PoissonSolver(bpd, h0, bc, bmean=0, tol_rel=0.0, tol_rhs=1e-11, max_iter=3000,
                      extrapolate=K).set_topology(blocks)
    x = torch.zeros(O.n, dtype=torch.float64, device=gpu)
    for t in range(K + 4):
        b = O.spmv(p0 + sum(t ** (j + 1) * d[j] for j in range(K)))
        prev = x.cpu().numpy()
        rep = S.solve_device(dev(b), x)
        assert rep["status"] == 0, (t, rep)
        r_prev = np.linalg.norm(b - O.spmv(prev))
        if t >= K + 1:
            assert abs(rep["guess_weight_time"] - K) < 1e-4 * K, (t, rep)
            # (the guess residual comes from |r0|^2 - g.f, exact to ~1e-8 |r0|)
            assert rep["guess_resid"] < 1e-6 * r_prev, (t, rep["guess_resid"], r_prev)
        x.copy_(dev(mean_free(x.cpu().numpy(), blocks, h0)))


NOISE_WEIGHT = 200.0  # the library's weight of the history-noise estimate


@pytest.mark.parametrize("tol", [1e-10, 1e-3])
@pytest.mark.parametrize("K", [3, 6, 8])
@pytest.mark.parametrize("name", ["amr3_walls", "tiny_amr3_walls"])
def test_history_weights_are_the_least_squares_solution(gpu, name, K, tol):
    """On a smooth sequence that is no polynomial in time, the chosen depth,
    the reported guess residual and the weight of x - h1 are those of the
    least-squares problems over the nested direction sets {d1}, {d1, d2},
    ..., {d1 .. dK}, each scored by its squared residual plus NOISE_WEIGHT^2
    times the history noise sum_j c_j^2 rho_j^2 (c: the option's coefficients
    on x, h1 .. hm; rho: the true residual of the solve each pressure came
    from), a larger set taken only if it lowers the score by at least 1 %;
    recomputed on the CPU with the oracle's operator; never worse than the
    given guess. Solved to 1e-10 the noise is negligible and every direction
    is taken; solved to 1e-3 it decides the depth."""
    blocks, bpd, h0, bc = CASES[name]
    O = pyoracle.Oracle(blocks, bpd, h0, bc, 0)
    S = PoissonSolver(bpd, h0, bc, bmean=0, tol_rel=0.0, tol_rhs=tol, max_iter=3000,
                      extrapolate=K).set_topology(blocks)
    # twelve components with their own frequencies: no K differences span
    # the sequence exactly (the residuals stay well above rounding)
    nb = 12
    base = [mean_free(random_field(len(blocks), 80 + k), blocks, h0) for k in range(nb)]
    x = torch.zeros(O.n, dtype=torch.float64, device=gpu)
    handed, rho_handed = [], []
    rho_last = 0.0  # true residual of the last solve (none yet)
    deep, noise_decided = 0, 0
    for t in range(K + 5):
        c = [np.cos(0.3 * (1 + 0.25 * k) * t + k) for k in range(nb)]
        b = O.spmv(sum(ck * bk for ck, bk in zip(c, base)))
        xin = x.cpu().numpy()
        rep = S.solve_device(dev(b), x)
        assert rep["status"] == 0
        hist = handed[::-1][:K]  # most recent first
        rho_h = rho_handed[::-1][:K]
        if hist:
            r0 = b - O.spmv(xin)
            pts = [xin] + hist
            U = [O.spmv(pts[k] - pts[k + 1]) for k in range(len(hist))]
            a2 = r0 @ r0
            nw2 = NOISE_WEIGHT ** 2
            best, g1, cond, depth = a2, 0.0, 1.0, 0
            score_best = a2 + nw2 * rho_last ** 2
            plain_depth, plain_best = 0, a2  # the same without the noise term
            for m in range(1, len(U) + 1):
                Um = np.stack(U[:m])
                G = Um @ Um.T
                g = np.linalg.solve(G, Um @ r0)
                r = r0 - g @ Um
                r2 = r @ r
                cc = np.concatenate([[1 + g[0]], np.diff(g), [-g[-1]]])
                rho = np.array([rho_last] + rho_h[:m])
                score = r2 + nw2 * np.sum(cc ** 2 * rho ** 2)
                if score < 0.99 * score_best:
                    score_best, best, g1, cond, depth = score, r2, g[0], np.linalg.cond(G), m
                if r2 < 0.99 * plain_best:
                    plain_best, plain_depth = r2, m
            deep = max(deep, depth)
            noise_decided += int(depth != plain_depth)
            assert rep["guess_depth"] == depth, (t, rep, depth)
            assert abs(rep["guess_resid"] - np.sqrt(best)) <= 1e-7 * np.sqrt(a2), (t, rep, best)
            # the weights are fixed to about cond(G) times the rounding of G
            bar = max(1e-6, 1e-13 * cond) * max(1.0, abs(g1))
            assert abs(rep["guess_weight_time"] - g1) <= bar, (t, rep, g1, cond)
            assert rep["guess_resid"] <= np.linalg.norm(r0)
        handed.append(xin)
        rho_handed.append(rho_last)
        rho_last = rep["true_rel_resid"] * rep["rhs_norm"]
        x.copy_(dev(mean_free(x.cpu().numpy(), blocks, h0)))
    # the test has teeth: at 1e-10 all K directions were taken; at 1e-3 the
    # noise term changed the choice at least once
    if tol < 1e-8:
        assert deep == K
    else:
        assert noise_decided > 0


@pytest.mark.parametrize("prec", ["fp64", "mixed_fp16"])
@pytest.mark.parametrize("bmean", [1, 3])
def test_history_with_zeroed_origin_cell(gpu, prec, bmean):
    """With bMean 1 or 3 getVec zeroes the origin cell of the guess, and every
    warm-start direction is zero there too: the solve starts from the residual
    the weights minimised (its own err_init equals the reported guess
    residual), never above the given guess's (a handle without the history,
    same inputs)."""
    blocks, bpd, h0, bc = CASES["tiny_amr3_walls"]
    O = pyoracle.Oracle(blocks, bpd, h0, bc, bmean)
    kw = dict(bmean=bmean, tol_rel=0.0, tol_rhs=1e-9, precision=prec, max_iter=3000)
    S = PoissonSolver(bpd, h0, bc, extrapolate=3, **kw).set_topology(blocks)
    P = PoissonSolver(bpd, h0, bc, **kw).set_topology(blocks)
    base = [random_field(len(blocks), 90 + k) for k in range(4)]
    x = torch.zeros(O.n, dtype=torch.float64, device=gpu)
    taken = 0
    for t in range(7):
        c = [np.cos(0.3 * (1 + 0.25 * k) * t + k) for k in range(4)]
        b = dev(O.spmv(sum(ck * bk for ck, bk in zip(c, base))))
        xg = x.clone()
        rep = S.solve_device(b, x)
        rep2 = P.solve_device(b, xg)
        assert rep["status"] == 0 and rep2["status"] == 0, (t, rep, rep2)
        assert abs(rep["err_init"] - rep["guess_resid"]) <= 1e-8 * rep2["err_init"], (t, rep)
        assert rep["err_init"] <= rep2["err_init"] * (1 + 1e-12), (t, rep, rep2)
        taken += rep["guess_depth"] > 0
    assert taken >= 3  # the test has teeth: the history was used
