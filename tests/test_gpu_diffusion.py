"""Implicit diffusion on the GPU (cup3d_diffusion_*: AdvectionDiffusionImplicit,
DiffusionSolverAMRKernels) against the oracle's restatement, through the C ABI.

Bars (FP64): the operator and the DST block inverse within 1e-13 / 1e-12 of
the field's magnitude (the oracle assembles the interface rows and inverts
the per-level block matrix by Cholesky, the GPU evaluates the stencil
matrix-free and applies sine transforms); the classical solve at the same
iteration count within 1e-10 relative L2 (this system is well conditioned,
so rounding is not amplified); converged solves and whole steps within 1e-9
of the oracle's pipelined reference algorithm at a 1e-12 residual; at the
reference's default tolerances both reach their stopping rule. Full size: one
step on the uniform 256x128x128 walls grid and on the 3-level AMR sphere grid
with the blob velocity."""
import numpy as np
import pyoracle
import pytest
import torch

from cup3d_amd import workloads as W
from cup3d_amd.poisson import PoissonSolver, solver_for
from helpers import FREE, grid_cases, random_field, school4_small, tiny_amr

pytestmark = pytest.mark.gpu

CASES = dict(grid_cases())
CASES["tiny_amr_free"] = tiny_amr()[:3] + ((FREE, FREE, FREE),)
CASES["tiny_amr3_walls"] = tiny_amr(three_levels=True)
CASES["school4_small"] = school4_small()

NU, DT = 2e-3, 0.05


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a, np.float64)).cuda()


def pair(name):
    blocks, bpd, h0, bc = CASES[name]
    S = PoissonSolver(bpd, h0, bc, bmean=0).set_topology(blocks)
    O = pyoracle.Oracle(blocks, bpd, h0, bc, 0)
    return S, O, blocks


@pytest.mark.parametrize("name", list(CASES))
def test_operator_matches_oracle(gpu, name):
    S, O, blocks = pair(name)
    x = random_field(len(blocks), seed=21).reshape(-1)
    y = torch.zeros(len(x), dtype=torch.float64, device=gpu)
    for comp in range(3):
        for nu, dt in ((NU, DT), (1e-5, 1e-3)):
            want = O.diffusion_lhs(x, comp, 1.0 / (nu * dt))
            S.diffusion_apply(comp, nu, dt, dev(x), y)
            got = y.cpu().numpy()
            assert np.abs(got - want).max() <= 1e-13 * np.abs(want).max(), (comp, nu)


@pytest.mark.parametrize("name", ["uniform_walls_3x2x2", "amr3_walls", "school4_small"])
def test_preconditioner_matches_oracle(gpu, name):
    S, O, blocks = pair(name)
    p = random_field(len(blocks), seed=22).reshape(-1)
    z = torch.zeros(len(p), dtype=torch.float64, device=gpu)
    for nu, dt in ((NU, DT), (1.0, 1e3)):
        want = O.diffusion_precond(p, 1.0 / (nu * dt))
        S.diffusion_precond(nu, dt, dev(p), z)
        got = z.cpu().numpy()
        assert np.abs(got - want).max() <= 1e-12 * np.abs(want).max()


@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("iters", [6, 12])
def test_solve_same_iterations_matches_oracle(gpu, name, iters):
    """The classical iteration with the exact block inverse, run for a fixed
    number of iterations on both sides (best iterate of each)."""
    blocks, bpd, h0, bc = CASES[name]
    S = PoissonSolver(bpd, h0, bc, bmean=0, max_iter=iters).set_topology(blocks)
    O = pyoracle.Oracle(blocks, bpd, h0, bc, 0)
    b = random_field(len(blocks), seed=23).reshape(-1)
    comp = 1
    xo, info = O.diffusion_bicgstab(b, comp, 1.0 / (NU * DT), tol_rel=0.0, max_iter=iters,
                                    max_restarts=1 << 30)
    x = torch.zeros(len(b), dtype=torch.float64, device=gpu)
    rep = S.diffusion_solve(comp, NU, DT, dev(b), x, tol_abs=0.0, tol_rel=0.0)
    assert rep["iterations"] == info["iters"] == iters
    got = x.cpu().numpy()
    assert np.linalg.norm(got - xo) <= 1e-10 * np.linalg.norm(xo)


@pytest.mark.parametrize("name", ["uniform_mixed_bc", "amr2_periodic_offcentre", "tiny_amr_free",
                                  "school4_small"])
def test_converged_solve_matches_reference_algorithm(gpu, name):
    S, O, blocks = pair(name)
    b = random_field(len(blocks), seed=24).reshape(-1)
    sigma = 1.0 / (NU * DT)
    for comp in range(3):
        xo, info = O.diffusion_pipelined(b, comp, sigma, tol_rel=1e-12, max_iter=1000)
        assert info["status"] == 0
        x = torch.zeros(len(b), dtype=torch.float64, device=gpu)
        rep = S.diffusion_solve(comp, NU, DT, dev(b), x, tol_abs=0.0, tol_rel=1e-12)
        assert rep["status"] == 0 and rep["true_rel_resid"] <= 2e-12, rep
        got = x.cpu().numpy()
        assert np.linalg.norm(got - xo) <= 1e-9 * np.linalg.norm(xo)


def step_inputs(blocks, seed):
    n = len(blocks) * 512
    rng = np.random.default_rng(seed)
    u_old = rng.standard_normal((n, 3))
    u_star = u_old + 0.05 * rng.standard_normal((n, 3))
    return u_old, u_star


@pytest.mark.parametrize("name", ["uniform_periodic_2x3x2", "uniform_mixed_bc", "amr3_walls",
                                  "tiny_amr_free", "school4_small"])
def test_step_matches_oracle(gpu, name):
    S, O, blocks = pair(name)
    u_old, u_star = step_inputs(blocks, 25)
    want, _ = O.diffusion_step(u_old, u_star, NU, DT, tol_abs=0.0, tol_rel=1e-12)
    v = dev(u_star)
    reps = S.diffusion_step(dev(u_old), v, NU, DT, tol_abs=0.0, tol_rel=1e-12)
    assert all(r["status"] == 0 for r in reps), reps
    got = v.cpu().numpy()
    assert np.linalg.norm(got - want) <= 1e-9 * np.linalg.norm(want)


@pytest.mark.parametrize("name", ["uniform_walls_3x2x2", "amr2_walls"])
def test_step_at_reference_tolerances(gpu, name):
    """sim.DiffusionErrorTol = 1e-6 and DiffusionErrorTolRel = 1e-4
    (SimulationData.cpp:92-93): both sides stop by that rule, and both land
    within the distance it allows from the tight solution."""
    S, O, blocks = pair(name)
    u_old, u_star = step_inputs(blocks, 26)
    tight, _ = O.diffusion_step(u_old, u_star, NU, DT, tol_abs=0.0, tol_rel=1e-12)
    ref, info = O.diffusion_step(u_old, u_star, NU, DT)
    v = dev(u_star)
    reps = S.diffusion_step(dev(u_old), v, NU, DT)
    for r in reps:
        assert r["status"] == 0
        assert r["err_final"] < 1e-6 or r["err_final"] / r["err_init"] < 1e-4, r
    inc = np.linalg.norm(tight - u_star)
    e_gpu = np.linalg.norm(v.cpu().numpy() - tight) / inc
    e_ref = np.linalg.norm(ref - tight) / inc
    print(f"{name}: increment error GPU {e_gpu:.2e}, reference algorithm {e_ref:.2e}, "
          f"iterations GPU {[r['iterations'] for r in reps]} reference {info['iters']}")
    assert e_gpu <= 1e-3 and e_ref <= 1e-3


def blob_step(w, nu, dt):
    vel = W.vel_blobs(w.blocks, w.h0, w.bpd, w.bc, 256, (0.02, 0.08), sphere=W.SPHERE)
    u_old = torch.from_numpy(vel).cuda()
    u_star = torch.from_numpy(W.vel_blobs(w.blocks, w.h0, w.bpd, w.bc, 256, (0.02, 0.08),
                                          sphere=W.SPHERE, step=1)).cuda()
    S = solver_for(w, bmean=0)
    v = u_star.clone()
    reps = S.diffusion_step(u_old, v, nu, dt, tol_abs=0.0, tol_rel=1e-10)
    return S, u_old, u_star, v, reps


@pytest.mark.parametrize("wl", ["c2", "c3"])
@pytest.mark.parametrize("nu", [1e-3, 1e-4])
def test_full_size_step_solves_backward_euler(gpu, wl, nu):
    """One step at full size: for every component the result satisfies
    A_lap u^{n+1} = h^3 (u^{n+1} - u^n) / (nu dt) to the solver's residual,
    checked with the GPU operator (pinned against the oracle above)."""
    w = W.c2_uniform_walls() if wl == "c2" else W.c3_amr_sphere()
    dt = 0.1
    S, u_old, u_star, v, reps = blob_step(w, nu, dt)
    print(f"{w.name} nu {nu}: iterations {[r['iterations'] for r in reps]}, "
          f"ms {[round(r['solve_ms'], 2) for r in reps]}")
    h3 = torch.from_numpy(np.repeat((w.h0 / 2.0 ** w.blocks[:, 0]) ** 3, 512)).cuda()
    s = h3 / (nu * dt)
    lap = torch.zeros_like(h3)
    for c in range(3):
        assert reps[c]["status"] == 0 and reps[c]["true_rel_resid"] <= 2e-10, reps[c]
        # nu dt = 1e300: the shift h^3 / (nu dt) vanishes, leaving h lap
        S.diffusion_apply(c, 1.0, 1e300, v[:, c].contiguous(), lap)
        res = lap - s * (v[:, c] - u_old[:, c])
        S.diffusion_apply(c, 1.0, 1e300, u_star[:, c].contiguous(), lap)
        b = -lap + s * (u_star[:, c] - u_old[:, c])
        assert float(torch.linalg.norm(res)) <= 1e-9 * float(torch.linalg.norm(b))
