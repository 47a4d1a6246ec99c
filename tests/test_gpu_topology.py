"""The device topology build (set_topology) against the library's host build,
table for table and bit for bit, on every test grid and at the full 3-level
sphere and 4-level school sizes; the same first-offender error messages for
rejected block lists, which leave the previous topology in place; a mesh
change that reuses the handle's buffers."""
import ctypes

import numpy as np
import pytest
import torch

from cup3d_amd import poisson as P
from cup3d_amd import workloads as W
from cup3d_amd.poisson import PoissonSolver
from test_topology_build import BAD, CASES, host_topology

pytestmark = pytest.mark.gpu


def device_topology(S):
    L = P.lib()
    f = L.cup3d_debug_device_topology
    f.argtypes = [ctypes.c_void_p] + [ctypes.c_void_p] * 8
    nb = S.nb
    out = dict(nbr=np.zeros((nb, 6), np.int32), kind=np.zeros((nb, 6), np.int8),
               fnb=np.zeros((nb, 24), np.int32), bpar=np.zeros(nb, np.int8),
               h=np.zeros(nb), invh=np.zeros(nb), h3=np.zeros(nb), info=np.zeros(3, np.int64))
    assert f(S.h, *[out[k].ctypes.data for k in ("nbr", "kind", "fnb", "bpar", "h", "invh",
                                                   "h3", "info")]) == 0
    return out


def full_cases():
    c = dict(CASES)
    for name, fn in (("c3_amr_sphere", W.c3_amr_sphere), ("c5_amr_school", W.c5_amr_school),
                     ("c2_uniform_walls", W.c2_uniform_walls)):
        w = fn()
        c[name] = (w.blocks, w.bpd, w.h0, w.bc)
    return c


@pytest.mark.parametrize("name", list(full_cases()))
def test_device_build_equals_host_build(gpu, name):
    blocks, bpd, h0, bc = full_cases()[name]
    S = PoissonSolver(bpd, h0, bc, bmean=1).set_topology(blocks)
    D = device_topology(S)
    T, err = host_topology(blocks, bpd, h0, bc)
    assert T is not None, err
    for k in ("nbr", "kind", "fnb", "bpar", "h", "invh", "h3", "info"):
        assert np.array_equal(D[k], T[k]), k


@pytest.mark.parametrize("blocks,message", BAD)
def test_device_build_rejects_like_host(gpu, blocks, message):
    good = np.array([[0, 0, 0, 0], [0, 1, 0, 0]], np.int32)
    S = PoissonSolver((2, 1, 1), 1 / 16, (2, 2, 2), bmean=0).set_topology(good)
    with pytest.raises(RuntimeError) as e:
        S.set_topology(np.array(blocks, np.int32))
    assert message in str(e.value), str(e.value)
    # the previous topology is still in place and solves
    assert S.nb == 2
    D = device_topology(S)
    assert D["nbr"][0, 1] == 1 and D["nbr"][1, 0] == 0
    n = 2 * 512
    b = torch.from_numpy(np.random.default_rng(1).standard_normal(n)).cuda()
    b -= b.mean()
    x = torch.zeros(n, dtype=torch.float64, device=gpu)
    assert S.solve_device(b, x)["status"] == 0


def test_mesh_change_sequence(gpu):
    """One handle taken through a sequence of meshes on the same base grid,
    growing and shrinking (its buffers grow only), solves each exactly like
    a fresh handle, bit for bit."""
    bpd, h0, bc = (4, 2, 2), 1.0 / 16, (W.BC_WALL,) * 3
    grids = [W.amr_school_blocks(bpd, h0, lmax, (0.18, 0.09, 0.045)[:lmax], 6, W.SCHOOL_AXES,
                                 W.SEED) if lmax else W.uniform_blocks(*bpd)
             for lmax in (1, 3, 0, 2, 3)]
    assert len(set(len(g) for g in grids)) == 4
    S = PoissonSolver(bpd, h0, bc, bmean=0, tol_rel=0.0, max_iter=12)
    for blocks in grids:
        S.set_topology(blocks)
        S1 = PoissonSolver(bpd, h0, bc, bmean=0, tol_rel=0.0, max_iter=12).set_topology(blocks)
        n = len(blocks) * 512
        h3 = np.repeat((h0 / 2.0 ** np.asarray(blocks)[:, 0]) ** 3, 512)
        r = np.random.default_rng(3).standard_normal(n)
        b = torch.from_numpy(r - h3 * r.sum() / h3.sum()).cuda()
        x, x1 = (torch.zeros(n, dtype=torch.float64, device=gpu) for _ in range(2))
        r0, r1 = S.solve_device(b, x), S1.solve_device(b, x1)
        assert r0["iterations"] == r1["iterations"] == 12
        assert torch.equal(x, x1), len(blocks)
