"""The allocation-free mean removal (cup3d_remove_mean_noalloc: no handle, its
partial sums in the library's static device array) against the handle's
cup3d_poisson_remove_mean, through the C ABI: bitwise equal results and
means on two calls in a row, the second on a smaller grid, with and without
reading the mean back; a block count beyond the array's limit is refused on
the host with a message naming the limit."""
import ctypes

import numpy as np
import pytest
import torch

from cup3d_amd import workloads as W
from cup3d_amd.poisson import PoissonSolver, lib
from helpers import random_field, school4_small

pytestmark = pytest.mark.gpu

MAX_BLOCKS = 1 << 20  # CUP3D_NOALLOC_MAX_BLOCKS


def noalloc(nb, h3, x, read_mean=True):
    L = lib()
    f = L.cup3d_remove_mean_noalloc
    f.argtypes = [ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                  ctypes.POINTER(ctypes.c_double)]
    m = ctypes.c_double(float("nan"))
    torch.cuda.synchronize()
    rc = f(nb, ctypes.c_void_p(h3.data_ptr()), ctypes.c_void_p(x.data_ptr()), None,
           ctypes.byref(m) if read_mean else None)
    torch.cuda.synchronize()
    return rc, m.value


def grid(blocks, bpd, h0, bc, seed):
    S = PoissonSolver(bpd, h0, bc, bmean=0).set_topology(blocks)
    x = torch.from_numpy(random_field(len(blocks), seed).reshape(-1) + 0.3).cuda()
    h3 = torch.from_numpy((h0 / 2.0 ** np.asarray(blocks)[:, 0]) ** 3).cuda()
    return S, x, h3


def test_noalloc_matches_the_handle_entry_point(gpu):
    small = (W.uniform_blocks(3, 2, 2), (3, 2, 2), 1.0 / 24, (2, 2, 2))
    for k, (blocks, bpd, h0, bc) in enumerate([school4_small(), small]):
        S, x, h3 = grid(blocks, bpd, h0, bc, 40 + k)
        y = x.clone()
        z = x.clone()
        m_ref = S.remove_mean(x)
        rc, m = noalloc(len(blocks), h3, y)
        assert rc == 0, lib().cup3d_poisson_last_error(None)
        assert m == m_ref
        assert torch.equal(x, y)
        rc, _ = noalloc(len(blocks), h3, z, read_mean=False)  # enqueue only
        assert rc == 0
        assert torch.equal(x, z)
        assert abs(float((h3.repeat_interleave(512) * y).sum())) <= 1e-12 * float(
            (h3.repeat_interleave(512) * y.abs()).sum())


def test_noalloc_refuses_more_blocks_than_its_array(gpu):
    x = torch.zeros(512, dtype=torch.float64, device=gpu)
    h3 = torch.ones(1, dtype=torch.float64, device=gpu)
    rc, _ = noalloc(MAX_BLOCKS + 1, h3, x)
    assert rc < 0
    assert b"CUP3D_NOALLOC_MAX_BLOCKS" in lib().cup3d_poisson_last_error(None)
    rc, _ = noalloc(0, h3, x)
    assert rc < 0
