"""The host topology build of libcup3d_poisson (the one the explicit-matrix
export uses) against an independent dictionary walk of the block list, on
the CPU: same-level neighbours with the periodic wrap, coarser and finer
partners, index parities, h, and the face counts. The device build is
checked against this host build, table for table, in
test_gpu_topology.py."""
import ctypes

import numpy as np
import pytest

from cup3d_amd import poisson as P
from cup3d_amd import workloads as W
from helpers import face_tables, grid_cases, school4_small, tiny_amr

CASES = dict(grid_cases())
CASES["tiny_amr3_walls"] = tiny_amr(three_levels=True)
CASES["tiny_amr_periodic"] = tiny_amr(periodic=True)
CASES["school4_small"] = school4_small()


def config(bpd, h0, bc, bmean=0):
    c = P.Config()
    for d in range(3):
        c.bpd[d], c.bc[d] = bpd[d], bc[d]
    c.h0, c.bmean, c.max_iter = h0, bmean, 10
    return c


def host_topology(blocks, bpd, h0, bc):
    """Tables of cup3d_debug_host_topology, or the error message."""
    L = P.lib()
    f = L.cup3d_debug_host_topology
    f.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p] + [ctypes.c_void_p] * 8 + \
        [ctypes.c_char_p, ctypes.c_int]
    blocks = np.ascontiguousarray(blocks, np.int32)
    nb = len(blocks)
    out = dict(nbr=np.zeros((nb, 6), np.int32), kind=np.zeros((nb, 6), np.int8),
               fnb=np.zeros((nb, 24), np.int32), bpar=np.zeros(nb, np.int8),
               h=np.zeros(nb), invh=np.zeros(nb), h3=np.zeros(nb), info=np.zeros(3, np.int64))
    err = ctypes.create_string_buffer(256)
    c = config(bpd, h0, bc)
    rc = f(ctypes.byref(c), nb, blocks.ctypes.data,
           *[out[k].ctypes.data for k in ("nbr", "kind", "fnb", "bpar", "h", "invh", "h3",
                                          "info")], err, 256)
    return (out if rc == 0 else None), err.value.decode()


@pytest.mark.parametrize("name", list(CASES))
def test_host_build_matches_dictionary_walk(name):
    blocks, bpd, h0, bc = CASES[name]
    T, err = host_topology(blocks, bpd, h0, bc)
    assert T is not None, err
    kind, nbr, fnb, bpar = face_tables(blocks, bpd, bc)
    assert np.array_equal(T["kind"], kind)
    assert np.array_equal(T["nbr"], nbr)
    assert np.array_equal(T["fnb"], fnb)
    assert np.array_equal(T["bpar"], bpar)
    h = h0 / 2.0 ** np.asarray(blocks)[:, 0]
    assert np.array_equal(T["h"], h) and np.array_equal(T["h3"], h * h * h)
    assert np.array_equal(T["invh"], 1.0 / h)
    assert T["info"][1] == np.sum(kind == 2) and T["info"][2] == np.sum(kind == 3)
    origin = [i for i, b in enumerate(np.asarray(blocks).tolist()) if b[1:] == [0, 0, 0]]
    assert T["info"][0] == (origin[-1] if origin else -1)


def test_host_build_c5_counts():
    """The 4-level school grid: 59 012 blocks, every face classified."""
    w = W.c5_amr_school()
    T, err = host_topology(w.blocks, w.bpd, w.h0, w.bc)
    assert T is not None, err
    assert len(w.blocks) == 59012
    kind, nbr, fnb, bpar = face_tables(w.blocks, w.bpd, w.bc)
    assert np.array_equal(T["kind"], kind) and np.array_equal(T["fnb"], fnb)
    assert np.array_equal(T["nbr"], nbr)


BAD = [
    ([[0, 0, 0, 0], [0, 1, 0, 0], [0, 0, 0, 0]], "duplicate block 2"),
    ([[0, 0, 0, 0], [0, 2, 0, 0], [0, 1, 0, 0]], "block 1: index outside the domain"),
    ([[0, 1, 0, 0], [-1, 0, 0, 0], [0, 0, 0, 0]], "block 1: invalid level/index"),
    ([[0, 1, 0, 0], [16, 0, 0, 0]], "block 1: invalid level/index"),
    ([[2, 0, 0, 0], [0, 1, 0, 0]], "block 0 face 1: neighbour is neither same-level"),
    # the first offending block decides, whatever the kind of its fault
    ([[0, 1, 0, 0], [0, 1, 0, 0], [0, 9, 0, 0]], "duplicate block 1"),
    ([[0, 1, 0, 0], [0, 9, 0, 0], [0, 1, 0, 0]], "block 1: index outside the domain"),
]


@pytest.mark.parametrize("blocks,message", BAD)
def test_host_build_rejects(blocks, message):
    T, err = host_topology(np.array(blocks, np.int32), (2, 1, 1), 1 / 16, (2, 2, 2))
    assert T is None and err.startswith(message), err


def test_host_build_rejects_too_fine_levels():
    # 2^17 base blocks at level 4 exceed the 2^20 index width of the key
    T, err = host_topology(np.array([[4, 0, 0, 0]], np.int32), (1 << 17, 1, 1), 1.0, (2, 2, 2))
    assert T is None and "too fine" in err, err
