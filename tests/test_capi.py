"""C ABI and host logic of libcup3d_poisson, without a GPU: the library loads,
exports every function the public headers declare, rejects bad input, and its
host-side topology build yields the same operator as the oracle's assembly of
PoissonSolverExp::getMat (two independent implementations)."""
import ctypes
import glob
import os
import subprocess
import re

import numpy as np
import pyoracle
import pytest

from cup3d_amd import poisson as P
from cup3d_amd._lib import load_library
from helpers import school4_small, grid_cases, tiny_amr

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions(headers=None):
    names = set()
    for h in headers or glob.glob(os.path.join(ROOT, "include", "*.h")):
        src = open(h).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        for m in re.finditer(r"^\s*[A-Za-z_][\w\s\*]*?\b(cup3d_\w+)\s*\(", src, re.M):
            names.add(m.group(1))
    return names


def test_library_exports_every_declared_symbol():
    """include/cup3d_h5.h is libcup3d_h5.so's; every other header is
    libcup3d_poisson.so's."""
    inc = os.path.join(ROOT, "include")
    h5 = os.path.join(inc, "cup3d_h5.h")
    L = P.lib()
    names = header_functions([h for h in glob.glob(os.path.join(inc, "*.h")) if h != h5])
    assert len(names) >= 16
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    assert set(P.EXPORTS) <= names
    L5 = load_library("libcup3d_h5.so")
    names5 = header_functions([h5])
    assert len(names5) == 3
    missing = [n for n in names5 if not hasattr(L5, n)]
    assert not missing, missing


def test_create_rejects_bad_config():
    L = P.lib()
    c = P.Config()
    h = ctypes.c_void_p()
    assert L.cup3d_poisson_create(ctypes.byref(c), ctypes.byref(h)) == -1
    assert b"bpd" in L.cup3d_poisson_last_error(None)
    for d in range(3):
        c.bpd[d] = 2
        c.bc[d] = 7
    c.h0 = 0.1
    c.max_iter = 10
    assert L.cup3d_poisson_create(ctypes.byref(c), ctypes.byref(h)) == -1
    assert b"bc" in L.cup3d_poisson_last_error(None)


@pytest.mark.parametrize("field,value,message", [
    ("bmean", 4, b"bmean"), ("max_iter", 0, b"max_iter"), ("precision", 3, b"precision"),
    ("guess", 3, b"guess"), ("extrapolate", 9, b"extrapolate"), ("h0", 0.0, b"h0"),
    ("precond", 2, b"precond")])
def test_create_rejects_each_bad_field(field, value, message):
    """Every configuration field is checked before any device is touched."""
    L = P.lib()
    c = P.Config()
    for d in range(3):
        c.bpd[d] = 2
        c.bc[d] = 2
    c.h0, c.max_iter = 0.1, 10
    setattr(c, field, value)
    h = ctypes.c_void_p()
    assert L.cup3d_poisson_create(ctypes.byref(c), ctypes.byref(h)) == -1
    assert message in L.cup3d_poisson_last_error(None)


def test_null_handle_calls_fail_cleanly():
    L = P.lib()
    assert L.cup3d_poisson_solve_device(None, None, None, None) < 0
    assert L.cup3d_poisson_remove_mean(None, None, None) < 0
    assert L.cup3d_poisson_num_coarse_fine_faces(None) < 0
    L.cup3d_poisson_destroy(None)


def debug_csr(blocks, bpd, h0, bc, bmean):
    L = load_library("libcup3d_poisson.so")
    f = L.cup3d_debug_operator_csr
    f.restype = ctypes.c_int64
    f.argtypes = [ctypes.POINTER(P.Config), ctypes.c_int64, ctypes.POINTER(P.BlockDesc),
                  ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64]
    c = P.Config()
    for d in range(3):
        c.bpd[d], c.bc[d] = bpd[d], bc[d]
    c.h0, c.bmean = h0, bmean
    arr = (P.BlockDesc * len(blocks))()
    for i, (l, a, b, cc) in enumerate(blocks):
        arr[i].level = int(l)
        arr[i].index[0], arr[i].index[1], arr[i].index[2] = int(a), int(b), int(cc)
    nnz = f(ctypes.byref(c), len(blocks), arr, None, None, None, 0)
    assert nnz > 0
    n = len(blocks) * 512
    rowptr = np.zeros(n + 1, np.int64)
    col = np.zeros(nnz, np.int32)
    val = np.zeros(nnz, np.float64)
    assert f(ctypes.byref(c), len(blocks), arr, rowptr.ctypes.data, col.ctypes.data,
             val.ctypes.data, nnz) == nnz
    import scipy.sparse as sp
    return sp.csr_matrix((val, col, rowptr), shape=(n, n))


CASES = dict(grid_cases())
CASES["tiny_amr3_walls"] = tiny_amr(three_levels=True)
CASES["school4_small"] = school4_small()
CASES["tiny_amr_periodic"] = tiny_amr(periodic=True)


@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("bmean", [0, 3])
def test_topology_operator_matches_oracle_assembly(name, bmean):
    blocks, bpd, h0, bc = CASES[name]
    A = debug_csr(blocks, bpd, h0, bc, bmean)
    O = pyoracle.Oracle(blocks, bpd, h0, bc, bmean)
    B = O.scipy_matrix()
    D = (A - B).tocsr()
    D.eliminate_zeros()
    scale = np.abs(B.data).max()
    assert (abs(D).max() if D.nnz else 0.0) <= 1e-14 * scale


@pytest.mark.parametrize("blocks", [
    [[0, 0, 0, 0], [0, 0, 0, 0]],   # duplicate leaf
    [[0, 2, 0, 0], [0, 1, 0, 0]],   # index outside the 2x1x1 domain
    [[-1, 0, 0, 0], [0, 1, 0, 0]],  # invalid level
])
def test_topology_rejects_bad_block_lists(blocks):
    with pytest.raises(AssertionError):
        debug_csr(np.array(blocks, np.int32), (2, 1, 1), 1 / 16, (2, 2, 2), 0)


def test_topology_rejects_unbalanced_grid():
    # a level-2 block next to level-0 blocks violates 2:1 balance
    blocks = np.array([[2, 0, 0, 0]] + [[0, 1, 0, 0]], np.int32)
    with pytest.raises(AssertionError):
        debug_csr(blocks, (2, 1, 1), 1 / 16, (2, 2, 2), 0)


def test_adapter_binds_into_reference_factory(tmp_path):
    """The header-only adapter compiles inside a translation unit that has
    declared the reference's PoissonSolverBase, SimulationData, BCflag and
    BlockInfo first (shaped like source/poisson/PoissonSolverBase.cpp), and
    the result links against libcup3d_poisson.so with no undefined symbol."""
    out = tmp_path / "ref_tu.so"
    r = subprocess.run(
        ["g++", "-std=c++17", "-Wall", "-Wextra", "-Werror", "-fPIC", "-shared",
         "-Wl,--no-undefined", "-I", os.path.join(ROOT, "include"),
         "-I", os.path.join(ROOT, "cup3d_amd", "host"),
         os.path.join(ROOT, "tests", "adapter_tu", "reference_factory.cpp"), "-o", str(out),
         "-L", os.path.join(ROOT, "cup3d_amd", "lib"), "-lcup3d_poisson",
         "-Wl,-rpath," + os.path.join(ROOT, "cup3d_amd", "lib")],
        capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    L = ctypes.CDLL(str(out))
    assert L.reference_style_projection is not None
    exe = os.path.join(ROOT, "cup3d_amd", "lib", "example_projection")
    assert os.access(exe, os.X_OK)
