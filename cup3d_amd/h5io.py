"""Block fields on disk in CubismUP_3D's multi-block HDF5 layout (ctypes
binding of libcup3d_h5.so, include/cup3d_h5.h): write a field of the
solver's block layout for the reference's tools, and read a dump
(Simulation.cpp:245-258, read by tools/uniform_conversion/main.cpp:59-150)
into the solver's layout, e.g. a dumped pressure as the guess and rhs of a
solve."""
from __future__ import annotations

import ctypes

import numpy as np

from ._lib import load_library

_i32p = ctypes.POINTER(ctypes.c_int32)
_f64p = ctypes.POINTER(ctypes.c_double)


def _lib():
    L = load_library("libcup3d_h5.so")
    if not getattr(L, "_typed", False):
        L.cup3d_h5_write_blocks.argtypes = [ctypes.c_char_p, ctypes.c_int64, _i32p,
                                            ctypes.c_double, _f64p, ctypes.c_int, ctypes.c_int,
                                            ctypes.c_char_p, ctypes.c_int]
        L.cup3d_h5_count_blocks.restype = ctypes.c_int64
        L.cup3d_h5_count_blocks.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int]
        L.cup3d_h5_read_blocks.restype = ctypes.c_int64
        L.cup3d_h5_read_blocks.argtypes = [ctypes.c_char_p, ctypes.c_int64, _i32p, _f64p, _f64p,
                                           ctypes.c_char_p, ctypes.c_int]
        L._typed = True
    return L


def write_blocks(base, blocks, h0, values, float32=True, group_x=1):
    """Write <base>.h5 and <base>-groups.h5 from a block list (nb x 4:
    level, index[3]) and values (nb x 512, [iz][iy][ix])."""
    blocks = np.ascontiguousarray(blocks, np.int32).reshape(-1, 4)
    values = np.ascontiguousarray(values, np.float64).reshape(len(blocks), 512)
    err = ctypes.create_string_buffer(512)
    rc = _lib().cup3d_h5_write_blocks(str(base).encode(), len(blocks),
                                      blocks.ctypes.data_as(_i32p), float(h0),
                                      values.ctypes.data_as(_f64p), int(bool(float32)),
                                      int(group_x), err, len(err))
    if rc != 0:
        raise RuntimeError(f"cup3d_h5_write_blocks: {err.value.decode()}")


def read_blocks(base):
    """Returns (blocks [nb, 4] int32, values [nb, 512] float64, h0)."""
    L = _lib()
    err = ctypes.create_string_buffer(512)
    nb = L.cup3d_h5_count_blocks(str(base).encode(), err, len(err))
    if nb < 0:
        raise RuntimeError(f"cup3d_h5_count_blocks: {err.value.decode()}")
    blocks = np.zeros((nb, 4), np.int32)
    values = np.zeros((nb, 512), np.float64)
    h0 = ctypes.c_double()
    got = L.cup3d_h5_read_blocks(str(base).encode(), nb, blocks.ctypes.data_as(_i32p),
                                 values.ctypes.data_as(_f64p), ctypes.byref(h0), err, len(err))
    if got != nb:
        raise RuntimeError(f"cup3d_h5_read_blocks: {err.value.decode()}")
    return blocks, values, h0.value
