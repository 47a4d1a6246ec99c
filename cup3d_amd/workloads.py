"""Synthetic Poisson workloads (block lists and right-hand sides).

The five workloads below are the solver inputs of a CubismUP_3D pressure
projection (rhs in ``sim.lhs``, guess in ``sim.pres``) built without the fluid
solver. The arithmetic lives in ``csrc/workloads.cpp`` (C++/OpenMP); this
module only binds it with ctypes and names the workloads.

Block arrays are ``int32[nblocks, 4]`` rows ``(level, ix, iy, iz)`` in the
order the solver numbers blocks (blockID); fields are ``float64[nblocks, 8, 8,
8]`` laid out ``[iz][iy][ix]`` like Cubism's ``ScalarBlock``
(source/Definitions.h:524-531).
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass, field

import numpy as np

from ._lib import load_library

BC_FREESPACE, BC_PERIODIC, BC_WALL = 0, 1, 2  # BCflag, source/Definitions.h:38

SEED = 20261202
FRESH_SEED = 20261203

_i32p = ctypes.POINTER(ctypes.c_int)
_f64p = ctypes.POINTER(ctypes.c_double)


def _lib():
    lib = load_library("libcup3d_workloads.so")
    if not getattr(lib, "_typed", False):
        lib.cw_uniform_blocks.restype = ctypes.c_int
        lib.cw_uniform_blocks.argtypes = [ctypes.c_int] * 3 + [_i32p, ctypes.c_int]
        lib.cw_amr_sphere_blocks.restype = ctypes.c_int
        lib.cw_amr_sphere_blocks.argtypes = [_i32p, ctypes.c_double, ctypes.c_int, _f64p, _f64p,
                                             ctypes.c_double, _i32p, ctypes.c_int]
        lib.cw_amr_school_blocks.restype = ctypes.c_int
        lib.cw_amr_school_blocks.argtypes = [_i32p, ctypes.c_double, ctypes.c_int, _f64p,
                                             ctypes.c_int, _f64p, ctypes.c_uint64, _i32p,
                                             ctypes.c_int]
        lib.cw_rhs_fourier.restype = None
        lib.cw_rhs_fourier.argtypes = [ctypes.c_int, _i32p, ctypes.c_double, ctypes.c_int,
                                       ctypes.c_int, ctypes.c_int, ctypes.c_uint64, _f64p]
        lib.cw_rhs_blobs.restype = None
        lib.cw_rhs_blobs.argtypes = [ctypes.c_int, _i32p, ctypes.c_double, _i32p, _i32p,
                                     ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                     ctypes.c_uint64, ctypes.c_int, ctypes.c_double, _f64p,
                                     ctypes.c_double, ctypes.c_int, _f64p, ctypes.c_uint64,
                                     _f64p]
        lib.cw_vel_blobs.restype = None
        lib.cw_vel_blobs.argtypes = lib.cw_rhs_blobs.argtypes
        lib.cw_blob_params.restype = None
        lib.cw_blob_params.argtypes = [_f64p, _i32p, ctypes.c_int, ctypes.c_double,
                                       ctypes.c_double, ctypes.c_uint64, _f64p]
        lib.cw_chi.restype = None
        lib.cw_chi.argtypes = [ctypes.c_int, _i32p, ctypes.c_double, _i32p, _f64p,
                               ctypes.c_double, ctypes.c_int, _f64p, ctypes.c_uint64, _f64p]
        lib._typed = True
    return lib


def _p(a, t):
    return a.ctypes.data_as(ctypes.POINTER(t))


def uniform_blocks(bx: int, by: int, bz: int) -> np.ndarray:
    lib = _lib()
    n = lib.cw_uniform_blocks(bx, by, bz, None, 0)
    out = np.zeros((n, 4), np.int32)
    lib.cw_uniform_blocks(bx, by, bz, _p(out, ctypes.c_int), n)
    return out


def amr_sphere_blocks(bpd, h0, lmax, thresholds, centre=(0.5, 0.5, 0.5), radius=0.1):
    lib = _lib()
    bpd = np.ascontiguousarray(bpd, np.int32)
    thr = np.ascontiguousarray(thresholds, np.float64)
    c = np.ascontiguousarray(centre, np.float64)
    n = lib.cw_amr_sphere_blocks(_p(bpd, ctypes.c_int), h0, lmax, _p(thr, ctypes.c_double),
                                 _p(c, ctypes.c_double), radius, None, 0)
    out = np.zeros((n, 4), np.int32)
    lib.cw_amr_sphere_blocks(_p(bpd, ctypes.c_int), h0, lmax, _p(thr, ctypes.c_double),
                             _p(c, ctypes.c_double), radius, _p(out, ctypes.c_int), n)
    return out


def amr_school_blocks(bpd, h0, lmax, thresholds, n_ell=50, axes=(0.05, 0.02, 0.015),
                      seed=SEED):
    lib = _lib()
    bpd = np.ascontiguousarray(bpd, np.int32)
    thr = np.ascontiguousarray(thresholds, np.float64)
    ax = np.ascontiguousarray(axes, np.float64)
    args = (_p(bpd, ctypes.c_int), h0, lmax, _p(thr, ctypes.c_double), n_ell,
            _p(ax, ctypes.c_double), seed)
    n = lib.cw_amr_school_blocks(*args, None, 0)
    out = np.zeros((n, 4), np.int32)
    lib.cw_amr_school_blocks(*args, _p(out, ctypes.c_int), n)
    return out


def rhs_fourier(blocks, h0, nmodes=16, kmin=1, kmax=4, seed=SEED):
    blocks = np.ascontiguousarray(blocks, np.int32)
    out = np.zeros((len(blocks), 8, 8, 8), np.float64)
    _lib().cw_rhs_fourier(len(blocks), _p(blocks, ctypes.c_int), h0, nmodes, kmin, kmax, seed,
                          _p(out, ctypes.c_double))
    return out


def rhs_blobs(blocks, h0, bpd, bc, nblobs, sigma, seed=SEED, step=0, speed=0.005,
              sphere=None, school=None):
    """rhs = h^2/2 * div((1-chi) u_blobs); sphere=(centre, r); school=(n, axes, seed)."""
    blocks = np.ascontiguousarray(blocks, np.int32)
    bpd = np.ascontiguousarray(bpd, np.int32)
    bc = np.ascontiguousarray(bc, np.int32)
    smin, smax = (sigma, sigma) if np.isscalar(sigma) else sigma
    sc = np.ascontiguousarray(sphere[0] if sphere else (0, 0, 0), np.float64)
    sr = float(sphere[1]) if sphere else 0.0
    n_ell, axes, eseed = school if school else (0, (0.0, 0.0, 0.0), 0)
    ax = np.ascontiguousarray(axes, np.float64)
    out = np.zeros((len(blocks), 8, 8, 8), np.float64)
    _lib().cw_rhs_blobs(len(blocks), _p(blocks, ctypes.c_int), h0, _p(bpd, ctypes.c_int),
                        _p(bc, ctypes.c_int), nblobs, smin, smax, seed, step, speed,
                        _p(sc, ctypes.c_double), sr, n_ell, _p(ax, ctypes.c_double), eseed,
                        _p(out, ctypes.c_double))
    return out


def vel_blobs(blocks, h0, bpd, bc, nblobs, sigma, seed=SEED, step=0, speed=0.005,
              sphere=None, school=None):
    """The penalised blob velocity (1 - chi) u_blobs at the cell centres,
    [nblocks * 512, 3] (VectorBlock layout), the field whose divergence
    rhs_blobs takes."""
    blocks = np.ascontiguousarray(blocks, np.int32)
    bpd = np.ascontiguousarray(bpd, np.int32)
    bc = np.ascontiguousarray(bc, np.int32)
    smin, smax = (sigma, sigma) if np.isscalar(sigma) else sigma
    sc = np.ascontiguousarray(sphere[0] if sphere else (0, 0, 0), np.float64)
    sr = float(sphere[1]) if sphere else 0.0
    n_ell, axes, eseed = school if school else (0, (0.0, 0.0, 0.0), 0)
    ax = np.ascontiguousarray(axes, np.float64)
    out = np.zeros((len(blocks) * 512, 3), np.float64)
    _lib().cw_vel_blobs(len(blocks), _p(blocks, ctypes.c_int), h0, _p(bpd, ctypes.c_int),
                        _p(bc, ctypes.c_int), nblobs, smin, smax, seed, step, speed,
                        _p(sc, ctypes.c_double), sr, n_ell, _p(ax, ctypes.c_double), eseed,
                        _p(out, ctypes.c_double))
    return out


def blob_params(bpd, h0, bc, nblobs, sigma, seed=SEED):
    """The blobs rhs_blobs draws: (nblobs, 10) rows of centre[3] at step 0,
    unit direction[3], sigma, amplitude[3]."""
    ext = np.ascontiguousarray([8 * bpd[a] * h0 for a in range(3)], np.float64)
    bc = np.ascontiguousarray(bc, np.int32)
    smin, smax = (sigma, sigma) if np.isscalar(sigma) else sigma
    out = np.zeros((nblobs, 10), np.float64)
    _lib().cw_blob_params(_p(ext, ctypes.c_double), _p(bc, ctypes.c_int), nblobs, smin, smax,
                          seed, _p(out, ctypes.c_double))
    return out


def chi(blocks, h0, bpd, sphere=None, school=None):
    blocks = np.ascontiguousarray(blocks, np.int32)
    bpd = np.ascontiguousarray(bpd, np.int32)
    sc = np.ascontiguousarray(sphere[0] if sphere else (0, 0, 0), np.float64)
    sr = float(sphere[1]) if sphere else 0.0
    n_ell, axes, eseed = school if school else (0, (0.0, 0.0, 0.0), 0)
    ax = np.ascontiguousarray(axes, np.float64)
    out = np.zeros((len(blocks), 8, 8, 8), np.float64)
    _lib().cw_chi(len(blocks), _p(blocks, ctypes.c_int), h0, _p(bpd, ctypes.c_int),
                  _p(sc, ctypes.c_double), sr, n_ell, _p(ax, ctypes.c_double), eseed,
                  _p(out, ctypes.c_double))
    return out


@dataclass
class Workload:
    """A named solver input: grid, BCs and an rhs generator for step t."""

    name: str
    bpd: tuple
    h0: float
    bc: tuple
    level_max: int
    blocks: np.ndarray
    rhs_fn: object = field(repr=False)

    @property
    def ncells(self) -> int:
        return int(len(self.blocks)) * 512

    def rhs(self, step: int = 0, seed: int = SEED) -> np.ndarray:
        return self.rhs_fn(step, seed)


SPHERE = ((0.5, 0.5, 0.5), 0.1)
SCHOOL_AXES = (0.05, 0.02, 0.015)


def c1_periodic_32() -> Workload:
    """4x4x4 blocks (32^3 cells), unit cube, periodic, 16 Fourier modes."""
    bpd, h0, bc = (4, 4, 4), 1.0 / 32, (BC_PERIODIC,) * 3
    blocks = uniform_blocks(*bpd)
    return Workload("uniform_32^3_periodic_fourier16", bpd, h0, bc, 0, blocks,
                    lambda step, seed: rhs_fourier(blocks, h0, 16, 1, 4, seed + step))


def c2_uniform_walls(bpd=(32, 16, 16)) -> Workload:
    """32x16x16 blocks (256x128x128 cells), [0,2]x[0,1]^2, walls, 256 blobs + sphere."""
    h0 = 1.0 / (8 * bpd[1])
    bc = (BC_WALL,) * 3
    blocks = uniform_blocks(*bpd)
    return Workload(f"uniform_{8*bpd[0]}x{8*bpd[1]}x{8*bpd[2]}_walls_blobs256_sphere", bpd, h0,
                    bc, 0, blocks,
                    lambda step, seed: rhs_blobs(blocks, h0, bpd, bc, 256, (0.02, 0.08), seed,
                                                 step, 0.005, SPHERE))


def c3_amr_sphere(bpd=(16, 8, 8)) -> Workload:
    """Base 16x8x8 blocks, 3 levels refined within 0.3/0.15 of the sphere surface."""
    h0 = 1.0 / (8 * bpd[1])
    bc = (BC_WALL,) * 3
    blocks = amr_sphere_blocks(bpd, h0, 2, (0.3, 0.15), *SPHERE)
    return Workload("amr3_sphere_blobs256", bpd, h0, bc, 2, blocks,
                    lambda step, seed: rhs_blobs(blocks, h0, bpd, bc, 256, (0.02, 0.08), seed,
                                                 step, 0.005, SPHERE))


def c4_uniform_sequence(bpd=(32, 16, 16)) -> Workload:
    """Uniform as C2; 64 blobs, sigma 0.04, moving 0.005 per step."""
    h0 = 1.0 / (8 * bpd[1])
    bc = (BC_WALL,) * 3
    blocks = uniform_blocks(*bpd)
    return Workload(f"uniform_{8*bpd[0]}x{8*bpd[1]}x{8*bpd[2]}_walls_blobs64_sequence", bpd, h0,
                    bc, 0, blocks,
                    lambda step, seed: rhs_blobs(blocks, h0, bpd, bc, 64, 0.04, seed, step,
                                                 0.005, None))


def c5_amr_school(bpd=(16, 8, 8), n_ell=50, lmax=3) -> Workload:
    """Base 16x8x8 blocks, up to 3 refinement levels near 50 ellipsoids: level
    l+1 wherever a block centre is within 0.18 / 2^l of an ellipsoid surface
    (59 012 blocks, 30.2 M cells at the defaults)."""
    h0 = 1.0 / (8 * bpd[1])
    bc = (BC_WALL,) * 3
    thr = (0.18, 0.09, 0.045)[:lmax]
    school = (n_ell, SCHOOL_AXES, SEED)
    blocks = amr_school_blocks(bpd, h0, lmax, thr, n_ell, SCHOOL_AXES, SEED)
    return Workload(f"amr{lmax+1}_school{n_ell}_blobs64_sequence", bpd, h0, bc, lmax, blocks,
                    lambda step, seed: rhs_blobs(blocks, h0, bpd, bc, 64, 0.04, seed, step,
                                                 0.005, None, school))
