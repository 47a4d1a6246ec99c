"""ctypes binding of the CPU oracle (oracle/oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py, as the checker. The product path never loads it.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_libs: dict = {}

_i32p = ctypes.POINTER(ctypes.c_int)
_i64p = ctypes.POINTER(ctypes.c_int64)
_f64p = ctypes.POINTER(ctypes.c_double)


def _load(fast: bool = False):
    name = "liboracle_fast.so" if fast else "liboracle.so"
    if name in _libs:
        return _libs[name]
    path = os.path.join(HERE, name)
    if not os.path.exists(path):
        raise RuntimeError(f"{path} missing: run `make -C oracle`")
    lib = ctypes.CDLL(path)
    lib.orc_create.restype = ctypes.c_void_p
    lib.orc_create.argtypes = [_i32p, ctypes.c_int, _i32p, ctypes.c_double, ctypes.c_int,
                               ctypes.c_int, _i32p]
    lib.orc_destroy.argtypes = [ctypes.c_void_p]
    lib.orc_nnz.restype = ctypes.c_int64
    lib.orc_nnz.argtypes = [ctypes.c_void_p]
    lib.orc_bmean_row.restype = ctypes.c_int
    lib.orc_bmean_row.argtypes = [ctypes.c_void_p]
    lib.orc_get_csr.argtypes = [ctypes.c_void_p, _i64p, _i64p, _f64p]
    lib.orc_get_pinv.argtypes = [ctypes.c_void_p, _f64p]
    lib.orc_pinv.argtypes = [_f64p]
    for f in ("orc_spmv", "orc_spmv_raw", "orc_precond", "orc_block_cg", "orc_lhs_cpu"):
        getattr(lib, f).argtypes = [ctypes.c_void_p, _f64p, _f64p]
    lib.orc_bicgstab.restype = ctypes.c_int
    lib.orc_bicgstab.argtypes = [ctypes.c_void_p, _f64p, _f64p, _f64p, ctypes.c_double,
                                 ctypes.c_double, ctypes.c_int, ctypes.c_int, _i32p, _f64p,
                                 _f64p]
    lib.orc_pipelined.restype = ctypes.c_int
    lib.orc_pipelined.argtypes = [ctypes.c_void_p, _f64p, _f64p, _f64p, ctypes.c_double,
                                  ctypes.c_double, ctypes.c_int, _i32p, _f64p]
    lib.orc_pressure_rhs.argtypes = [ctypes.c_void_p, _f64p, _f64p, _f64p, ctypes.c_double,
                                     _f64p]
    lib.orc_grad_p_update.argtypes = [ctypes.c_void_p, _f64p, ctypes.c_double, _f64p]
    lib.orc_diffusion_lhs.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_double, _f64p,
                                      _f64p]
    lib.orc_diffusion_lap_vec.argtypes = [ctypes.c_void_p, ctypes.c_int, _f64p, _f64p]
    for f in ("orc_diffusion_precond", "orc_diffusion_block_cg"):
        getattr(lib, f).argtypes = [ctypes.c_void_p, ctypes.c_double, _f64p, _f64p]
    lib.orc_diffusion_bicgstab.restype = ctypes.c_int
    lib.orc_diffusion_bicgstab.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_double,
                                           _f64p, _f64p, _f64p, ctypes.c_double,
                                           ctypes.c_double, ctypes.c_int, ctypes.c_int, _i32p,
                                           _f64p, _f64p]
    lib.orc_diffusion_pipelined.restype = ctypes.c_int
    lib.orc_diffusion_pipelined.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_double,
                                            _f64p, _f64p, _f64p, ctypes.c_double,
                                            ctypes.c_double, ctypes.c_int, _i32p, _f64p]
    lib.orc_diffusion_step.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_double,
                                       ctypes.c_double, ctypes.c_double, ctypes.c_int, _f64p,
                                       _f64p, _i32p, _f64p]
    lib.orc_last_error.restype = ctypes.c_char_p
    lib.orc_num_threads.restype = ctypes.c_int
    _libs[name] = lib
    return lib


def _p(a, t=ctypes.c_double):
    return None if a is None else a.ctypes.data_as(ctypes.POINTER(t))


def pinv() -> np.ndarray:
    """P_inv (512x512, row-major) by Cholesky-Crout + Gauss-Jordan."""
    P = np.zeros((512, 512), np.float64)
    _load().orc_pinv(_p(P))
    return P


class Oracle:
    """FP64 CPU restatement of the reference solver on one block grid."""

    def __init__(self, blocks, bpd, h0, bc, bmean=1, level_max=0, fast=False):
        self.lib = _load(fast)
        self.blocks = np.ascontiguousarray(blocks, np.int32)
        bpd = np.ascontiguousarray(bpd, np.int32)
        bc = np.ascontiguousarray(bc, np.int32)
        self.n = len(self.blocks) * 512
        self.h = h0 / (2.0 ** self.blocks[:, 0])
        self.bmean = bmean
        self.obj = self.lib.orc_create(_p(bpd, ctypes.c_int), level_max, _p(bc, ctypes.c_int),
                                       h0, bmean, len(self.blocks), _p(self.blocks, ctypes.c_int))
        if not self.obj:
            raise RuntimeError(self.lib.orc_last_error().decode())

    def __del__(self):
        if getattr(self, "obj", None):
            self.lib.orc_destroy(self.obj)
            self.obj = None

    @property
    def bmean_row(self) -> int:
        return self.lib.orc_bmean_row(self.obj)

    def csr(self):
        nnz = self.lib.orc_nnz(self.obj)
        rowptr = np.zeros(self.n + 1, np.int64)
        col = np.zeros(nnz, np.int64)
        val = np.zeros(nnz, np.float64)
        self.lib.orc_get_csr(self.obj, _p(rowptr, ctypes.c_int64), _p(col, ctypes.c_int64),
                             _p(val))
        return rowptr, col, val

    def scipy_matrix(self):
        import scipy.sparse as sp
        rowptr, col, val = self.csr()
        return sp.csr_matrix((val, col, rowptr), shape=(self.n, self.n))

    def _vec(self, x):
        return np.ascontiguousarray(np.asarray(x, np.float64).reshape(-1))

    def _apply(self, fn, x):
        x = self._vec(x)
        y = np.zeros(self.n, np.float64)
        getattr(self.lib, fn)(self.obj, _p(x), _p(y))
        return y

    def spmv(self, x):
        """y = A x with the GPU solver's bMean row (BiCGSTAB.cu:392-413)."""
        return self._apply("orc_spmv", x)

    def spmv_raw(self, x):
        return self._apply("orc_spmv_raw", x)

    def precond(self, p):
        return self._apply("orc_precond", p)

    def block_cg(self, p):
        return self._apply("orc_block_cg", p)

    def lhs_cpu(self, x):
        return self._apply("orc_lhs_cpu", x)

    def pressure_rhs(self, vel, dt, udef=None, chi=None):
        """KernelPressureRHS: vel, udef [n, 3]; chi [n] (udef, chi: obstacle)."""
        vel = np.ascontiguousarray(np.asarray(vel, np.float64).reshape(self.n, 3))
        udef = None if udef is None else np.ascontiguousarray(
            np.asarray(udef, np.float64).reshape(self.n, 3))
        chi = None if chi is None else self._vec(chi)
        out = np.zeros(self.n, np.float64)
        self.lib.orc_pressure_rhs(self.obj, _p(vel), _p(udef), _p(chi), float(dt), _p(out))
        return out

    def grad_p_update(self, p, vel, dt):
        """KernelGradP + velocity update: returns vel + o / h^3, [n, 3]."""
        v = np.ascontiguousarray(np.asarray(vel, np.float64).reshape(self.n, 3)).copy()
        self.lib.orc_grad_p_update(self.obj, _p(self._vec(p)), float(dt), _p(v))
        return v

    def bicgstab(self, rhs, x0=None, tol_abs=0.0, tol_rel=1e-7, max_restarts=100,
                 max_iter=1000, history=False):
        rhs = self._vec(rhs)
        x0 = None if x0 is None else self._vec(x0)
        x = np.zeros(self.n, np.float64)
        ir = np.zeros(2, np.int32)
        errs = np.zeros(2, np.float64)
        hist = np.zeros(max_iter, np.float64) if history else None
        st = self.lib.orc_bicgstab(self.obj, _p(rhs), _p(x0), _p(x), tol_abs, tol_rel,
                                   max_restarts, max_iter, _p(ir, ctypes.c_int), _p(errs),
                                   _p(hist))
        info = dict(status=st, iters=int(ir[0]), restarts=int(ir[1]), err_init=float(errs[0]),
                    err_final=float(errs[1]))
        if history:
            info["history"] = hist[: info["iters"]].copy()
        return x, info

    def pipelined(self, rhs, x0=None, tol_abs=0.0, tol_rel=1e-7, max_iter=1000):
        rhs = self._vec(rhs)
        x0 = None if x0 is None else self._vec(x0)
        x = np.zeros(self.n, np.float64)
        ir = np.zeros(2, np.int32)
        errs = np.zeros(2, np.float64)
        st = self.lib.orc_pipelined(self.obj, _p(rhs), _p(x0), _p(x), tol_abs, tol_rel, max_iter,
                                    _p(ir, ctypes.c_int), _p(errs))
        return x, dict(status=st, iters=int(ir[0]), restarts=int(ir[1]),
                       err_init=float(errs[0]), err_final=float(errs[1]))

    # ---- implicit diffusion (sigma = 1 / (nu dt)) ----
    def diffusion_lhs(self, x, comp, sigma):
        """KernelLHSDiffusion on a scalar field, boundary ghosts of velocity
        component `comp`; sigma = 0 gives KernelDiffusionRHS's h lap."""
        x = self._vec(x)
        y = np.zeros(self.n, np.float64)
        self.lib.orc_diffusion_lhs(self.obj, comp, float(sigma), _p(x), _p(y))
        return y

    def diffusion_lap_vec(self, vel, comp):
        vel = np.ascontiguousarray(np.asarray(vel, np.float64).reshape(self.n, 3))
        y = np.zeros(self.n, np.float64)
        self.lib.orc_diffusion_lap_vec(self.obj, comp, _p(vel), _p(y))
        return y

    def diffusion_precond(self, p, sigma):
        y = np.zeros(self.n, np.float64)
        self.lib.orc_diffusion_precond(self.obj, float(sigma), _p(self._vec(p)), _p(y))
        return y

    def diffusion_block_cg(self, p, sigma):
        y = np.zeros(self.n, np.float64)
        self.lib.orc_diffusion_block_cg(self.obj, float(sigma), _p(self._vec(p)), _p(y))
        return y

    def diffusion_bicgstab(self, rhs, comp, sigma, x0=None, tol_abs=0.0, tol_rel=1e-7,
                           max_restarts=100, max_iter=1000, history=False):
        rhs = self._vec(rhs)
        x0 = None if x0 is None else self._vec(x0)
        x = np.zeros(self.n, np.float64)
        ir = np.zeros(2, np.int32)
        errs = np.zeros(2, np.float64)
        hist = np.zeros(max_iter, np.float64) if history else None
        st = self.lib.orc_diffusion_bicgstab(self.obj, comp, float(sigma), _p(rhs), _p(x0),
                                             _p(x), tol_abs, tol_rel, max_restarts, max_iter,
                                             _p(ir, ctypes.c_int), _p(errs), _p(hist))
        info = dict(status=st, iters=int(ir[0]), restarts=int(ir[1]), err_init=float(errs[0]),
                    err_final=float(errs[1]))
        if history:
            info["history"] = hist[: info["iters"]].copy()
        return x, info

    def diffusion_pipelined(self, rhs, comp, sigma, x0=None, tol_abs=0.0, tol_rel=1e-7,
                            max_iter=1000):
        rhs = self._vec(rhs)
        x0 = None if x0 is None else self._vec(x0)
        x = np.zeros(self.n, np.float64)
        ir = np.zeros(2, np.int32)
        errs = np.zeros(2, np.float64)
        st = self.lib.orc_diffusion_pipelined(self.obj, comp, float(sigma), _p(rhs), _p(x0),
                                              _p(x), tol_abs, tol_rel, max_iter,
                                              _p(ir, ctypes.c_int), _p(errs))
        return x, dict(status=st, iters=int(ir[0]), restarts=int(ir[1]),
                       err_init=float(errs[0]), err_final=float(errs[1]))

    def diffusion_step(self, vel_old, vel, nu, dt, tol_abs=1e-6, tol_rel=1e-4, max_iter=1000):
        """AdvectionDiffusionImplicit::euler after the advection kernel:
        returns u^{n+1} [n, 3] from u^n = vel_old and u* = vel."""
        vo = np.ascontiguousarray(np.asarray(vel_old, np.float64).reshape(self.n, 3))
        v = np.ascontiguousarray(np.asarray(vel, np.float64).reshape(self.n, 3)).copy()
        its = np.zeros(3, np.int32)
        errs = np.zeros(6, np.float64)
        self.lib.orc_diffusion_step(self.obj, float(nu), float(dt), tol_abs, tol_rel, max_iter,
                                    _p(vo), _p(v), _p(its, ctypes.c_int), _p(errs))
        return v, dict(iters=its.tolist(), err_init=errs[0::2].tolist(),
                       err_final=errs[1::2].tolist())


def _load_h5():
    if "h5" in _libs:
        return _libs["h5"]
    path = os.path.join(HERE, "liboracle_h5.so")
    if not os.path.exists(path):
        raise RuntimeError(f"{path} missing: run `make -C oracle` (needs HDF5)")
    lib = ctypes.CDLL(path)
    lib.orc_h5_groups.restype = ctypes.c_int64
    lib.orc_h5_groups.argtypes = [ctypes.c_char_p, _f64p, _i32p, ctypes.c_int64]
    lib.orc_h5_dataset.restype = ctypes.c_int64
    lib.orc_h5_dataset.argtypes = [ctypes.c_char_p, ctypes.POINTER(ctypes.c_float),
                                   ctypes.c_int64]
    lib.orc_h5_to_uniform.restype = ctypes.c_int64
    lib.orc_h5_to_uniform.argtypes = [ctypes.c_char_p, _i32p, ctypes.POINTER(ctypes.c_float),
                                      ctypes.c_int64]
    _libs["h5"] = lib
    return lib


def h5_groups(base, cap=1 << 22):
    """get_amr_groups (tools/uniform_conversion/main.cpp:59-99): origins
    [g, 4] and indices [g, 7] of <base>-groups.h5."""
    o = np.zeros((cap, 4), np.float64)
    ix = np.zeros((cap, 7), np.int32)
    n = _load_h5().orc_h5_groups(str(base).encode(), _p(o), _p(ix, ctypes.c_int), cap)
    if n < 0:
        raise RuntimeError("orc_h5_groups failed")
    return o[:n].copy(), ix[:n].copy()


def h5_dataset(base, cap=1 << 28):
    """get_amr_dataset (:102-150): dset of <base>.h5 as FP32."""
    out = np.zeros(cap, np.float32)
    n = _load_h5().orc_h5_dataset(str(base).encode(), _p(out, ctypes.c_float), cap)
    if n < 0:
        raise RuntimeError("orc_h5_dataset failed")
    return out[:n].copy()


def h5_to_uniform(base, cap=1 << 26):
    """convert_to_uniform (:152-238) on one rank: the uniform grid [z][y][x]
    at 1/8 of the finest resolution."""
    dims = np.zeros(3, np.int32)
    out = np.zeros(cap, np.float32)
    n = _load_h5().orc_h5_to_uniform(str(base).encode(), _p(dims, ctypes.c_int),
                                     _p(out, ctypes.c_float), cap)
    if n < 0:
        raise RuntimeError("orc_h5_to_uniform failed")
    nx, ny, nz = (int(d) for d in dims)
    return out[:n].reshape(nz, ny, nx).copy()


def num_threads(fast=True) -> int:
    return _load(fast).orc_num_threads()
