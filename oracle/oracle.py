"""ctypes wrapper of the CPU oracle (oracle/libdhara_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg, as the checker / CPU baseline.  The product package never imports it.
"""
import ctypes as C
import os
import subprocess

import numpy as np

from dhara_amd import _abi
from dhara_amd.model import BaseModel, DharaError

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libdhara_oracle.so")

LIN_PCG_JACOBI, LIN_BICGSTAB_REF, LIN_DENSE, LIN_PCG_LINE = 0, 1, 2, 3
OLF_PCG_JACOBI, OLF_CG_REF, OLF_DENSE = 0, 1, 2

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        _abi.declare(L, "dor_")
        P = C.POINTER
        dp, ip = P(C.c_double), P(C.c_int32)
        L.dor_create.argtypes = [P(_abi.Grid), P(_abi.Params), C.c_int32, C.c_int32, P(C.c_void_p)]
        L.dor_eval_constitutive_refform.argtypes = [C.c_void_p, dp, dp, dp, dp]
        L.dor_assemble_subsurface.argtypes = [C.c_void_p, dp, dp, dp]
        L.dor_overland_conductance.argtypes = [C.c_void_p, dp, dp, dp, dp]
        L.dor_assemble_overland.argtypes = [C.c_void_p, dp, dp]
        L.dor_set_threads.argtypes = [C.c_int32]
        L.dor_vg_derivatives.argtypes = [C.c_void_p, dp, dp, dp, dp, dp, dp]
        L.dor_poisson5pt.argtypes = [C.c_int32, C.c_int32, ip, dp]
        L.dor_poisson7pt.argtypes = [C.c_int32, C.c_int32, C.c_int32, ip, dp]
        L.dor_dia_spmv.argtypes = [C.c_int32, C.c_int32, ip, dp, dp, dp]
        for n in ("dor_dia_cg", "dor_dia_bicgstab"):
            getattr(L, n).argtypes = [C.c_int32, C.c_int32, ip, dp, dp, dp, C.c_int32, C.c_double,
                                      C.c_double, ip, dp, ip]
        L.dor_dia_pcg_jacobi.argtypes = [C.c_int32, C.c_int32, ip, dp, dp, dp, C.c_int32,
                                         C.c_double, ip, dp, ip]
        L.dor_line_solve.argtypes = [C.c_int32, C.c_int32, C.c_int32, dp, dp, dp]
        L.dor_dia_pcg_line.argtypes = [C.c_int32, C.c_int32, C.c_int32, dp, dp, dp, C.c_int32, C.c_double,
                                       ip, dp, ip]
        L.dor_dense_solve.argtypes = [C.c_int32, dp, dp, dp]
        L.dor_thomas.argtypes = [C.c_int32, dp, dp, dp, dp, dp]
        L.dor_column_run.argtypes = [C.c_int32, C.c_double, P(_abi.Params), C.c_double, C.c_double,
                                     C.c_int32, dp, dp, ip]
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int32))


def set_threads(n):
    return lib().dor_set_threads(n)


class OracleModel(BaseModel):
    def __init__(self, grid, params=None, lin_solver=LIN_PCG_JACOBI, olf_solver=OLF_PCG_JACOBI):
        L = lib()
        params = params if params is not None else _abi.default_params()
        h = C.c_void_p()
        rc = L.dor_create(C.byref(grid), C.byref(params), lin_solver, olf_solver, C.byref(h))
        if rc != 0:
            raise DharaError(rc, "dor_create failed")
        super().__init__(L, "dor_", h, grid, params)

    def eval_constitutive_refform(self, psi):
        psi = np.ascontiguousarray(psi, dtype=np.float64)
        n = psi.size
        th, k, c = (np.empty(n) for _ in range(3))
        self._check(self._lib.dor_eval_constitutive_refform(self._h, _dp(psi), _dp(th), _dp(k), _dp(c)),
                    "refform")
        return th, k, c

    def vg_derivatives(self, psi):
        """dK/dpsi, dtheta/dalpha, dtheta/dn, dK/dalpha, dK/dn of every cell at psi."""
        psi = np.ascontiguousarray(psi, dtype=np.float64)
        out = [np.empty(psi.size) for _ in range(5)]
        self._check(self._lib.dor_vg_derivatives(self._h, _dp(psi), *[_dp(a) for a in out]), "vg_derivatives")
        return out

    def assemble(self):
        n = self.size("PSI")
        a, d, b = np.empty(7 * n), np.empty(n), np.empty(n)
        self._check(self._lib.dor_assemble_subsurface(self._h, _dp(a), _dp(d), _dp(b)), "assemble")
        return a.reshape(7, n), d, b

    def overland_conductance(self):
        s = self.grid.nx * self.grid.ny
        ke, kn, kw, ks = (np.empty(s) for _ in range(4))
        self._check(self._lib.dor_overland_conductance(self._h, _dp(ke), _dp(kn), _dp(kw), _dp(ks)),
                    "overland_conductance")
        return ke, kn, kw, ks

    def assemble_overland(self):
        s = self.grid.nx * self.grid.ny
        a, b = np.empty(5 * s), np.empty(s)
        self._check(self._lib.dor_assemble_overland(self._h, _dp(a), _dp(b)), "assemble_overland")
        return a.reshape(5, s), b


# ---- CUSP-semantics building blocks ------------------------------------------------------
def poisson5pt(m, n):
    off = np.empty(5, np.int32)
    val = np.empty(5 * m * n)
    assert lib().dor_poisson5pt(m, n, _ip(off), _dp(val)) == 0
    return off, val.reshape(5, m * n)


def poisson7pt(m, n, k):
    off = np.empty(7, np.int32)
    val = np.empty(7 * m * n * k)
    assert lib().dor_poisson7pt(m, n, k, _ip(off), _dp(val)) == 0
    return off, val.reshape(7, m * n * k)


def dia_spmv(off, val, x):
    off = np.ascontiguousarray(off, np.int32)
    val = np.ascontiguousarray(val, np.float64)
    x = np.ascontiguousarray(x, np.float64)
    y = np.empty_like(x)
    assert lib().dor_dia_spmv(x.size, off.size, _ip(off), _dp(val), _dp(x), _dp(y)) == 0
    return y


def dia_to_dense(off, val):
    n = val.shape[1]
    A = np.zeros((n, n))
    for d, o in enumerate(off):
        for r in range(n):
            c = r + int(o)
            if 0 <= c < n:
                A[r, c] += val[d, r]
    return A


def _krylov(name, off, val, b, x0, maxit, rtol, atol=None):
    off = np.ascontiguousarray(off, np.int32)
    val = np.ascontiguousarray(val, np.float64)
    b = np.ascontiguousarray(b, np.float64)
    x = np.array(x0, dtype=np.float64, copy=True)
    it = np.zeros(1, np.int32)
    rn = np.zeros(1)
    cv = np.zeros(1, np.int32)
    fn = getattr(lib(), name)
    if atol is None:
        rc = fn(b.size, off.size, _ip(off), _dp(val), _dp(b), _dp(x), maxit, rtol, _ip(it), _dp(rn), _ip(cv))
    else:
        rc = fn(b.size, off.size, _ip(off), _dp(val), _dp(b), _dp(x), maxit, rtol, atol, _ip(it), _dp(rn),
                _ip(cv))
    assert rc == 0
    return x, int(it[0]), float(rn[0]), bool(cv[0])


def dia_cg(off, val, b, x0, maxit, rtol, atol=0.0):
    return _krylov("dor_dia_cg", off, val, b, x0, maxit, rtol, atol)


def dia_bicgstab(off, val, b, x0, maxit, rtol, atol=0.0):
    return _krylov("dor_dia_bicgstab", off, val, b, x0, maxit, rtol, atol)


def dia_pcg_jacobi(off, val, b, x0, maxit, rtol):
    return _krylov("dor_dia_pcg_jacobi", off, val, b, x0, maxit, rtol)


def line_solve(shape, val, r):
    """z = T^-1 r, T the vertical tridiagonal part of a 7-point DIA matrix (shape nx, ny, nz)."""
    nx, ny, nz = shape
    val = np.ascontiguousarray(val, np.float64)
    r = np.ascontiguousarray(r, np.float64)
    z = np.empty_like(r)
    assert lib().dor_line_solve(nx, ny, nz, _dp(val), _dp(r), _dp(z)) == 0
    return z


def dia_pcg_line(shape, val, b, x0, maxit, rtol):
    nx, ny, nz = shape
    val = np.ascontiguousarray(val, np.float64)
    b = np.ascontiguousarray(b, np.float64)
    x = np.array(x0, dtype=np.float64, copy=True)
    it, cv = np.zeros(1, np.int32), np.zeros(1, np.int32)
    rn = np.zeros(1)
    assert lib().dor_dia_pcg_line(nx, ny, nz, _dp(val), _dp(b), _dp(x), maxit, rtol, _ip(it), _dp(rn), _ip(cv)) == 0
    return x, int(it[0]), float(rn[0]), bool(cv[0])


def dense_solve(A, b):
    A = np.array(A, dtype=np.float64, order="C", copy=True)
    b = np.ascontiguousarray(b, np.float64)
    x = np.empty_like(b)
    rc = lib().dor_dense_solve(b.size, _dp(A), _dp(b), _dp(x))
    if rc != 0:
        raise DharaError(rc, "singular")
    return x


def thomas(a, b, c, d):
    a, b, c, d = (np.ascontiguousarray(v, np.float64) for v in (a, b, c, d))
    x = np.empty_like(d)
    assert lib().dor_thomas(d.size, _dp(a), _dp(b), _dp(c), _dp(d), _dp(x)) == 0
    return x


def column_run(nz, dz, params, Ks, qtop, nsteps, psi0):
    psi = np.array(psi0, dtype=np.float64, copy=True)
    th = np.empty(nz)
    it = np.zeros(1, np.int32)
    rc = lib().dor_column_run(nz, dz, C.byref(params), Ks, qtop, nsteps, _dp(psi), _dp(th), _ip(it))
    assert rc == 0
    return psi, th, int(it[0])
