"""ctypes wrapper of the CPU FP64 oracle (oracle/hbem_oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker / timed CPU baseline.  The product
path (helmholtzbem_amd) never imports this module.

All matrices are numpy complex128 arrays in column-major (Fortran) order, the
layout of the reference's Eigen::MatrixXcd.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

LAYER_DOUBLE, LAYER_SINGLE, LAYER_HYPER, LAYER_ALL = 1, 2, 4, 7
QUIRK_SLP_COINC_DER = 1

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


def build(force: bool = False) -> str:
    """Compile liboracle.so with the committed Makefile."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-B" if force else "-s", "-C", _HERE, "liboracle.so"], check=True,
                       stdout=subprocess.DEVNULL)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.orc_solop_create.restype = ctypes.c_void_p
        L.orc_solop_create.argtypes = [_dp, ctypes.c_int, _dp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.orc_solop_destroy.argtypes = [ctypes.c_void_p]
        for name in ("orc_solop_factor", "orc_solop_mass", "orc_gen_sol_op", "orc_projection", "orc_sigma_sweep"):
            getattr(L, name).argtypes = None
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(_dp) if a is not None else None


def _check(rc):
    if rc != 0:
        buf = ctypes.create_string_buffer(512)
        lib().orc_last_error(buf, 512)
        raise RuntimeError(buf.value.decode())


def _geom(mesh):
    pan = np.ascontiguousarray(mesh.packed, dtype=np.float64)
    coefs = np.ascontiguousarray(mesh.coefs, dtype=np.float64)
    return pan, pan.shape[0], coefs, coefs.shape[0]


def gauss_rule(n, a=0.0, b=1.0):
    x = np.empty(n)
    w = np.empty(n)
    _check(lib().orc_gauss_rule(ctypes.c_double(a), ctypes.c_double(b), n, _p(x), _p(w)))
    return x, w


def composite_rule(n):
    m = n * (n + 1) // 2 - 1
    x = np.empty(m)
    w = np.empty(m)
    _check(lib().orc_composite_rule(n, _p(x), _p(w)))
    return x, w


def jy01(x):
    """Returns (J0, Y0, J1, Y1) arrays."""
    x = np.ascontiguousarray(np.atleast_1d(np.asarray(x, dtype=np.float64)))
    out = np.empty((x.size, 4))
    _check(lib().orc_jy01(_p(x), x.size, _p(out)))
    return out[:, 0], out[:, 1], out[:, 2], out[:, 3]


def hankel01_cplx(z):
    """H0(z), H1(z) of the first kind for complex z (Im z >= 0), as the
    builder's complex path computes them; returns two complex arrays."""
    z = np.ascontiguousarray(np.atleast_1d(np.asarray(z, dtype=np.complex128)))
    out = np.empty(2 * z.size, dtype=np.complex128)
    _check(lib().orc_hankel01_cplx(_p(z), z.size, _p(out)))
    return out[0::2], out[1::2]


def panel_eval(mesh, i, what, s):
    pan, n, coefs, nc = _geom(mesh)
    s = np.ascontiguousarray(np.atleast_1d(np.asarray(s, dtype=np.float64)))
    out = np.empty((s.size, 2))
    _check(lib().orc_panel_eval(_p(pan), n, _p(coefs), nc, i, what, _p(s), s.size, _p(out)))
    return out[:, 0] + 1j * out[:, 1]


def classify(mesh):
    pan, n, coefs, nc = _geom(mesh)
    cls = np.empty(n * (n + 1) // 2, dtype=np.int32)
    _check(lib().orc_classify(_p(pan), n, _p(coefs), nc, cls.ctypes.data_as(_ip)))
    return cls


def pair_local(mesh, order, k, c, i, j, d=2, flags=0):
    """Local matrices of pair (i >= j): dict K[d,o,m,n], W[d,m,n], V[d,m,n]."""
    pan, n, coefs, nc = _geom(mesh)
    out = np.zeros(96)
    k = complex(k)
    _check(lib().orc_pair_local(_p(pan), n, _p(coefs), nc, order, ctypes.c_double(k.real), ctypes.c_double(k.imag),
                                ctypes.c_double(c), i, j, d, flags, _p(out)))
    z = out[0::2] + 1j * out[1::2]
    return {"K": z[:24].reshape(3, 2, 2, 2), "W": z[24:36].reshape(3, 2, 2), "V": z[36:48].reshape(3, 2, 2)}


def assemble(mesh, order, k, c, layers=LAYER_ALL, d=0, flags=0, nthreads=8):
    """GalerkinBuilder::assembleDense; returns (K, W, V), each (d+1, N, N) column-major per slice."""
    pan, n, coefs, nc = _geom(mesh)
    shp = (n, n, d + 1)
    K = np.zeros(shp, dtype=np.complex128, order="F")
    W = np.zeros(shp, dtype=np.complex128, order="F")
    V = np.zeros(shp, dtype=np.complex128, order="F")
    k = complex(k)
    _check(lib().orc_assemble(_p(pan), n, _p(coefs), nc, order, ctypes.c_double(k.real), ctypes.c_double(k.imag),
                              ctypes.c_double(c), layers, d, flags, nthreads, _p(K), _p(W), _p(V)))
    return (np.moveaxis(K, 2, 0), np.moveaxis(W, 2, 0), np.moveaxis(V, 2, 0))


def mass_matrix(mesh, order):
    pan, n, coefs, nc = _geom(mesh)
    M = np.zeros((n, n), order="F")
    _check(lib().orc_mass_matrix(_p(pan), n, _p(coefs), nc, order, _p(M)))
    return M


def ldlt_factor(A):
    A = np.asfortranarray(A, dtype=np.float64)
    n = A.shape[0]
    F = np.zeros((n, n), order="F")
    perm = np.zeros(n, dtype=np.int32)
    _check(lib().orc_ldlt_factor(_p(A), n, _p(F), perm.ctypes.data_as(_ip)))
    return F, perm


class SolOp:
    """SolutionsOperator (gen_sol_op.cpp:23-196) on the oracle."""

    def __init__(self, mesh, order, flags=0, proj=True):
        pan, n, coefs, nc = _geom(mesh)
        self._keep = (pan, coefs)
        self.N = n
        self.proj = proj
        self.h = lib().orc_solop_create(_p(pan), n, _p(coefs), nc, order, flags, 1 if proj else 0)
        if not self.h:
            _check(-1)

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_solop_destroy(self.h)
            self.h = None

    def factor(self):
        n = 2 * self.N
        F = np.zeros((n, n), order="F")
        perm = np.zeros(n, dtype=np.int32)
        _check(lib().orc_solop_factor(ctypes.c_void_p(self.h), _p(F), perm.ctypes.data_as(_ip)))
        return F, perm

    def mass_matrix(self):
        M = np.zeros((self.N, self.N), order="F")
        _check(lib().orc_solop_mass(ctypes.c_void_p(self.h), _p(M)))
        return M

    def gen_sol_op(self, k, c_o, c_i, d=0, nthreads=8):
        n = 2 * self.N
        Ts = [np.zeros((n, n), dtype=np.complex128, order="F") for _ in range(d + 1)]
        ptrs = [_p(T) for T in Ts] + [None] * (2 - d)
        k = complex(k)
        _check(lib().orc_gen_sol_op(ctypes.c_void_p(self.h), ctypes.c_double(k.real), ctypes.c_double(k.imag),
                                    ctypes.c_double(c_o), ctypes.c_double(c_i), ctypes.c_int(d),
                                    ctypes.c_int(nthreads), *ptrs))
        return Ts

    def projection(self, T):
        T = np.array(T, dtype=np.complex128, order="F", copy=True)
        _check(lib().orc_projection(ctypes.c_void_p(self.h), _p(T)))
        return T

    def sigma_sweep(self, ks, c_o, c_i, R, q, nthreads=8):
        ks = np.ascontiguousarray(ks, dtype=np.float64)
        R = np.asfortranarray(R, dtype=np.complex128)
        out = np.zeros(ks.size)
        _check(lib().orc_sigma_sweep(ctypes.c_void_p(self.h), _p(ks), ctypes.c_int(ks.size), ctypes.c_double(c_o),
                                     ctypes.c_double(c_i), _p(R), ctypes.c_int(R.shape[1]), ctypes.c_int(q),
                                     ctypes.c_int(nthreads), _p(out)))
        return out


def sv(T, R, q=0):
    T = np.asfortranarray(T, dtype=np.complex128)
    R = np.asfortranarray(R, dtype=np.complex128)
    s = ctypes.c_double()
    _check(lib().orc_sv(_p(T), T.shape[0], _p(R), R.shape[1], q, ctypes.byref(s)))
    return s.value


def sv_der(T, Td, R, q=0):
    T = np.asfortranarray(T, dtype=np.complex128)
    Td = np.asfortranarray(Td, dtype=np.complex128)
    R = np.asfortranarray(R, dtype=np.complex128)
    out = np.zeros(2)
    _check(lib().orc_sv_der(_p(T), _p(Td), T.shape[0], _p(R), R.shape[1], q, _p(out)))
    return out


def sv_der2(T, Td, Tdd, R, q):
    T = np.asfortranarray(T, dtype=np.complex128)
    Td = np.asfortranarray(Td, dtype=np.complex128)
    Tdd = np.asfortranarray(Tdd, dtype=np.complex128)
    R = np.asfortranarray(R, dtype=np.complex128)
    out = np.zeros(3)
    _check(lib().orc_sv_der2(_p(T), _p(Td), _p(Tdd), T.shape[0], _p(R), R.shape[1], q, _p(out)))
    return out


def svd_values(B):
    B = np.asfortranarray(B, dtype=np.complex128)
    out = np.zeros(B.shape[1])
    _check(lib().orc_svd_values(_p(B), B.shape[0], B.shape[1], _p(out)))
    return out


def tp_A_rhs(mesh, order, k, c_o, c_i, U, proj=False, nthreads=8):
    """tp::direct_second_kind::A_rhs (solvers.cpp:106-131) composed from the
    oracle's operators: A = gen_sol_op(k, c_o, c_i) and rhs = A_o U with
    A_o = [[-K_o + M/2, V_o], [W_o, K_o^T + M/2]], K_o, W_o, V_o assembled at
    (k, c_o).  U: 2N x nrhs interpolation coefficients of the incoming waves."""
    S = SolOp(mesh, order, proj=proj)
    A = S.gen_sol_op(k, c_o, c_i, d=0, nthreads=nthreads)[0]
    M = S.mass_matrix()
    K, W, V = assemble(mesh, order, k, c_o, nthreads=nthreads)
    Ko = K[0]
    Ao = np.block([[-Ko + 0.5 * M, V[0]], [W[0], Ko.T + 0.5 * M]])
    return A, Ao @ np.asarray(U, dtype=np.complex128)


def tp_solve(mesh, order, k, c_o, c_i, U, nthreads=8):
    """tp::direct_second_kind::solve (solvers.cpp:133-152): the unprojected
    system of tp_A_rhs solved by LAPACK's zgesv (numpy.linalg.solve), the
    routine the reference calls (solvers.cpp:25-51), after its finiteness
    checks."""
    A, rhs = tp_A_rhs(mesh, order, k, c_o, c_i, U, proj=False, nthreads=nthreads)
    if not np.all(np.isfinite(rhs)):
        raise RuntimeError("Failed to solve for traces, right-hand side is undefined")
    if not np.all(np.isfinite(A)):
        raise RuntimeError("Failed to solve for traces, solutions operator is undefined")
    return np.linalg.solve(A, rhs)
