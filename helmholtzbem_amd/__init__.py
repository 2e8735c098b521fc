"""helmholtzbem_amd -- MI355X-native sigma_min(k) path of HelmholtzBEM.

Host-side mirror of the reference's C++ interface for the path (names,
argument meaning, defaults and error messages), over the C-ABI of
include/hbem.h:

* ``BuilderData`` / ``GalerkinBuilder`` / ``LayerType`` (galerkin_builder.hpp)
* ``SolutionsOperator`` (gen_sol_op.hpp)
* ``randomized_svd.{sv, sv_der, sv_der2, randGaussian}`` (randsvd.hpp)
* ``sigma_sweep`` -- the batched k-sweep the reference drivers run with
  ``std::transform(par_unseq, ...)`` (examples/roots_brent_square_rsvd.cpp:167-171)

Matrices are numpy complex128 arrays; results are Fortran-ordered like
Eigen::MatrixXcd.  Everything runs on the GPU through libhbem.so; nothing here
computes operator entries on the CPU.
"""
from __future__ import annotations

import ctypes
import enum
import math

import numpy as np

from . import _lib
from ._lib import HbemError, check, lib
from .geometry import (ParametrizedCircularArc, ParametrizedLine, ParametrizedMesh, ParametrizedTrigCurve,
                       Xoshiro256ss, circle_mesh, gaussian_test_matrix, polygon_mesh, random_polygon_instances,
                       random_polygon_mesh, square_mesh, star_curve, star_mesh)

__all__ = ["LayerType", "ContinuousSpace1", "BuilderData", "GalerkinBuilder", "SolutionsOperator",
           "randomized_svd", "sigma_sweep", "sigma_sweep_der2", "sigma_sweep_mixed", "sigma_sweep_multi", "sigma_sweep_batched",
           "random_polygon_instances", "HbemError", "ParametrizedLine", "ParametrizedCircularArc",
           "ParametrizedTrigCurve", "ParametrizedMesh", "square_mesh", "star_mesh", "circle_mesh",
           "polygon_mesh", "random_polygon_mesh", "gaussian_test_matrix", "Xoshiro256ss", "star_curve",
           "FLAG_QUIRK_SLP_COINC_DER", "FLAG_FAR_FIELD"]

FLAG_QUIRK_SLP_COINC_DER = 1
FLAG_FAR_FIELD = 2          # lower-order Gauss rules on separated panel pairs (HBEM_FLAG_FAR_FIELD)

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


def _dptr(a):
    return a.ctypes.data_as(_dp)


def _vptr(a):
    return ctypes.c_void_p(a.ctypes.data) if a is not None else None


def _square(name, T, n=None):
    """T as a Fortran-ordered complex n x n matrix; ValueError otherwise (the
    C-ABI reads n^2 entries from the pointer it is given)."""
    T = np.asfortranarray(T, dtype=np.complex128)
    if T.ndim != 2 or T.shape[0] != T.shape[1] or T.shape[0] < 1:
        raise ValueError("%s must be a square matrix, got shape %s" % (name, T.shape))
    if n is not None and T.shape[0] != n:
        raise ValueError("%s must be %d x %d, got %s" % (name, n, n, T.shape))
    return T


def _test_matrix(R, n):
    """R as a Fortran-ordered complex n x nc matrix with 1 <= nc <= min(8, n)."""
    R = np.asfortranarray(R, dtype=np.complex128)
    if R.ndim != 2 or R.shape[0] != n:
        raise ValueError("R must have %d rows, got shape %s" % (n, R.shape))
    if not 1 <= R.shape[1] <= min(8, n):
        raise ValueError("R must have between 1 and %d columns, got %d" % (min(8, n), R.shape[1]))
    return R


class LayerType(enum.IntFlag):
    """galerkin_builder.hpp:100-106."""
    NONE = 0
    DOUBLE = 1
    SINGLE = 2
    HYPERSINGULAR = 4
    ALL = 7


class ContinuousSpace1:
    """ContinuousSpace<1> (continuous_space.hpp:33-152): P1 on a closed curve."""

    q_ = 2

    def getQ(self):
        return 2

    def getSpaceDim(self, numpanels):
        return numpanels * (self.q_ - 1)

    def LocGlobMap(self, q, n, N):
        if not (q <= self.q_ and n <= N):
            raise IndexError("Panel/RSF index out of range!")
        return n if q == 2 else (1 if n % N == 0 else n + 1)

    def Interpolate_helmholtz(self, func, mesh):
        """Coefficients of the P1 interpolant: func(x1, x2) at the start point
        of every panel (continuous_space.hpp:67-82)."""
        panels = mesh.getPanels()
        out = np.zeros(self.getSpaceDim(len(panels)), dtype=np.complex128)
        for i, p in enumerate(panels):
            x = p[0.0]
            out[i] = func(x[0], x[1])
        return out

    def Interpolate_helmholtz_neu(self, func, mesh):
        """Normal derivatives n . func(x1, x2) at the panel start points, with
        func the gradient and n = (t_y, -t_x) / |t| from the panel tangent t
        there (continuous_space.hpp:84-101)."""
        panels = mesh.getPanels()
        out = np.zeros(self.getSpaceDim(len(panels)), dtype=np.complex128)
        for i, p in enumerate(panels):
            x = p[0.0]
            t = p.Derivative_01(0.0)
            nrm = math.sqrt(t[1] * t[1] + t[0] * t[0])
            g = func(x[0], x[1])
            out[i] = (t[1] / nrm) * g[0] + (-t[0] / nrm) * g[1]
        return out


class _Geometry:
    """Owns one device context (hbem_geometry_t)."""

    def __init__(self, mesh, order, proj, flags):
        self.mesh = mesh
        pan = np.ascontiguousarray(mesh.packed, dtype=np.float64)
        coefs = np.ascontiguousarray(mesh.coefs if mesh.coefs.size else np.zeros((1, 3)), dtype=np.float64)
        self._keep = (pan, coefs)
        h = ctypes.c_void_p()
        check(lib().hbem_geometry_create(_dptr(pan), pan.shape[0], _dptr(coefs), int(mesh.coefs.shape[0]), int(order),
                                         1 if proj else 0, int(flags), ctypes.byref(h)))
        self.h = h
        self.N = pan.shape[0]

    def __del__(self):
        h = getattr(self, "h", None)
        if h is not None and h.value:
            try:
                lib().hbem_geometry_destroy(h)
            except Exception:
                pass
            self.h = None


class BuilderData:
    """BuilderData(mesh, test_space, trial_space, order) (galerkin_builder.hpp:43-46).

    Test and trial space must be the same ContinuousSpace1 object (the only
    configuration of the reference's sigma_min drivers).  ``flags`` selects
    reference quirks (FLAG_QUIRK_SLP_COINC_DER reproduces the coinciding
    single-layer derivative copy of galerkin_builder.cpp:744-746).
    """

    def __init__(self, mesh, test_space, trial_space, order, flags=0):
        if test_space is not trial_space or not isinstance(test_space, ContinuousSpace1):
            raise ValueError("test and trial space must be the same ContinuousSpace1 object")
        self.mesh = mesh
        self.test_space = test_space
        self.trial_space = trial_space
        self.order = int(order)
        self.flags = int(flags)
        self._geo = _Geometry(mesh, order, True, flags)
        self._geo_noproj = None

    def _ctx(self, proj=True):
        if proj:
            return self._geo
        if self._geo_noproj is None:
            self._geo_noproj = _Geometry(self.mesh, self.order, False, self.flags)
        return self._geo_noproj

    def release_work(self):
        """Frees the device work buffers of this geometry's contexts (the
        geometry and its node tables stay); the next call allocates them
        again."""
        for g in (self._geo, self._geo_noproj):
            if g is not None:
                check(lib().hbem_geometry_release_work(g.h))

    def getTestSpaceDimension(self, n_panels=0):
        return self.test_space.getSpaceDim(n_panels or self.getNumPanels())

    def getTrialSpaceDimension(self, n_panels=0):
        return self.trial_space.getSpaceDim(n_panels or self.getNumPanels())

    def testTrialSpacesAreEqual(self):
        return self.test_space is self.trial_space

    def getQtest(self):
        return 2

    def getQtrial(self):
        return 2

    def getGaussQROrder(self):
        return self.order

    def getCGaussQROrder(self):
        return self.order * (self.order + 1) // 2 - 1

    def getNumPanels(self):
        return self.mesh.getNumPanels()

    def getPanelsAreLines(self):
        return self.mesh.isPolygonal()


class GalerkinBuilder:
    """GalerkinBuilder (galerkin_builder.hpp:117-277)."""

    def __init__(self, builder_data: BuilderData):
        self.data = builder_data
        self._dense = {}

    def getData(self):
        return self.data

    def assembleDense(self, k, c, layers=LayerType.ALL, d=0, row=0, col=0, rows=0, cols=0):
        k = complex(k)
        if rows or cols or row or col:
            raise NotImplementedError("submatrix assembly is not provided by the device path")
        N = self.data.getNumPanels()
        layers = LayerType(int(layers))
        outs = {}
        args = []
        for lt in (LayerType.DOUBLE, LayerType.HYPERSINGULAR, LayerType.SINGLE):
            if layers & lt:
                outs[lt] = np.zeros((d + 1, N, N), dtype=np.complex128)   # slices written column-major below
                args.append(outs[lt])
            else:
                args.append(None)
        check(lib().hbem_assemble_dense(self.data._geo.h, k.real, k.imag, float(c), int(layers), int(d),
                                        *[_vptr(a) for a in args]))
        self._dense = {}
        for lt, a in outs.items():
            # C-ABI slices are column-major N x N: reinterpret as (d+1, N, N) transposed per slice
            self._dense[lt] = [np.asfortranarray(a[t].T) for t in range(d + 1)]
        self._d = d

    def _get(self, layer, d):
        if d > 2:
            raise RuntimeError("Invalid order of derivative")
        if layer not in self._dense or d >= len(self._dense[layer]):
            raise RuntimeError("Requested layer is not computed, so cannot be retrieved")
        return self._dense[layer][d]

    def getDoubleLayer(self, d=0):
        return self._get(LayerType.DOUBLE, d)

    def getHypersingular(self, d=0):
        return self._get(LayerType.HYPERSINGULAR, d)

    def getSingleLayer(self, d=0):
        return self._get(LayerType.SINGLE, d)


class SolutionsOperator:
    """SolutionsOperator (gen_sol_op.hpp:40-127).

    The derivative variants return (T, T') and (T, T', T'') -- the intended
    semantics; the reference writes every assemble_operator call into T
    (gen_sol_op.cpp:150-151, 193-195).
    """

    def __init__(self, builder_data: BuilderData, proj=True):
        self.builder_data = builder_data
        self.do_projection = bool(proj)
        self._g = builder_data._ctx(self.do_projection)
        self._last = None
        self._ops = {}

    def _gen(self, k, c_o, c_i, d):
        k = complex(k)
        n = 2 * self.builder_data.getNumPanels()
        outs = [np.zeros((n, n), dtype=np.complex128, order="F") for _ in range(d + 1)]
        ptrs = [_vptr(o) for o in outs] + [None] * (2 - d)
        check(lib().hbem_gen_sol_op(self._g.h, k.real, k.imag, float(c_o), float(c_i), int(d), *ptrs))
        self._remember(k, c_o, c_i)
        return outs

    def _remember(self, k, c_o, c_i):
        self._last = (complex(k), float(c_o), float(c_i))
        self._ops = {}

    def _layer(self, layer, inner):
        """Operator of the last gen_sol_op* call (gen_sol_op.hpp:120-125),
        assembled on request on the device."""
        if self._last is None:
            raise RuntimeError("no operator has been generated yet")
        key = (layer, inner)
        if key not in self._ops:
            k, c_o, c_i = self._last
            b = GalerkinBuilder(self.builder_data)
            b.assembleDense(k, c_i if inner else c_o, layer)
            self._ops[key] = b._get(layer, 0)
        return self._ops[key]

    def K_int(self):
        return self._layer(LayerType.DOUBLE, True)

    def K_ext(self):
        return self._layer(LayerType.DOUBLE, False)

    def W_int(self):
        return self._layer(LayerType.HYPERSINGULAR, True)

    def W_ext(self):
        return self._layer(LayerType.HYPERSINGULAR, False)

    def V_int(self):
        return self._layer(LayerType.SINGLE, True)

    def V_ext(self):
        return self._layer(LayerType.SINGLE, False)

    def gen_sol_op(self, k, c_o, c_i, builder=None):
        return self._gen(k, c_o, c_i, 0)[0]

    def gen_sol_op_1st_der(self, k, c_o, c_i, builder=None):
        return tuple(self._gen(k, c_o, c_i, 1))

    def gen_sol_op_2nd_der(self, k, c_o, c_i, builder=None):
        return tuple(self._gen(k, c_o, c_i, 2))

    def mass_matrix(self):
        N = self.builder_data.getNumPanels()
        M = np.zeros((N, N), order="F")
        check(lib().hbem_mass_matrix(self._g.h, _dptr(M)))
        return M

    def projection_factor(self):
        n = 2 * self.builder_data.getNumPanels()
        F = np.zeros((n, n), order="F")
        tr = np.zeros(n, dtype=np.int32)
        check(lib().hbem_projection_factor(self._g.h, _dptr(F), tr.ctypes.data_as(_ip)))
        return F, tr

    def projection(self, T):
        if not self.do_projection:
            raise RuntimeError("projection is disabled for this SolutionsOperator")
        T = _square("T", T, 2 * self.builder_data.getNumPanels())
        out = np.zeros_like(T, order="F")
        check(lib().hbem_projection(self._g.h, _vptr(T), _vptr(out)))
        return out

    def getBuilderData(self):
        return self.builder_data


class randomized_svd:
    """namespace randomized_svd (randsvd.hpp:23-77)."""

    @staticmethod
    def randGaussian(nr, nc=None, seed=0):
        """Complex Gaussian test matrix, reproducible from ``seed`` (the
        reference's draws are unseeded, randsvd.cpp:49-69)."""
        if nc is None:
            # the vector overload omits the matrix form's 1/sqrt2 (randsvd.cpp:60-69)
            return gaussian_test_matrix(nr, 1, seed)[:, 0] * math.sqrt(2.0)
        return gaussian_test_matrix(nr, nc, seed)

    @staticmethod
    def benchmarking(yes):
        """randsvd.cpp:27-30: record the phase timers of sv / sv_der / sv_der2."""
        check(lib().hbem_rsvd_benchmarking(1 if yes else 0))

    @staticmethod
    def _timers():
        out = np.zeros(4)
        check(lib().hbem_rsvd_timers(_dptr(out)))
        return out

    @staticmethod
    def get_lu_time():
        return float(randomized_svd._timers()[0])

    @staticmethod
    def get_qr_time():
        return float(randomized_svd._timers()[1])

    @staticmethod
    def get_sub_iter_time():
        return float(randomized_svd._timers()[2])

    @staticmethod
    def get_svd_time():
        return float(randomized_svd._timers()[3])

    @staticmethod
    def reset_timer():
        check(lib().hbem_rsvd_reset_timer())

    @staticmethod
    def sv(T, R, q=0):
        T = _square("T", T)
        R = _test_matrix(R, T.shape[0])
        s = ctypes.c_double()
        check(lib().hbem_rsvd_sv(_vptr(T), T.shape[0], _vptr(R), R.shape[1], int(q), ctypes.byref(s)))
        return s.value

    @staticmethod
    def sv_der(T, T_der, R, q=0):
        T = _square("T", T)
        Td = _square("T_der", T_der, T.shape[0])
        R = _test_matrix(R, T.shape[0])
        out = np.zeros(2)
        check(lib().hbem_rsvd_sv_der(_vptr(T), _vptr(Td), T.shape[0], _vptr(R), R.shape[1], int(q), _dptr(out)))
        return out

    @staticmethod
    def sv_der2(T, T_der, T_der2, R, q):
        T = _square("T", T)
        Td = _square("T_der", T_der, T.shape[0])
        Tdd = _square("T_der2", T_der2, T.shape[0])
        R = _test_matrix(R, T.shape[0])
        out = np.zeros(3)
        check(lib().hbem_rsvd_sv_der2(_vptr(T), _vptr(Td), _vptr(Tdd), T.shape[0], _vptr(R), R.shape[1], int(q),
                                      _dptr(out)))
        return out


def assemble_dense_batch(builder_data: BuilderData, ks, c, layers=LayerType.ALL, d=0):
    """GalerkinBuilder::assembleDense(k, c, layers, d) for every real k of `ks`
    in one batched pass.  Returns {LayerType: array (nk, d+1, N, N)} with the
    matrices GalerkinBuilder.getDoubleLayer / getHypersingular / getSingleLayer
    would return for each k."""
    g = builder_data._ctx(True)
    ks = np.ascontiguousarray(ks, dtype=np.float64)
    N = builder_data.getNumPanels()
    layers = LayerType(int(layers))
    outs, args = {}, []
    for lt in (LayerType.DOUBLE, LayerType.HYPERSINGULAR, LayerType.SINGLE):
        if layers & lt:
            outs[lt] = np.zeros((ks.size, d + 1, N, N), dtype=np.complex128)
            args.append(outs[lt])
        else:
            args.append(None)
    check(lib().hbem_assemble_dense_batch(g.h, ks.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), ks.size,
                                          float(c), int(layers), int(d), *[_vptr(a) for a in args]))
    # slices are column-major N x N
    return {lt: np.swapaxes(a, 2, 3).copy() for lt, a in outs.items()}


def sigma_sweep(builder_data: BuilderData, ks, c_o, c_i, R, q=2, der=0, proj=True):
    """sigma_min(k) for every k (and dsigma/dk when der >= 1; der = 2 assembles
    the operators with their first and second k-derivatives)."""
    ks = np.ascontiguousarray(ks, dtype=np.float64)
    if ks.ndim != 1:
        raise ValueError("ks must be a 1-D array of wavenumbers")
    R = _test_matrix(R, 2 * builder_data.getNumPanels())
    g = builder_data._ctx(proj)
    sig = np.zeros(ks.size)
    dsig = np.zeros(ks.size) if der else None
    check(lib().hbem_sigma_sweep(g.h, _dptr(ks), ks.size, float(c_o), float(c_i), _vptr(R), R.shape[1], int(q),
                                 int(der), _dptr(sig), _dptr(dsig) if der else None))
    return (sig, dsig) if der else sig


def sigma_sweep_der2(builder_data: BuilderData, ks, c_o, c_i, R, q=2, proj=True):
    """(sigma, dsigma/dk, d2sigma/dk2) for every k: the operators with their
    first and second k-derivatives and randomized_svd::sv_der2's bordered
    system, batched on the device (the evaluations of the Newton refinement)."""
    ks = np.ascontiguousarray(ks, dtype=np.float64)
    if ks.ndim != 1:
        raise ValueError("ks must be a 1-D array of wavenumbers")
    R = _test_matrix(R, 2 * builder_data.getNumPanels())
    g = builder_data._ctx(proj)
    out = [np.zeros(ks.size) for _ in range(3)]
    check(lib().hbem_sigma_sweep_der2(g.h, _dptr(ks), ks.size, float(c_o), float(c_i), _vptr(R), R.shape[1], int(q),
                                      *[_dptr(o) for o in out]))
    return tuple(out)


def sigma_sweep_mixed(builder_data: BuilderData, ks, c_o, c_i, R, q=2, refine=2, refine_inner=1, proj=True):
    """sigma_min(k) for every k with the dense stage in mixed precision (FP32
    LU on the matrix cores; FP64 refinement: `refine` steps for the last
    solve of the randomized SVD, `refine_inner` for the others, -1 = as the
    last; the defaults 2 / 1 held sigma within 2e-10 of FP64 on the 2000-panel
    polygon's 128-k grid).  The assembly is FP64, with the lower-order rules
    on separated pairs when builder_data was made with FLAG_FAR_FIELD.
    Returns (sigma, fp64_used): fp64_used[i] is True where the last solve's
    refinement could not certify k (or an earlier solve's refinement did not
    contract) and it was recomputed with the FP64 factorisation."""
    ks = np.ascontiguousarray(ks, dtype=np.float64)
    if ks.ndim != 1:
        raise ValueError("ks must be a 1-D array of wavenumbers")
    R = _test_matrix(R, 2 * builder_data.getNumPanels())
    g = builder_data._ctx(proj)
    sig = np.zeros(ks.size)
    used = np.zeros(ks.size, dtype=np.int32)
    check(lib().hbem_sigma_sweep_mixed(g.h, _dptr(ks), ks.size, float(c_o), float(c_i), _vptr(R), R.shape[1], int(q),
                                       int(refine), int(refine_inner), _dptr(sig), used.ctypes.data_as(_ip)))
    return sig, used.astype(bool)


def sigma_sweep_batched(instances, q=2, proj=True, bucket=0, tune_cache=None):
    """sigma_min(k) of many scatterers as one batched computation
    (hbem_sigma_sweep_multi): instances are (builder_data, ks, c_o, c_i, R)
    tuples with k grids of one length, one quadrature order and one c_o.
    Instances whose orders 2N round up to the same multiple of `bucket` (0:
    256; "auto": the autotuned choice, see helmholtzbem_amd.autotune, cached
    in `tune_cache` or the default cache file) share one assembly launch per
    pair class and one batched LU and randomized SVD (T padded with the
    identity).  Returns the list of sigma arrays."""
    items = list(instances)
    if not items:
        return []
    if bucket == "auto":
        from . import autotune
        bucket = autotune.tune_bucket(items, q, proj, cache=tune_cache or autotune.DEFAULT_CACHE)
    nk = len(items[0][1])
    c_o = float(items[0][2])
    nc = np.asarray(items[0][4]).shape[1]
    for bd, ks, co, ci, R in items:
        if len(ks) != nk or float(co) != c_o or np.asarray(R).shape[1] != nc:
            raise ValueError("instances must share the k-grid length, c_o and the number of test columns")
    Rs = [_test_matrix(R, 2 * bd.getNumPanels()) for bd, _, _, _, R in items]
    geos = [bd._ctx(proj) for bd, *_ in items]
    hs = (ctypes.c_void_p * len(items))(*[g.h.value for g in geos])
    ks_all = np.ascontiguousarray(np.concatenate([np.asarray(ks, dtype=np.float64) for _, ks, _, _, _ in items]))
    ci = np.ascontiguousarray([float(ci) for _, _, _, ci, _ in items], dtype=np.float64)
    rp = (ctypes.c_void_p * len(items))(*[R.ctypes.data for R in Rs])
    sig = np.zeros(ks_all.size)
    check(lib().hbem_sigma_sweep_multi(hs, len(items), _dptr(ks_all), nk, c_o, _dptr(ci), rp, nc, int(q), int(bucket),
                                       _dptr(sig)))
    return [sig[i * nk:(i + 1) * nk].copy() for i in range(len(items))]


def sigma_sweep_multi(instances, q=2, der=0, proj=True, max_workers=None, bucket=0):
    """sigma_sweep for many independent scatterers at once -- the drivers'
    loop over instances of uneven size.  instances: (builder_data, ks, c_o,
    c_i, R) tuples.  Sigma-only sweeps with k grids of one length (>= 32,
    real positive), one quadrature order and one c_o run as one batched
    computation (sigma_sweep_batched), whose results are then the same bits
    as sigma_sweep's.  Otherwise every geometry context's sweep is
    issued from its own host thread (ctypes releases the GIL) onto its own
    stream, and the GPU overlaps them.  Returns the list of sigma_sweep
    results in the order of the instances."""
    from concurrent.futures import ThreadPoolExecutor
    items = list(instances)
    if not items:
        return []
    if der == 0:
        orders = {bd.getGaussQROrder() for bd, *_ in items}
        if len(orders) == 1 and len({len(it[1]) for it in items}) == 1 and len({float(it[2]) for it in items}) == 1 \
                and len({np.asarray(it[4]).shape[1] for it in items}) == 1 and len(items[0][1]) >= 32 \
                and all(np.all(np.asarray(it[1]) > 0) for it in items):
            return sigma_sweep_batched(items, q, proj, bucket)
    for bd, *_ in items:
        bd._ctx(proj)                          # contexts are created on this thread
    workers = max_workers or min(len(items), 8)
    with ThreadPoolExecutor(max_workers=workers) as ex:
        futs = [ex.submit(sigma_sweep, bd, ks, c_o, c_i, R, q, der, proj) for bd, ks, c_o, c_i, R in items]
        return [f.result() for f in futs]


def sweep_batch_size(npanels, order, d, nc, avail_bytes, nk=0):
    """Wavenumbers per batch of sigma_sweep for a closed curve of `npanels`
    panels when `avail_bytes` of device memory are usable (host-only)."""
    B = ctypes.c_int()
    check(lib().hbem_sweep_batch_size(int(npanels), int(order), int(d), int(nc), int(avail_bytes), int(nk),
                                      ctypes.byref(B)))
    return B.value


def profile_enable(builder_data: BuilderData, on=True, proj=True):
    check(lib().hbem_profile_enable(builder_data._ctx(proj).h, 1 if on else 0))


def profile_query(builder_data: BuilderData, which, proj=True):
    ms = ctypes.c_double()
    n = ctypes.c_longlong()
    fl = ctypes.c_double()
    check(lib().hbem_profile_query(builder_data._ctx(proj).h, int(which), ctypes.byref(ms), ctypes.byref(n),
                                   ctypes.byref(fl)))
    return ms.value, n.value, fl.value
