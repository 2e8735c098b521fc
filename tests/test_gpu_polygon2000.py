"""The headline workload at full size: a random star-shaped 12-vertex polygon
(seed 3) in 2000 line panels, Gauss order 11, c_i = 20, c_o = 1 -- a
4000 x 4000 transmission matrix -- on the 128-point grid k = linspace(5, 40).

At this size the oracle's per-k dense work takes minutes, so the full-size
checks are size-independent properties and independent implementations:
exact V / W symmetry, the batched sweep against the per-k API path
(gen_sol_op + randomized_svd.sv) at 1e-10, sigma against a numpy/LAPACK
restatement of the randomized SVD (the same R and q) at 1e-10, sigma as an
upper bound of the converged sigma_min, and bitwise repeatability across
batch splits.  Matrix and sigma parity with the oracle run on the same
generator at 500 panels (T at 1e-12 relative Frobenius, sigma at 1e-10).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

N, ORDER, C_I, C_O = 2000, 11, 20.0, 1.0
KS = np.linspace(5.0, 40.0, 128)
MAT_TOL = 1e-12
SIG_TOL = 1e-10


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def numpy_rsvd(T, R, q):
    """randomized_svd::sv (randsvd.cpp:99-146) with LAPACK solves and QR:
    sigma depends on R, q and T only through the spans of the iterates."""
    Q = np.linalg.qr(np.linalg.solve(T, R))[0]
    Th = T.conj().T
    for _ in range(q):
        Q = np.linalg.qr(np.linalg.solve(Th, Q))[0]
        Q = np.linalg.qr(np.linalg.solve(T, Q))[0]
    return 1.0 / np.linalg.svd(np.linalg.solve(Th, Q), compute_uv=False)[0]


@pytest.fixture(scope="module")
def poly(H):
    mesh = H.random_polygon_mesh(3, 12, N)
    sp = H.ContinuousSpace1()
    return mesh, H.BuilderData(mesh, sp, sp, ORDER)


def test_polygon2000_mesh(poly):
    mesh, bd = poly
    assert mesh.getNumPanels() == N and mesh.isPolygonal()
    P = mesh.packed
    # consecutive panels meet (to rounding at the polygon's vertices, where the
    # reference's adjacency test allows 100 eps)
    assert np.abs(P[1:, 1:3] - P[:-1, 3:5]).max() <= 1e-14 and np.abs(P[0, 1:3] - P[-1, 3:5]).max() <= 1e-14


def test_polygon2000_symmetry(poly, H):
    mesh, bd = poly
    gb = H.GalerkinBuilder(bd)
    gb.assembleDense(22.5, C_I, H.LayerType.SINGLE | H.LayerType.HYPERSINGULAR, 0)
    V, W = gb.getSingleLayer(0), gb.getHypersingular(0)
    assert np.array_equal(V, V.T) and np.array_equal(W, W.T)
    assert np.all(np.isfinite(V)) and np.all(np.isfinite(W))


def test_polygon2000_sweep_vs_api_and_lapack(poly, H):
    mesh, bd = poly
    R = H.gaussian_test_matrix(2 * N, 2, 3)
    idx = [0, 61, 127]
    ks = KS[idx]
    sig = H.sigma_sweep(bd, ks, C_O, C_I, R, 2)
    so = H.SolutionsOperator(bd)
    for i, k in enumerate(ks):
        T = so.gen_sol_op(k, C_O, C_I)
        s_api = H.randomized_svd.sv(T, R, 2)
        assert abs(s_api - sig[i]) <= SIG_TOL * sig[i], (k, s_api, sig[i])
        s_conv = H.randomized_svd.sv(T, R, 30)
        assert sig[i] >= s_conv * (1 - 1e-10), (k, sig[i], s_conv)
        if i == 1:
            s_np = numpy_rsvd(T, R, 2)
            assert abs(s_np - sig[i]) <= SIG_TOL * s_np, (k, s_np, sig[i])


def test_polygon2000_grid_repeatable_across_batches(poly, H):
    mesh, bd = poly
    R = H.gaussian_test_matrix(2 * N, 2, 3)
    full = H.sigma_sweep(bd, KS, C_O, C_I, R, 2)
    assert np.all(np.isfinite(full)) and np.all(full > 0)
    again = H.sigma_sweep(bd, KS, C_O, C_I, R, 2)
    assert np.array_equal(full, again)
    # batches of 32 k (64 wavenumber slots) still run on the paired-lane kernel
    parts = np.concatenate([H.sigma_sweep(bd, KS[a:a + 32], C_O, C_I, R, 2) for a in range(0, 128, 32)])
    assert np.array_equal(full, parts)


def test_polygon500_oracle_parity(H, O):
    """The same generator at 500 panels: unprojected T and sigma (projected and
    not) against the oracle."""
    mesh = H.random_polygon_mesh(3, 12, 500)
    sp = H.ContinuousSpace1()
    bd = H.BuilderData(mesh, sp, sp, ORDER)
    R = H.gaussian_test_matrix(1000, 2, 3)
    so = H.SolutionsOperator(bd, proj=False)
    ref = O.SolOp(mesh, ORDER, proj=False)
    for k in (5.0, 40.0):
        T = so.gen_sol_op(k, C_O, C_I)
        To = ref.gen_sol_op(k, C_O, C_I)[0]
        assert rel(T, To) <= MAT_TOL
    ks = np.array([5.0, 22.5, 40.0])
    sig = H.sigma_sweep(bd, ks, C_O, C_I, R, 2, proj=False)
    sig_o = ref.sigma_sweep(ks, C_O, C_I, R, 2)
    assert np.all(np.abs(sig - sig_o) <= SIG_TOL * sig_o), (sig, sig_o)
    sp_ = H.sigma_sweep(bd, ks[1:2], C_O, C_I, R, 2)
    sp_o = O.SolOp(mesh, ORDER).sigma_sweep(ks[1:2], C_O, C_I, R, 2)
    assert abs(sp_[0] - sp_o[0]) <= SIG_TOL * sp_o[0], (sp_, sp_o)
