"""Pins the oracle's mass matrix, projection factor and randomized SVD
(mass_matrix.cpp, gen_sol_op.cpp:23-74, randsvd.cpp) with closed forms,
numpy/LAPACK and finite differences."""
import math

import numpy as np
import pytest


def test_mass_matrix_closed_form_lines(O, H):
    mesh = H.random_polygon_mesh(7, 7, 23)
    N = mesh.getNumPanels()
    M = O.mass_matrix(mesh, 11)
    L = np.array([math.hypot(p[3] - p[1], p[4] - p[2]) for p in mesh.packed])
    ref = np.zeros((N, N))
    for p in range(N):
        a, b = (p + 1) % N, p          # dof of b0 = t (end vertex), b1 = 1 - t (start vertex)
        ref[a, a] += L[p] / 3
        ref[b, b] += L[p] / 3
        ref[a, b] += L[p] / 6
        ref[b, a] += L[p] / 6
    assert np.abs(M - ref).max() <= 4e-16 * L.max()


def test_mass_matrix_arcs(O, H):
    mesh = H.circle_mesh(11, 0.7)
    M = O.mass_matrix(mesh, 9)
    L = 0.7 * 2 * math.pi / 11
    assert np.allclose(np.diag(M), 2 * L / 3, rtol=0, atol=1e-15)
    assert abs(M[1, 0] - L / 6) < 1e-15 and abs(M[0, 10] - L / 6) < 1e-15


def test_ldlt_factor_properties(O):
    rng = np.random.default_rng(3)
    n = 24
    A = np.zeros((n, n))
    for i in range(n):
        A[i, i] = 1.0 + rng.random()
        A[i, (i + 1) % n] = A[(i + 1) % n, i] = 0.2 * rng.random()
    F, tr = O.ldlt_factor(A)
    assert np.abs(F @ F.T - A).max() < 1e-14
    # left-looking diagonal pivoting: step k takes the largest remaining original diagonal
    diag = list(np.diag(A))
    for k in range(n):
        rest = diag[k:]
        assert tr[k] == k + int(np.argmax(np.abs(rest)))
        diag[k], diag[tr[k]] = diag[tr[k]], diag[k]
    # at most two off-diagonal entries per column of L (cycle graph elimination)
    assert np.max(np.sum(np.abs(F) > 0, axis=0)) <= 3


def test_ldlt_ties_keep_order(O):
    # exactly equal diagonal (uniform P1 mass matrix): first index wins -> identity order
    n = 16
    A = np.diag(np.full(n, 2.0 / 3.0)) + np.diag(np.full(n - 1, 1.0 / 6.0), 1) + np.diag(np.full(n - 1, 1.0 / 6.0), -1)
    A[0, n - 1] = A[n - 1, 0] = 1.0 / 6.0
    F, tr = O.ldlt_factor(A)
    assert np.array_equal(tr, np.arange(n))
    assert np.abs(F @ F.T - A).max() < 1e-15


def test_ldlt_pivot_rule_on_mass_matrix_bits(O, H):
    """On a real mesh the diagonal ties are broken by the last bits of M; the
    pivot order must follow the selection rule on those exact values."""
    mesh = H.circle_mesh(10, 1.0)
    S = O.SolOp(mesh, 6)
    F, tr = S.factor()
    M = S.mass_matrix()
    diag = list(np.concatenate([np.diag(M), np.diag(M)]))
    for k in range(len(diag)):
        assert tr[k] == k + int(np.argmax(np.abs(diag[k:])))
        diag[k], diag[tr[k]] = diag[tr[k]], diag[k]
    A = np.block([[M, np.zeros_like(M)], [np.zeros_like(M), M.T]])
    assert np.abs(F @ F.T - A).max() < 1e-15


def test_projection_definition(O, H):
    mesh = H.square_mesh(0.5, 3)
    S = O.SolOp(mesh, 6)
    T = S.gen_sol_op(2.0, 1.0, 20.0, d=0)[0]
    Su = O.SolOp(mesh, 6, proj=False)
    T0 = Su.gen_sol_op(2.0, 1.0, 20.0, d=0)[0]
    F, _ = S.factor()
    ref = np.linalg.solve(F, np.linalg.solve(F, T0.T).T)          # F^-1 T F^-T
    assert np.abs(T - ref).max() <= 1e-13 * np.abs(ref).max()
    assert np.allclose(np.linalg.svd(T, compute_uv=False), np.linalg.svd(ref, compute_uv=False), rtol=1e-12)


def _numpy_rsvd(T, R, q):
    """Randomized subspace iteration with numpy's solve/QR/SVD (randsvd.cpp:99-146)."""
    Q = np.linalg.qr(np.linalg.solve(T, R))[0]
    for _ in range(q):
        Q = np.linalg.qr(np.linalg.solve(T.conj().T, Q))[0]
        Q = np.linalg.qr(np.linalg.solve(T, Q))[0]
    return 1.0 / np.linalg.svd(np.linalg.solve(T.conj().T, Q), compute_uv=False)[0]


@pytest.fixture(scope="module")
def small_T(O, H):
    mesh = H.star_mesh(14)
    S = O.SolOp(mesh, 6)
    return S, S.gen_sol_op(3.0, 1.0, 10.0, d=2)


def test_rsvd_against_numpy(O, H, small_T):
    _, (T, Td, Tdd) = small_T
    n = T.shape[0]
    smin = np.linalg.svd(T, compute_uv=False)[-1]
    for q, nc in ((0, 1), (2, 2), (5, 4)):
        R = H.gaussian_test_matrix(n, nc, 11 + q)
        s = O.sv(T, R, q)
        assert abs(s - _numpy_rsvd(T, R, q)) <= 1e-12 * s
        assert s >= smin * (1 - 1e-13)
    R = H.gaussian_test_matrix(n, 2, 5)
    assert abs(O.sv(T, R, 40) - smin) <= 1e-12 * smin


def test_svd_values(O):
    rng = np.random.default_rng(1)
    B = rng.standard_normal((50, 4)) + 1j * rng.standard_normal((50, 4))
    assert np.allclose(O.svd_values(B), np.linalg.svd(B, compute_uv=False), rtol=1e-14)


def test_sv_derivatives_vs_finite_differences(O, H):
    """sv_der / sv_der2 (randsvd.cpp:148-216) against central differences of the
    exact sigma_min(T(k)) at converged subspace iteration."""
    mesh = H.circle_mesh(12, 0.6)
    S = O.SolOp(mesh, 6)
    k, h = 2.7, 1e-3
    T, Td, Tdd = S.gen_sol_op(k, 1.0, 9.0, d=2)
    R = H.gaussian_test_matrix(T.shape[0], 2, 3)
    s = lambda kk: np.linalg.svd(S.gen_sol_op(kk, 1.0, 9.0, d=0)[0], compute_uv=False)[-1]
    sp, s0, sm = s(k + h), s(k), s(k - h)
    d1 = O.sv_der(T, Td, R, 40)
    d2 = O.sv_der2(T, Td, Tdd, R, 40)
    assert abs(d1[0] - s0) <= 1e-12 * s0 and abs(d2[0] - s0) <= 1e-12 * s0
    assert abs(d1[1] - (sp - sm) / (2 * h)) <= 1e-5 * max(1.0, abs(d1[1]))
    assert abs(d2[1] - d1[1]) <= 1e-10 * max(1.0, abs(d1[1]))
    assert abs(d2[2] - (sp - 2 * s0 + sm) / h ** 2) <= 1e-3 * max(1.0, abs(d2[2]))
