"""Pins the oracle's Galerkin assembly (galerkin_builder.cpp) without the reference.

* Local matrices of single panel pairs against the textbook Galerkin integrals,
  written here from the kernel definitions (G = (i/4) H0(kappa r), its normal
  derivative, Maue's form of the hypersingular operator and their k-derivatives
  via scipy's Hankel derivatives) and integrated by independent quadrature:
  tensor Gauss for separated pairs, an exact 1-D reduction + mpmath tanh-sinh
  for coinciding straight panels, Duffy splitting at the shared vertex with a
  dyadically graded radial Gauss rule for adjacent pairs.  Panel points and
  tangents are evaluated here from the panel parameters.
* Structural invariants: symmetry of V and W, circulant operators on an
  equal-arc circle (exercises both adjacency orientations), translation
  invariance, central differences in k for the derivative orders, the curl
  part of W annihilating constants, and the reference's V' coinciding quirk.

Local matrices exclude the factor i/4 (applied at the end of assembleDense).
"""
import math

import numpy as np
import pytest
import scipy.special as sp

K0, C0 = 2.0, 5.0


def _h0d(z, n):
    return sp.h1vp(0, z, n) if n else sp.hankel1(0, z)


def _h1d(z, n):
    return sp.h1vp(1, z, n) if n else sp.hankel1(1, z)


def kernels(x, y, ti, tj, s, t, k, c, d):
    """Integrands (times the quadrature-free Jacobians) of V, K (direct,
    transposed) and W for test point x = gamma_i(s), trial y = gamma_j(t).
    Returns dict of arrays [m][n]."""
    sq = math.sqrt(c)
    kap = k * sq
    v = x - y
    r = np.abs(v)
    b = lambda m, u: u if m == 0 else 1.0 - u
    db = (1.0, -1.0)
    z = kap * r
    # d^d/dk^d H0(k sqrt(c) r) = (sqrt(c) r)^d H0^(d)(kappa r)
    H = (sq * r) ** d * _h0d(z, d)
    # d^d/dk^d [kappa H1(kappa r)] = c^(d/2) [kappa r^d H1^(d) + d r^(d-1) H1^(d-1)]
    if d == 0:
        F = kap * _h1d(z, 0)
    else:
        F = c ** (d / 2) * (kap * r ** d * _h1d(z, d) + d * r ** (d - 1) * _h1d(z, d - 1))
    Nj = np.imag(tj) - 1j * np.real(tj)          # (Im tau_j, -Re tau_j) as complex
    Ni = np.imag(ti) - 1j * np.real(ti)
    vn_d = (np.real(v) * np.real(Nj) + np.imag(v) * np.imag(Nj)) / r * np.abs(ti)
    vn_t = (-np.real(v) * np.real(Ni) - np.imag(v) * np.imag(Ni)) / r * np.abs(tj)
    tdt = np.real(ti) * np.real(tj) + np.imag(ti) * np.imag(tj)
    h0, h0p, h0pp = _h0d(z, 0), _h0d(z, 1), _h0d(z, 2)
    out = {"V": np.zeros((2, 2), dtype=object), "Kd": np.zeros((2, 2), dtype=object),
           "Kt": np.zeros((2, 2), dtype=object), "W": np.zeros((2, 2), dtype=object)}
    for m in range(2):
        for n in range(2):
            out["V"][m, n] = b(m, s) * b(n, t) * np.abs(ti) * np.abs(tj) * H
            out["Kd"][m, n] = b(m, s) * b(n, t) * vn_d * F
            out["Kt"][m, n] = b(m, t) * b(n, s) * vn_t * F
            A = db[m] * db[n]
            B = tdt * b(m, s) * b(n, t)
            if d == 0:
                wv = h0 * (A - kap ** 2 * B)
            elif d == 1:
                wv = sq * (r * h0p * (A - kap ** 2 * B) - 2 * kap * B * h0)
            else:
                wv = c * (r ** 2 * h0pp * (A - kap ** 2 * B) - 4 * kap * r * B * h0p - 2 * B * h0)
            out["W"][m, n] = wv
    return out


def panel_fn(O, mesh, i):
    """Point and forward tangent of panel i, evaluated here from the panel
    parameters (independently of the oracle's panel code)."""
    p = mesh.packed[i]
    typ = int(p[0])
    if typ == 0:
        z1, z2 = p[1] + 1j * p[2], p[3] + 1j * p[4]
        return (lambda s: z1 + (z2 - z1) * np.asarray(s)), (lambda s: (z2 - z1) + 0 * np.asarray(s))
    if typ == 1:
        c, r, a, b = p[1] + 1j * p[2], p[3], p[4], p[5]
        return (lambda s: c + r * np.exp(1j * (a + (b - a) * np.asarray(s))),
                lambda s: 1j * r * (b - a) * np.exp(1j * (a + (b - a) * np.asarray(s))))
    if typ == 3:
        # spline piece: 9 doubles per cubic piece {t_j, x_j, y_j, bx, cx, dx, by, cy, dy}
        t0, t1, c0, npc = p[1], p[2], int(p[3]), int(p[4])
        rows = mesh.coefs[c0:c0 + 3 * npc + 1]
        knots = rows[0::3, 0]

        def piece(s):
            t = t0 + (t1 - t0) * np.asarray(s, dtype=float)
            j = np.clip(np.searchsorted(knots, t, side="right") - 1, 0, npc - 1)
            return t - rows[3 * j, 0], rows[3 * j], rows[3 * j + 1], rows[3 * j + 2]

        def zs(s):
            d, r0, r1, r2 = piece(s)
            return (r0[..., 1] + d * (r1[..., 0] + d * (r1[..., 1] + d * r1[..., 2]))
                    + 1j * (r0[..., 2] + d * (r2[..., 0] + d * (r2[..., 1] + d * r2[..., 2]))))

        def dzs(s):
            d, r0, r1, r2 = piece(s)
            return (t1 - t0) * ((r1[..., 0] + d * (2 * r1[..., 1] + 3 * d * r1[..., 2]))
                                + 1j * (r2[..., 0] + d * (2 * r2[..., 1] + 3 * d * r2[..., 2])))
        return zs, dzs
    t0, t1, c0, nc = p[1], p[2], int(p[3]), int(p[4])
    co = mesh.coefs[c0:c0 + nc]

    def z(s):
        t = t0 + (t1 - t0) * np.asarray(s)
        return sum((cm[1] + 1j * cm[2]) * np.exp(1j * cm[0] * t) for cm in co)

    def dz(s):
        t = t0 + (t1 - t0) * np.asarray(s)
        return (t1 - t0) * sum(1j * cm[0] * (cm[1] + 1j * cm[2]) * np.exp(1j * cm[0] * t) for cm in co)
    return z, dz


def oracle_local(O, mesh, order, i, j, d, k=K0):
    L = O.pair_local(mesh, order, k, C0, i, j, d=d)
    return {"V": L["V"][d], "W": L["W"][d], "Kd": L["K"][d, 0], "Kt": L["K"][d, 1]}


def _meshes(H):
    return {"square": H.square_mesh(1.0, 2), "circle": H.circle_mesh(8, 0.6, (0.1, -0.2)), "star": H.star_mesh(12)}


def _general_vs_textbook(O, H, name, i, j, d, k):
    mesh = _meshes(H)[name]
    assert O.classify(mesh)[i * (i + 1) // 2 + j] == 3
    xg, wg = np.polynomial.legendre.leggauss(48)
    xg, wg = 0.5 * (xg + 1), 0.5 * wg
    S, T = np.meshgrid(xg, xg, indexing="ij")
    W2 = np.outer(wg, wg)
    pi_, ti_ = panel_fn(O, mesh, i)
    pj_, tj_ = panel_fn(O, mesh, j)
    x = pi_(S.ravel()).reshape(S.shape)
    y = pj_(T.ravel()).reshape(T.shape)
    ti = ti_(S.ravel()).reshape(S.shape)
    tj = tj_(T.ravel()).reshape(T.shape)
    ker = kernels(x, y, ti, tj, S, T, k, C0, d)
    got = oracle_local(O, mesh, 20, i, j, d, k)
    for key in ("V", "Kd", "Kt", "W"):
        ref = np.array([[np.sum(W2 * ker[key][m, n]) for n in range(2)] for m in range(2)])
        scale = np.abs(ref).max()
        assert np.abs(got[key] - ref).max() <= 1e-11 * scale, (key, got[key], ref)


@pytest.mark.parametrize("name,i,j", [("square", 5, 0), ("square", 4, 1), ("circle", 4, 0), ("circle", 6, 1),
                                      ("star", 6, 0), ("star", 9, 2)])
@pytest.mark.parametrize("d", [0, 1, 2])
def test_general_pair_vs_textbook_integral(O, H, name, i, j, d):
    _general_vs_textbook(O, H, name, i, j, d, K0)


# Complex wavenumbers (galerkin_builder.cpp:224-235 path): Im k > 0, a negative
# real k (H of the negative real axis, principal branch) and a large imaginary
# part; the integrands use scipy's complex Hankel functions.
@pytest.mark.parametrize("name,i,j", [("circle", 4, 0), ("star", 6, 0), ("square", 5, 0)])
@pytest.mark.parametrize("k", [2.0 + 0.7j, -1.5 + 0.0j, 0.3 + 2.0j])
@pytest.mark.parametrize("d", [0, 1, 2])
def test_general_pair_complex_k_vs_textbook_integral(O, H, name, i, j, k, d):
    _general_vs_textbook(O, H, name, i, j, d, k)


def _coinciding_exact(L, kap, c, d, layer, m, n):
    """Exact Galerkin integral over a straight panel of length L (s, t in [0,1])
    by reduction to u = |s - t| (inner polynomial integral exact, outer by
    tanh-sinh in mpmath)."""
    import mpmath as mp
    mp.mp.dps = 20
    b = [lambda u: u, lambda u: 1 - u]
    db = (1.0, -1.0)
    gx, gw = np.polynomial.legendre.leggauss(6)

    def inner(u, fun):
        # int_0^{1-u} [fun(t+u, t) + fun(t, t+u)] dt, exact for the cubic integrands
        a = 1.0 - u
        ts = 0.5 * a * (gx + 1)
        return 0.5 * a * float(np.sum(gw * (fun(ts + u, ts) + fun(ts, ts + u))))

    sq = math.sqrt(c)

    def H0d(z, k):
        h = [mp.hankel1(0, z), -mp.hankel1(1, z)]
        if k <= 1:
            return h[k]
        return -(mp.hankel1(0, z) - mp.hankel1(1, z) / z)   # H0'' = -H1' = -(H0 - H1/z)

    def integrand(u):
        u = float(u)
        r = L * u
        z = kap * r
        if layer == "V":
            pm = inner(u, lambda s, t: b[m](s) * b[n](t)) * L * L
            return pm * (sq * r) ** d * H0d(z, d)
        A = inner(u, lambda s, t: db[m] * db[n] + 0 * s)
        B = inner(u, lambda s, t: b[m](s) * b[n](t)) * L * L
        h0, h0p, h0pp = H0d(z, 0), H0d(z, 1), H0d(z, 2)
        if d == 0:
            return h0 * (A - kap ** 2 * B)
        if d == 1:
            return sq * (r * h0p * (A - kap ** 2 * B) - 2 * kap * B * h0)
        return c * (r ** 2 * h0pp * (A - kap ** 2 * B) - 4 * kap * r * B * h0p - 2 * B * h0)

    val = mp.quad(integrand, [0, 0.25, 1])
    return complex(val)


@pytest.mark.parametrize("d,k", [(0, K0), (1, K0), (2, K0), (0, 1.5 + 0.5j), (2, 1.5 + 0.5j)])
def test_coinciding_straight_panel_exact(O, H, d, k):
    mesh = H.square_mesh(1.3, 3)
    i = 4
    p = mesh.packed[i]
    Lp = math.hypot(p[3] - p[1], p[4] - p[2])
    kap = k * math.sqrt(C0)
    # The composite rule omits [0, sigma^(n-1)] (gauleg.hpp:94-119): at order 16
    # that truncation is sigma^15 ~ 3e-12 of the integral, so 1e-9 pins the formulas.
    got = oracle_local(O, mesh, 16, i, i, d, k)
    assert np.all(got["Kd"] == 0) and np.all(got["Kt"] == 0)     # straight panel: v parallel to tangent
    for layer in ("V", "W"):
        ref = np.array([[_coinciding_exact(Lp, kap, C0, d, layer, m, n) for n in range(2)] for m in range(2)])
        assert np.abs(got[layer] - ref).max() <= 1e-9 * np.abs(ref).max(), (layer, got[layer], ref)


def _graded_rule():
    """Composite Gauss rule on (0, 1] graded geometrically toward 0 (the radial
    Duffy variable carries rho log rho): 10 points on each [2^-(k+1), 2^-k]."""
    gx, gw = np.polynomial.legendre.leggauss(10)
    xs, ws = [], []
    for k in range(48):
        lo, hi = 2.0 ** -(k + 1), 2.0 ** -k
        xs.append(lo + (hi - lo) * 0.5 * (gx + 1))
        ws.append((hi - lo) * 0.5 * gw)
    return np.concatenate(xs), np.concatenate(ws)


def _adjacent_exact(O, mesh, i, j, d, k=K0):
    """All local entries of an adjacent pair by Duffy splitting at the shared
    vertex: (sigma_i, sigma_j) = (rho u, rho) and (rho, rho u), Gauss-Legendre
    in u, the graded rule in rho.  sigma = parameter distance from the vertex."""
    cls = O.classify(mesh)[i * (i + 1) // 2 + j]
    swap = cls == 2
    pi_, ti_ = panel_fn(O, mesh, i)
    pj_, tj_ = panel_fn(O, mesh, j)
    ux, uw = np.polynomial.legendre.leggauss(40)
    ux, uw = 0.5 * (ux + 1), 0.5 * uw
    rx, rw = _graded_rule()
    U, RHO = np.meshgrid(ux, rx, indexing="ij")
    WW = np.outer(uw, rw) * RHO
    out = {}
    for tri in range(2):
        si_, sj_ = (RHO * U, RHO) if tri == 0 else (RHO, RHO * U)
        s = si_ if swap else 1.0 - si_
        t = 1.0 - sj_ if swap else sj_
        ker = kernels(pi_(s), pj_(t), ti_(s), tj_(t), s, t, k, C0, d)
        for key in ("V", "W", "Kd", "Kt"):
            val = np.array([[np.sum(WW * ker[key][m, n]) for n in range(2)] for m in range(2)])
            out[key] = out.get(key, 0) + val
    return out


@pytest.mark.parametrize("name,i,j,k", [("square", 2, 1, K0), ("square", 7, 0, K0), ("square", 1, 0, K0),
                                        ("circle", 3, 2, K0), ("circle", 7, 0, K0), ("circle", 3, 2, 2.0 + 0.7j),
                                        ("square", 2, 1, -1.2 + 0.4j)])
def test_adjacent_pair_vs_duffy_quadpack(O, H, name, i, j, k):
    mesh = _meshes(H)[name]
    d = 1
    got = oracle_local(O, mesh, 12, i, j, d, k)
    exact = _adjacent_exact(O, mesh, i, j, d, k)
    for key in ("V", "W", "Kd", "Kt"):
        if np.all(got[key] == 0):
            # skipped by the reference's all-|v.n| < eps test: collinear panels
            assert name == "square" and (i, j) == (1, 0) and key in ("Kd", "Kt")
            continue
        ref = exact[key]
        assert np.abs(got[key] - ref).max() <= 1e-7 * np.abs(ref).max(), (key, got[key], ref)


def test_symmetry_and_translation(O, H):
    m1 = H.square_mesh(0.5, 5)
    m2 = H.square_mesh(0.5, 5, center=(0.25, 0.25))
    K1, W1, V1 = O.assemble(m1, 8, 3.0, 4.0, d=2)
    K2, W2, V2 = O.assemble(m2, 8, 3.0, 4.0, d=2)
    for t in range(3):
        assert np.array_equal(V1[t], V1[t].T) and np.array_equal(W1[t], W1[t].T)
        for A, B in ((K1, K2), (W1, W2), (V1, V2)):
            assert np.abs(A[t] - B[t]).max() <= 1e-12 * np.abs(A[t]).max()


def test_circle_operators_circulant(O, H):
    """Equal arcs: every operator is circulant.  Interior adjacent pairs use
    swap=true, the wrap pair (N-1, 0) swap=false -- both paths must agree."""
    N = 10
    mesh = H.circle_mesh(N, 0.8)
    K, W, V = O.assemble(mesh, 8, 2.5, 3.0, d=2)
    for A in (K, W, V):
        for t in range(3):
            X = A[t]
            shifted = np.roll(np.roll(X, 1, axis=0), 1, axis=1)
            assert np.abs(shifted - X).max() <= 1e-12 * np.abs(X).max()


@pytest.mark.parametrize("which", ["square", "circle", "star"])
def test_derivatives_match_central_differences(O, H, which):
    mesh = {"square": H.square_mesh(0.6, 3), "circle": H.circle_mesh(9, 0.5), "star": H.star_mesh(15)}[which]
    k, c, h = 2.2, 6.0, 1e-3
    A0 = O.assemble(mesh, 8, k, c, d=2)
    Ap = O.assemble(mesh, 8, k + h, c, d=0)
    Am = O.assemble(mesh, 8, k - h, c, d=0)
    for a0, ap, am in zip(A0, Ap, Am):
        fd1 = (ap[0] - am[0]) / (2 * h)
        fd2 = (ap[0] - 2 * a0[0] + am[0]) / h ** 2
        assert np.abs(fd1 - a0[1]).max() <= 1e-5 * np.abs(a0[1]).max()
        assert np.abs(fd2 - a0[2]).max() <= 1e-4 * np.abs(a0[2]).max()


def test_hypersingular_curl_part_annihilates_constants(O, H):
    mesh = H.circle_mesh(12, 1.0)
    K, W, V = O.assemble(mesh, 8, 1e-3, 1.0, d=0)
    ones = np.ones(12)
    # W 1 = -kappa^2 (normal-normal term) = O(kappa^2 log kappa)
    assert np.abs(W[0] @ ones).max() <= 1e-4 * np.abs(W[0]).max()


def test_single_layer_coinciding_derivative_quirk(O, H):
    """galerkin_builder.cpp:744-746 copies the d=0 entry into the d>=1 (0,1)
    entries of the coinciding single-layer block; the default path uses the
    entry of the same order (what d/dk V requires)."""
    mesh = H.star_mesh(9)
    N = 9
    k, c, h = 2.0, 4.0, 1e-4
    _, _, V = O.assemble(mesh, 8, k, c, layers=O.LAYER_SINGLE, d=2)
    _, _, Vq = O.assemble(mesh, 8, k, c, layers=O.LAYER_SINGLE, d=2, flags=O.QUIRK_SLP_COINC_DER)
    assert np.array_equal(V[0], Vq[0])
    for t in (1, 2):
        diff = np.abs(V[t] - Vq[t])
        mask = np.zeros((N, N), bool)
        for i in range(N):
            mask[(i + 1) % N, i] = True
        assert np.all(diff[~mask] == 0) and np.all(diff[mask] > 0)
    _, _, Vp = O.assemble(mesh, 8, k + h, c, layers=O.LAYER_SINGLE, d=0)
    _, _, Vm = O.assemble(mesh, 8, k - h, c, layers=O.LAYER_SINGLE, d=0)
    fd = (Vp[0] - Vm[0]) / (2 * h)
    assert np.abs(fd - V[1]).max() <= 1e-6 * np.abs(V[1]).max()
    assert np.abs(fd - Vq[1]).max() > 1e-3 * np.abs(V[1]).max()


def test_reference_error_messages(O, H):
    mesh = H.square_mesh(0.5, 2)
    with pytest.raises(RuntimeError, match="Refraction index must not be smaller than 1"):
        O.assemble(mesh, 4, 1.0, 0.5)
    with pytest.raises(RuntimeError, match="Invalid order of derivative"):
        O.assemble(mesh, 4, 1.0, 2.0, d=3)
