"""Resonance search drivers (find_roots.cpp:19-215, 1426-1512, 1615-1726,
examples/roots_*_rsvd.cpp, find_resonances.cpp): zbrent and rtsafe against
scipy's brentq on analytic functions, their edge cases, lockstep execution
against individual runs, the C++ facade header against the Python host code
iterate for iterate; Brent's minimiser against known minima and scipy's
Brent, the Steffen spline's defining properties, opt_subdiv against the
minima of a known function; and (GPU) the batched minimum search against the
oracle's dsigma/dk, with the three refinement methods agreeing.

GSL (the reference's gsl_min_fminimizer_brent and gsl_interp_steffen, GSL 2.x,
not vendored) is absent here: Brent's minimiser and Steffen's interpolation
are restated from their publications, so their iterate-for-iterate parity with
GSL is unpinned; the tests pin what the methods fix."""
import math
import os
import subprocess

import numpy as np
import pytest
import scipy.optimize as so

from helmholtzbem_amd import resonances as RZ

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

CASES = [
    (lambda x: math.cos(x) - x, lambda x: -math.sin(x) - 1.0, 0.0, 1.0, 1e-12),
    (lambda x: x ** 3 - 2.0 * x - 5.0, lambda x: 3.0 * x * x - 2.0, 2.0, 3.0, 1e-14),
    (lambda x: math.tanh(3.0 * (x - 0.3)), lambda x: 3.0 * (1.0 - math.tanh(3.0 * (x - 0.3)) ** 2), -4.0, 5.0, 1e-10),
    (lambda x: math.exp(x) - 10.0, lambda x: math.exp(x), 0.0, 4.0, 1e-9),
]


@pytest.mark.parametrize("case", range(len(CASES)))
def test_zbrent_and_rtsafe_vs_brentq(case):
    f, df, a, b, tol = CASES[case]
    ref = so.brentq(f, a, b, xtol=1e-15, rtol=4 * np.finfo(float).eps)
    r, ok, it = RZ.zbrent(f, a, b, tol)
    assert ok and abs(r - ref) <= tol + 4e-16 * abs(ref) and 0 < it < RZ.MAXIT
    r2, ok2, it2 = RZ.rtsafe(f, lambda x: (f(x), df(x)), a, b, tol)
    assert ok2 and abs(r2 - ref) <= 10 * tol + 4e-16 * abs(ref)


def test_edge_cases():
    f = lambda x: x * x + 1.0
    assert RZ.zbrent(f, -1.0, 2.0, 1e-9) == (0.0, False, 0)          # no sign change
    assert RZ.rtsafe(f, lambda x: (f(x), 2 * x), -1.0, 2.0, 1e-9) == (2.0, False, 0)
    g = lambda x: x - 1.0
    assert RZ.rtsafe(g, lambda x: (g(x), 1.0), 1.0, 3.0, 1e-9) == (1.0, True, 0)
    assert RZ.zbrent(g, 0.0, 1.0, 1e-9)[:2] == (1.0, True)


def test_lockstep_equals_individual_runs():
    fs = [c[0] for c in CASES]

    def evaluate2(xs):
        return [fs[i](x) for i, x in zip(active, xs)]
    # generators with different functions, advanced round by round
    results = []
    gens = [RZ.zbrent_iter(c[2], c[3], c[4]) for c in CASES]
    pend = {i: next(g) for i, g in enumerate(gens)}
    while pend:
        active = list(pend)
        answers = evaluate2([pend[i] for i in active])
        for i, a in zip(active, answers):
            try:
                pend[i] = gens[i].send(a)
            except StopIteration as s:
                results.append((i, s.value))
                del pend[i]
    for i, res in results:
        c = CASES[i]
        assert res == RZ.zbrent(c[0], c[2], c[3], c[4])
    # and through RZ.lockstep with a vectorised evaluator over one function
    h = lambda x: math.sin(x)
    brackets = [(3.0, 3.5), (6.0, 6.5), (9.0, 9.9)]
    out = RZ.lockstep([RZ.zbrent_iter(a, b, 1e-12) for a, b in brackets], lambda xs: [h(x) for x in xs])
    for (a, b), r in zip(brackets, out):
        assert r == RZ.zbrent(h, a, b, 1e-12)


def test_bracket_minima():
    ks = np.linspace(0, 10, 101)
    sig = np.abs(np.sin(ks)) + 0.01
    br = RZ.bracket_minima(ks, sig)
    assert len(br) == 3
    for (a, b), m in zip(br, (np.pi, 2 * np.pi, 3 * np.pi)):
        assert a < m < b and abs((b - a) - 0.2) < 1e-12


def test_cpp_header_matches_python(tmp_path):
    exe = str(tmp_path / "roots_dump")
    subprocess.run(["g++", "-std=c++17", "-O2", "-ffp-contract=off", "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "cpp", "roots_dump.cpp"), "-o", exe],
                   check=True, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)
    lines = subprocess.run([exe], check=True, stdout=subprocess.PIPE, text=True).stdout.split("\n")
    cases = CASES + [(lambda x: x * x + 1.0, lambda x: 2.0 * x, -1.0, 2.0, 1e-9)]
    for k, (f, df, a, b, tol) in enumerate(cases):
        zr, zok, zit = lines[2 * k].split()
        tr, tok, tit = lines[2 * k + 1].split()
        r, ok, it = RZ.zbrent(f, a, b, tol)
        assert (float(zr), bool(int(zok)), int(zit)) == (r, ok, it), k
        r, ok, it = RZ.rtsafe(f, lambda x: (f(x), df(x)), a, b, tol)
        assert float(tr) == r and bool(int(tok)) == ok, k
        if ok:
            assert int(tit) == it
    off = 2 * len(cases)
    for k, (f, a, b, g, _) in enumerate(MIN_CASES):
        m, ic = lines[off + k].split()
        assert (float(m), int(ic)) == RZ.brent_gsl(a, b, g, f, 1e-7), k
    r, ok, it = lines[off + len(MIN_CASES)].split()
    f, both = nan_start_case()
    assert (float(r), bool(int(ok)), int(it)) == RZ.rtsafe(f, both, 1.0, 2.0, 1e-13)


def nan_start_case():
    """f(1) is NaN, so rtsafe asks f(NaN) for its first guess (1.25)."""
    def f(x):
        if x != x:
            return 1.25
        return math.nan if x == 1.0 else math.sin(x) - 0.9
    return f, lambda x: (f(x), math.cos(x))


def test_rtsafe_nan_start():
    """find_roots.cpp:159-162: with a NaN end value the first Newton iterate is
    fct(NaN) and the bracket is [x1, x2] as given."""
    f, both = nan_start_case()
    seen = []

    def fct(x):
        seen.append(x)
        return f(x)
    r, ok, it = RZ.rtsafe(fct, both, 1.0, 2.0, 1e-13)
    assert ok and abs(r - math.asin(0.9)) <= 1e-13
    assert seen[0] == 1.0 and seen[1] == 2.0 and seen[2] != seen[2]


@pytest.mark.gpu
def test_find_minima_vs_oracle(H, O):
    """The batched resonance search on a dielectric disk: every minimum found
    (Brent on dsigma/dk) has a vanishing oracle dsigma/dk at converged subspace
    iteration, and the Newton variant finds the same minima."""
    mesh = H.circle_mesh(16, 0.7, (0.0, 0.0))
    sp = H.ContinuousSpace1()
    bd = H.BuilderData(mesh, sp, sp, 8)
    R = H.gaussian_test_matrix(32, 2, 0)
    found = RZ.find_minima(bd, 2.0, 6.0, 41, 1.0, 9.0, R, q=2, q_refine=30, tol=1e-11)
    assert len(found) >= 1
    S = O.SolOp(mesh, 8)
    ks = np.linspace(2.0, 6.0, 41)
    scale = np.abs(H.sigma_sweep(bd, ks, 1.0, 9.0, R, 30, der=1)[1]).max()
    for k, s, it in found:
        T, Td = S.gen_sol_op(k, 1.0, 9.0, d=1)
        d = O.sv_der(T, Td, R, 40)
        assert abs(d[1]) <= 1e-6 * scale, (k, d, scale)
        assert abs(s - d[0]) <= 1e-8 * s
    newton = RZ.find_minima(bd, 2.0, 6.0, 41, 1.0, 9.0, R, q=2, q_refine=30, tol=1e-11, method="newton")
    assert len(newton) == len(found)
    for (k1, _, _), (k2, _, _) in zip(found, newton):
        assert abs(k1 - k2) <= 1e-7


MIN_CASES = [
    (lambda x: (x - 2.0) ** 2 + 1.0, 1.0, 3.5, 2.4, 2.0),
    (lambda x: math.cos(2.0 * x), 0.0, 2.0, 1.5, math.pi / 2),
    (lambda x: abs(x - 0.1234) + 1.0, -1.0, 1.0, 0.3, 0.1234),          # not smooth: golden steps
    (lambda x: math.exp(x) - 3.0 * x, 0.0, 3.0, 1.0, math.log(3.0)),
]


@pytest.mark.parametrize("case", range(len(MIN_CASES)))
def test_brent_minimiser_finds_known_minima(case):
    f, a, b, g, xstar = MIN_CASES[case]
    calls = []

    def fc(x):
        calls.append(x)
        return f(x)
    m, ic = RZ.brent_gsl(a, b, g, fc, 1e-7)
    assert abs(m - xstar) <= 1e-6 * max(1.0, abs(xstar)), (m, xstar)
    assert 0 < ic < RZ.MAXIT_MIN
    assert all(a <= x <= b for x in calls)
    ref = so.minimize_scalar(f, bracket=(a, g, b), method="brent", tol=1e-8).x
    assert abs(m - ref) <= 1e-6 * max(1.0, abs(ref))
    # known end values are not requested again and give the same iterates
    n = len(calls)
    calls.clear()
    m2, ic2 = RZ.brent_gsl_with_values(a, f(a), b, f(b), g, f(g), fc, 1e-7)
    assert (m2, ic2) == (m, ic) and len(calls) == n - 3


def test_brent_minimiser_edge_cases():
    assert math.isnan(RZ.brent_gsl(0.0, 1.0, 0.5, lambda x: x, 1e-8)[0])        # no enclosed minimum
    assert math.isnan(RZ.brent_gsl(-1.0, 1.0, 0.0, lambda x: x * x if x < 0.3 else math.nan, 1e-8)[0])
    assert RZ.min_test_interval(1.0, 1.0 + 1e-9, 0.0, 1e-8)
    assert not RZ.min_test_interval(-1e-9, 1e-9, 0.0, 1e-3)                      # bracket holds 0


def test_brent_minimisers_in_lockstep_equal_single_runs():
    f = lambda x: math.cos(3.0 * x) + 0.1 * x
    brackets = [(0.5, 1.5, 1.0), (2.5, 3.6, 3.1), (4.5, 5.6, 5.2)]
    out = RZ.lockstep([RZ.brent_min_iter(a, b, g, 1e-7) for a, b, g in brackets], lambda xs: [f(x) for x in xs])
    for (a, b, g), r in zip(brackets, out):
        assert r == RZ.brent_gsl(a, b, g, f, 1e-7)


def test_steffen_spline_properties():
    x = np.array([0.0, 0.4, 1.0, 1.3, 2.0, 3.1])
    y = np.array([0.0, 0.1, 0.9, 1.0, 1.0, 2.5])          # monotone data with a flat piece
    s = RZ.SteffenSpline(x, y)
    assert np.abs(s(x) - y).max() <= 1e-15
    t = np.linspace(0.0, 3.1, 2001)
    v = s(t)
    assert np.all(np.diff(v) >= -1e-15)                    # monotone interpolant
    assert np.all(np.abs(v[(t >= 1.3) & (t <= 2.0)] - 1.0) <= 1e-15)   # flat stays flat
    lin = RZ.SteffenSpline(x, 2.0 * x - 1.0)
    assert np.abs(lin(t) - (2.0 * t - 1.0)).max() <= 1e-14
    # C1 at the knots
    h = 1e-7
    for xi in x[1:-1]:
        assert abs((s(xi + h) - s(xi)) / h - (s(xi) - s(xi - h)) / h) <= 1e-5
    assert np.isnan(s(np.array([-0.1, 3.2]))).all()


def test_opt_subdiv_finds_every_minimum():
    f = lambda x: np.sin(7.0 * x) + 0.3 * np.cos(17.0 * x)
    t = np.linspace(0.0, 3.0, 300001)
    ft = f(t)
    true_min = t[1:-1][(ft[1:-1] < ft[:-2]) & (ft[1:-1] < ft[2:])]
    n, x, y, p = RZ.opt_subdiv(f, 0.0, 3.0, 32, 6, 1e-4, seed=3)
    assert n == x.size == y.size and np.all(np.diff(x) > 0) and np.array_equal(y, f(x))
    assert all(y[i] < y[i - 1] and y[i] < y[i + 1] for i in p)
    assert len(p) == len(true_min)
    for m in true_min:
        assert np.min(np.abs(x[p] - m)) < 0.05
    assert RZ.opt_subdiv(f, 0.0, 3.0, 32, 6, 1e-4, seed=3)[0] == n          # seeded: reproducible


@pytest.mark.gpu
def test_find_minima_brent_gsl_agrees(H):
    """Brent's minimiser on sigma (roots_brent_polygon_rsvd.cpp:288-311) finds
    the minima that Brent on dsigma/dk finds, to the resolution a minimiser
    has (sqrt(eps) relative)."""
    mesh = H.circle_mesh(16, 0.7, (0.0, 0.0))
    sp = H.ContinuousSpace1()
    bd = H.BuilderData(mesh, sp, sp, 8)
    R = H.gaussian_test_matrix(32, 2, 0)
    roots = RZ.find_minima(bd, 2.0, 6.0, 41, 1.0, 9.0, R, q=2, q_refine=30, tol=1e-11)
    mins = RZ.find_minima(bd, 2.0, 6.0, 41, 1.0, 9.0, R, q=2, q_refine=30, tol=1e-7, method="brent_gsl")
    assert len(mins) == len(roots) >= 1
    for (k1, s1, _), (k2, s2, _) in zip(roots, mins):
        assert abs(k1 - k2) <= 1e-5 * k1
        assert s2 <= s1 * (1 + 1e-9) + 1e-14 or abs(s2 - s1) <= 1e-9 * s1
