"""Resonance search drivers (find_roots.cpp:19-215, examples/roots_*_rsvd.cpp):
zbrent and rtsafe against scipy's brentq on analytic functions, their edge
cases, lockstep execution against individual runs, the C++ facade header
against the Python host code iterate for iterate, and (GPU) the batched
minimum search against the oracle's dsigma/dk."""
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
