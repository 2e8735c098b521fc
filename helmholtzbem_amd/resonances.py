"""Resonance search on top of the batched sigma_min(k) path (host side).

The reference's drivers (examples/roots_brent_*_rsvd.cpp,
examples/roots_newton_*_rsvd.cpp) sample sigma_min(k) on a k grid with the
randomized SVD, bracket the local minima, and refine each bracket with a
scalar root finder on dsigma/dk: Brent's method (``zbrent``,
find_roots.cpp:19-115) or safeguarded Newton with sigma'' (``rtsafe``,
find_roots.cpp:117-215).  Here both root finders are written as resumable
iterations (generators that yield the next abscissa and receive the function
values), so ``find_minima`` advances every bracket in lockstep and evaluates
all pending wavenumbers of a round in one batched device sweep
(``hbem_sigma_sweep``) instead of one k at a time.  ``zbrent`` and ``rtsafe``
keep the reference's scalar interface (root, root_found, num_iter).

Brent's method and the Newton/bisection safeguard follow the published
algorithms (R. P. Brent, Algorithms for Minimization without Derivatives,
1973, ch. 4; the rtsafe safeguard as in the reference), step for step as the
reference applies them, so the iterates are the same for the same function.
"""
from __future__ import annotations

import math
from typing import Callable, Generator, List, Optional, Sequence, Tuple

import numpy as np

MAXIT = 1000
EPS = 2.220446049250313e-16

Result = Tuple[float, bool, int]   # (root, root_found, num_iter)


def zbrent_iter(x1: float, x2: float, tol: float) -> Generator[float, float, Result]:
    """Brent's root finder on [x1, x2]: yields abscissae, receives f there.
    Without a sign change it returns (0.0, False, 0) as the reference does."""
    a, b, c = float(x1), float(x2), float(x2)
    fa = yield a
    fb = yield b
    if fa * fb > 0.0:
        return 0.0, False, 0
    fc = fb
    d = e = 0.0
    for it in range(1, MAXIT + 1):
        if (fb > 0.0 and fc > 0.0) or (fb < 0.0 and fc < 0.0):
            c, fc = a, fa
            e = d = b - a
        if abs(fc) < abs(fb):
            a, b, c = b, c, b
            fa, fb, fc = fb, fc, fb
        tol1 = 2.0 * EPS * abs(b) + 0.5 * tol
        xm = 0.5 * (c - b)
        if abs(xm) <= tol1 or fb == 0.0:
            return b, True, it
        if abs(e) >= tol1 and abs(fa) > abs(fb):
            s = fb / fa
            if a == c:
                p = 2.0 * xm * s
                q = 1.0 - s
            else:
                q = fa / fc
                r = fb / fc
                p = s * (2.0 * xm * q * (q - r) - (b - a) * (r - 1.0))
                q = (q - 1.0) * (r - 1.0) * (s - 1.0)
            if p > 0.0:
                q = -q
            p = abs(p)
            m1 = 3.0 * xm * q - abs(tol1 * q)
            m2 = abs(e * q)
            if 2.0 * p < (m1 if m1 < m2 else m2):
                e = d
                d = p / q
            else:
                d = xm
                e = d
        else:
            d = xm
            e = d
        a, fa = b, fb
        if abs(d) > tol1:
            b += d
        else:
            b += abs(tol1) if xm >= 0 else -abs(tol1)
        fb = yield b
    return 0.0, False, MAXIT


def rtsafe_iter(x1: float, x2: float, tol: float) -> Generator[Tuple[float, bool], object, Result]:
    """Safeguarded Newton on [x1, x2] for a root of f with derivative df:
    yields (x, want_derivative) and receives f(x) or (f(x), f'(x))."""
    x1, x2 = float(x1), float(x2)
    fl = yield (x1, False)
    fh = yield (x2, False)
    if fl * fh > 0.0:
        return (x1 if fl > fh else x2), False, 0
    if fl == 0.0:
        return x1, True, 0
    if fh == 0.0:
        return x2, True, 0
    xl, xh = (x1, x2) if fl < 0.0 else (x2, x1)
    rts = 0.5 * (x1 + x2)
    dxold = abs(x2 - x1)
    dx = dxold
    f, df = yield (rts, True)
    for j in range(1, MAXIT + 1):
        if ((rts - xh) * df - f) * ((rts - xl) * df - f) > 0.0 or abs(2.0 * f) > abs(dxold * df):
            dxold = dx
            dx = 0.5 * (xh - xl)
            rts = xl + dx
            if abs(xl - rts) < EPS:
                return rts, True, j
        else:
            dxold = dx
            dx = f / df
            temp = rts
            rts -= dx
            if temp == rts:
                return rts, True, j
        if abs(dx) < tol:
            return rts, True, j
        f, df = yield (rts, True)
        if f < 0.0:
            xl = rts
        else:
            xh = rts
    return rts, False, MAXIT


def _run(gen, answer):
    try:
        req = next(gen)
        while True:
            req = gen.send(answer(req))
    except StopIteration as stop:
        return stop.value


def zbrent(f: Callable[[float], float], x1: float, x2: float, tol: float) -> Result:
    """Root of f in [x1, x2] (find_roots.cpp:19-115): (root, root_found, num_iter)."""
    return _run(zbrent_iter(x1, x2, tol), f)


def rtsafe(fct: Callable[[float], float], fct_both: Callable[[float], Sequence[float]], x1: float, x2: float,
           tol: float) -> Result:
    """Root of fct in [x1, x2] by Newton steps with fct_both = (f, f'),
    bisecting when a step leaves the bracket or converges slowly
    (find_roots.cpp:117-215): (root, root_found, num_iter)."""
    def answer(req):
        x, both = req
        if both:
            v = fct_both(x)
            return float(v[0]), float(v[1])
        return fct(x)
    return _run(rtsafe_iter(x1, x2, tol), answer)


def lockstep(gens: List[Generator], evaluate: Callable[[list], list]) -> list:
    """Runs resumable root finders together: each round collects every pending
    request and answers them with one call of evaluate(requests)."""
    results: list = [None] * len(gens)
    pending = {}
    for i, g in enumerate(gens):
        try:
            pending[i] = next(g)
        except StopIteration as stop:
            results[i] = stop.value
    while pending:
        idx = list(pending)
        answers = evaluate([pending[i] for i in idx])
        for i, ans in zip(idx, answers):
            try:
                pending[i] = gens[i].send(ans)
            except StopIteration as stop:
                results[i] = stop.value
                del pending[i]
    return results


def bracket_minima(ks: Sequence[float], sig: Sequence[float]) -> List[Tuple[float, float]]:
    """Brackets [k_{i}, k_{i+2}] around every grid point that is lower than both
    neighbours (examples/roots_brent_square_rsvd.cpp:180-193)."""
    ks = np.asarray(ks, dtype=np.float64)
    out = []
    for i in range(len(sig) - 2):
        L = sig[i + 1] - sig[i]
        R = sig[i + 2] - sig[i + 1]
        if L < 0.0 and R > 0.0:
            out.append((float(ks[i]), float(ks[i + 2])))
    return out


def find_minima(builder_data, k_min: float, k_max: float, n_points: int, c_o: float, c_i: float, R=None,
                q: int = 2, q_refine: int = 20, tol: float = 1e-10, method: str = "brent"):
    """Local minima of sigma_min(k) on [k_min, k_max], the resonance search of
    the reference drivers: a batched sweep of n_points wavenumbers (rSVD with q
    iterations) brackets the minima, then every bracket is refined at once --
    Brent on dsigma/dk (method "brent") or safeguarded Newton with d2sigma/dk2
    (method "newton") -- with q_refine subspace iterations.  Returns a list of
    (k, sigma(k), iterations) for the brackets that converged."""
    from . import SolutionsOperator, gaussian_test_matrix, randomized_svd, sigma_sweep
    n = 2 * builder_data.getNumPanels()
    if R is None:
        R = gaussian_test_matrix(n, 2, 0)
    ks = np.linspace(k_min, k_max, n_points)
    sig = sigma_sweep(builder_data, ks, c_o, c_i, R, q)
    brackets = bracket_minima(ks, sig)
    if not brackets:
        return []
    if method == "brent":
        def evaluate(xs):
            return list(sigma_sweep(builder_data, np.array(xs, dtype=np.float64), c_o, c_i, R, q_refine, der=1)[1])
        res = lockstep([zbrent_iter(a, b, tol) for a, b in brackets], evaluate)
    elif method == "newton":
        so = SolutionsOperator(builder_data)

        def evaluate(reqs):
            xs = np.array([x for x, _ in reqs], dtype=np.float64)
            out = []
            d1 = sigma_sweep(builder_data, xs, c_o, c_i, R, q_refine, der=1)[1]
            for (x, both), s1 in zip(reqs, d1):
                if both:
                    T, Td, Tdd = so.gen_sol_op_2nd_der(x, c_o, c_i)
                    out.append((float(s1), float(randomized_svd.sv_der2(T, Td, Tdd, R, q_refine)[2])))
                else:
                    out.append(float(s1))
            return out
        res = lockstep([rtsafe_iter(a, b, tol) for a, b in brackets], evaluate)
    else:
        raise ValueError("method must be 'brent' or 'newton'")
    found = [(r, it) for (r, ok, it) in res if ok]
    if not found:
        return []
    vals = sigma_sweep(builder_data, np.array([r for r, _ in found]), c_o, c_i, R, q_refine)
    return [(r, float(v), it) for (r, it), v in zip(found, vals)]
