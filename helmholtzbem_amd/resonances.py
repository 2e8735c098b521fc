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
    if fl == fl and fh == fh:
        if fl * fh > 0.0:
            return (x1 if fl > fh else x2), False, 0
        if fl == 0.0:
            return x1, True, 0
        if fh == 0.0:
            return x2, True, 0
        xl, xh = (x1, x2) if fl < 0.0 else (x2, x1)
        rts = 0.5 * (x1 + x2)
    else:
        # an end value is NaN: the function's value at NaN is the first guess
        # (find_roots.cpp:159-162)
        rts = float((yield (math.nan, False)))
        xl, xh = x1, x2
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
                q: int = 2, q_refine: int = 20, tol: float = 1e-10, method: str = "brent", grid_sigma=None):
    """Local minima of sigma_min(k) on [k_min, k_max], the resonance search of
    the reference drivers: a batched sweep of n_points wavenumbers (rSVD with q
    iterations) brackets the minima, then every bracket is refined at once --
    Brent on dsigma/dk (method "brent"), safeguarded Newton with d2sigma/dk2
    (method "newton") or Brent's minimiser on sigma itself from the grid
    minimum (method "brent_gsl", examples/roots_brent_polygon_rsvd.cpp:288-311;
    tol is then relative) -- with q_refine subspace iterations.  grid_sigma:
    the grid's sigma values when the caller has them (e.g. from the
    mixed-precision sweep), instead of the FP64 sweep.  Returns a list of (k,
    sigma(k), iterations) for the brackets that converged."""
    from . import gaussian_test_matrix, sigma_sweep, sigma_sweep_der2
    n = 2 * builder_data.getNumPanels()
    if R is None:
        R = gaussian_test_matrix(n, 2, 0)
    ks = np.linspace(k_min, k_max, n_points)
    if grid_sigma is None:
        sig = sigma_sweep(builder_data, ks, c_o, c_i, R, q)
    else:
        sig = np.asarray(grid_sigma, dtype=np.float64)
        if sig.shape != ks.shape:
            raise ValueError("grid_sigma must hold one value per grid point")
    brackets = bracket_minima(ks, sig)
    if not brackets:
        return []
    if method == "brent":
        def evaluate(xs):
            return list(sigma_sweep(builder_data, np.array(xs, dtype=np.float64), c_o, c_i, R, q_refine, der=1)[1])
        res = lockstep([zbrent_iter(a, b, tol) for a, b in brackets], evaluate)
    elif method == "newton":
        def evaluate(reqs):
            # requests for (f, f') of all brackets in one batched sigma'' sweep,
            # the others (f only) in one dsigma/dk sweep
            out = [None] * len(reqs)
            both = [i for i, (_, b) in enumerate(reqs) if b]
            one = [i for i, (_, b) in enumerate(reqs) if not b]
            if both:
                xs = np.array([reqs[i][0] for i in both], dtype=np.float64)
                _, d1, d2 = sigma_sweep_der2(builder_data, xs, c_o, c_i, R, q_refine)
                for i, a, b in zip(both, d1, d2):
                    out[i] = (float(a), float(b))
            if one:
                xs = np.array([reqs[i][0] for i in one], dtype=np.float64)
                d1 = sigma_sweep(builder_data, xs, c_o, c_i, R, q_refine, der=1)[1]
                for i, a in zip(one, d1):
                    out[i] = float(a)
            return out
        res = lockstep([rtsafe_iter(a, b, tol) for a, b in brackets], evaluate)
    elif method == "brent_gsl":
        val = dict(zip(ks.tolist(), sig.tolist()))
        gens = []
        for a, b in brackets:
            g = float(ks[np.searchsorted(ks, a) + 1])
            gens.append(brent_min_iter(a, b, g, tol, val[a], val[b], val[g]))

        def evaluate(xs):
            return list(sigma_sweep(builder_data, np.array(xs, dtype=np.float64), c_o, c_i, R, q_refine))
        res = [(x, math.isfinite(x), ic) for x, ic in lockstep(gens, evaluate)]
    else:
        raise ValueError("method must be 'brent', 'newton' or 'brent_gsl'")
    found = [(r, it) for (r, ok, it) in res if ok]
    if not found:
        return []
    vals = sigma_sweep(builder_data, np.array([r for r, _ in found]), c_o, c_i, R, q_refine)
    return [(r, float(v), it) for (r, it), v in zip(found, vals)]


# ---------------------------------------------------------------------------
# Minimisation of sigma_min itself: brent_gsl and opt_subdiv
# ---------------------------------------------------------------------------
GOLDEN = 0.5 * (3.0 - math.sqrt(5.0))          # 0.381966...: golden-section fraction
SQRT_EPS = math.sqrt(EPS)
MAXIT_MIN = 100                                # brent_gsl's iteration cap (find_roots.cpp:1441)


def min_test_interval(a: float, b: float, epsabs: float, epsrel: float) -> bool:
    """Convergence of a bracket [a, b]: |b - a| < epsabs + epsrel min(|a|, |b|)
    when the bracket does not contain 0, else |b - a| < epsabs (the test the
    reference applies through GSL's gsl_min_test_interval)."""
    min_abs = min(abs(a), abs(b)) if (a > 0 and b > 0) or (a < 0 and b < 0) else 0.0
    return abs(b - a) < epsabs + epsrel * min_abs


def brent_min_iter(a: float, b: float, guess: float, tol: float, fa: Optional[float] = None,
                   fb: Optional[float] = None, fguess: Optional[float] = None,
                   max_iter: int = MAXIT_MIN) -> Generator[float, float, Tuple[float, int]]:
    """Brent's minimiser (R. P. Brent, Algorithms for Minimization without
    Derivatives, 1973, ch. 5: golden-section steps with parabolic
    interpolation through the three best points) on the bracket [a, b] with
    f(guess) below f(a) and f(b), in the form of GSL's fminimizer_brent that
    brent_gsl drives (find_roots.cpp:1426-1473): the second point is placed at
    a + golden (b - a), steps shorter than sqrt(eps) |x| are lengthened to it,
    and the loop stops when [a, b] passes min_test_interval(a, b, 0, tol) or
    after max_iter iterations.  Yields abscissae and receives f there;
    returns (x_minimum, iterations), x_minimum NaN when f is not finite.
    Values already known at a, b, guess may be passed in and are then not
    requested again."""
    lo, hi, z = float(a), float(b), float(guess)
    f_lo = (yield lo) if fa is None else float(fa)
    f_hi = (yield hi) if fb is None else float(fb)
    f_z = (yield z) if fguess is None else float(fguess)
    if not (f_z < f_lo and f_z < f_hi):
        return math.nan, 0                     # the bracket does not enclose a minimum
    v = w = lo + GOLDEN * (hi - lo)
    f_v = f_w = yield v
    d = e = 0.0
    for ic in range(1, max_iter + 1):
        if not all(math.isfinite(t) for t in (f_z, f_v, f_w)):
            return math.nan, ic
        tolerance = SQRT_EPS * abs(z)
        mid = 0.5 * (lo + hi)
        p = q = r = 0.0
        dd, ee = e, d                          # previous step lengths, shifted by one
        if abs(ee) > tolerance:
            r = (z - w) * (f_z - f_v)
            q = (z - v) * (f_z - f_w)
            p = (z - v) * q - (z - w) * r
            q = 2.0 * (q - r)
            if q > 0.0:
                p = -p
            else:
                q = -q
            r = ee
            ee = dd
        if abs(p) < abs(0.5 * q * r) and p < q * (z - lo) and p < q * (hi - z):
            dd = p / q
            u = z + dd
            if (u - lo) < 2.0 * tolerance or (hi - u) < 2.0 * tolerance:
                dd = tolerance if z < mid else -tolerance
        else:
            ee = (hi - z) if z < mid else -(z - lo)
            dd = GOLDEN * ee
        u = z + dd if abs(dd) >= tolerance else z + (tolerance if dd > 0.0 else -tolerance)
        e, d = ee, dd
        f_u = yield u
        if not math.isfinite(f_u):
            return math.nan, ic
        if f_u <= f_z:
            if u < z:
                hi = z
            else:
                lo = z
            v, f_v = w, f_w
            w, f_w = z, f_z
            z, f_z = u, f_u
        else:
            if u < z:
                lo = u
            else:
                hi = u
            if f_u <= f_w or w == z:
                v, f_v = w, f_w
                w, f_w = u, f_u
            elif f_u <= f_v or v == z or v == w:
                v, f_v = u, f_u
        if min_test_interval(lo, hi, 0.0, tol):
            return z, ic
    return z, max_iter


def brent_gsl(a: float, b: float, guess: float, f: Callable[[float], float], tol: float) -> Tuple[float, int]:
    """Local minimum of f in [a, b] from guess (find_roots.cpp:1426-1473):
    (x_minimum, iterations); NaN when the bracket does not enclose a minimum
    or f is not finite."""
    return _run(brent_min_iter(a, b, guess, tol), f)


def brent_gsl_with_values(a: float, fa: float, b: float, fb: float, guess: float, fguess: float,
                          f: Callable[[float], float], tol: float) -> Tuple[float, int]:
    """brent_gsl with the values at a, b and guess known (find_roots.cpp:1475-1512)."""
    return _run(brent_min_iter(a, b, guess, tol, fa, fb, fguess), f)


class SteffenSpline:
    """Monotone piecewise-cubic interpolation of M. Steffen (A simple method for
    monotonic interpolation in one dimension, Astron. Astrophys. 239, 1990):
    interior slopes (sign s_{i-1} + sign s_i) min(|s_{i-1}|, |s_i|, |p_i| / 2)
    with p_i the slope of the parabola through three points, end slopes the
    end secants ("simplest possibility", section 2.2).  The interpolant the
    reference's opt_subdiv fits through the samples (GSL's steffen type)."""

    def __init__(self, x, y):
        x = np.asarray(x, dtype=np.float64)
        y = np.asarray(y, dtype=np.float64)
        h = np.diff(x)
        s = np.diff(y) / h
        yp = np.empty_like(x)
        yp[0], yp[-1] = s[0], s[-1]
        p = (s[:-1] * h[1:] + s[1:] * h[:-1]) / (h[:-1] + h[1:])
        yp[1:-1] = (np.sign(s[:-1]) + np.sign(s[1:])) * np.minimum(np.minimum(np.abs(s[:-1]), np.abs(s[1:])),
                                                                   0.5 * np.abs(p))
        self.x, self.y = x, y
        self.c3 = (yp[:-1] + yp[1:] - 2.0 * s) / (h * h)
        self.c2 = (3.0 * s - 2.0 * yp[:-1] - yp[1:]) / h
        self.c1 = yp[:-1]

    def __call__(self, t):
        """Values at t; NaN outside [x_0, x_{n-1}] (GSL_EDOM in the reference)."""
        t = np.asarray(t, dtype=np.float64)
        i = np.clip(np.searchsorted(self.x, t, side="right") - 1, 0, len(self.x) - 2)
        dl = t - self.x[i]
        v = self.y[i] + dl * (self.c1[i] + dl * (self.c2[i] + dl * self.c3[i]))
        return np.where((t >= self.x[0]) & (t <= self.x[-1]), v, np.nan)


def opt_subdiv(f_batch: Callable[[np.ndarray], np.ndarray], x_min: float, x_max: float, batch_size: int, limit: int,
               tol: float, seed: int = 0, max_rounds: int = 1000):
    """Adaptive sampling of [x_min, x_max] until the count of local minima of
    the samples has not grown for `limit` rounds (find_roots.cpp:1615-1726).
    Every round adds batch_size points evaluated together by f_batch (one
    batched sweep): uniform on odd rounds, otherwise Gaussian around mu with
    width sigma where a Steffen spline through the earlier samples
    mispredicted the last batch (mu the error-weighted mean of their
    abscissae, sigma = d erravg (cnt - 1) / (6 (errmax - erravg))), rejecting
    points closer than tol d to a sample.  Returns (n, x, y, p): the sorted
    samples and the positions of their interior local minima.  The reference
    draws unseeded std::mt19937 numbers and loses its two endpoint samples to
    the first batch; here the draws are seeded and the endpoints are not
    sampled."""
    rng = np.random.default_rng(seed)
    d = x_max - x_min
    xs = np.zeros(0)
    ys = np.zeros(0)
    sigma, mu = -1.0, 0.0
    count = mc = 0
    p: List[int] = []
    it = 0
    while count < limit:
        it += 1
        if it >= max_rounds:
            break
        spl = SteffenSpline(xs, ys) if xs.size > 3 else None
        new = []
        for _ in range(batch_size):
            while True:
                if it % 2 or not (0.0 <= sigma <= 3.0 * d):
                    x0 = x_min + rng.random() * d
                else:
                    x0 = rng.normal(mu, sigma)
                    while x0 <= x_min or x0 >= x_max:
                        x0 = rng.normal(mu, sigma)
                near = np.any(np.abs(x0 - xs) < tol * d) or any(abs(x0 - t) < tol * d for t in new)
                if not near:
                    break
            new.append(x0)
        new = np.array(new)
        fy = np.asarray(f_batch(new), dtype=np.float64)
        if spl is not None:
            err = np.abs(fy - spl(new))
            ok = np.isfinite(err)
            cnt = int(ok.sum())
            if cnt > 0:
                err = err[ok]
                erravg, errmax = float(err.mean()), float(err.max())
                mu = float((err * new[ok]).sum() / (erravg * cnt)) if erravg > 0 else mu
                sigma = d * erravg * (cnt - 1.0) / (6.0 * (errmax - erravg)) if errmax > erravg else math.inf
        xs = np.concatenate([xs, new])
        ys = np.concatenate([ys, fy])
        order = np.argsort(xs, kind="stable")
        xs, ys = xs[order], ys[order]
        p = [i for i in range(1, xs.size - 1) if ys[i] < min(ys[i - 1], ys[i + 1])]
        if len(p) > mc:
            count, mc = 0, len(p)
        else:
            count += 1
    return xs.size, xs, ys, p


def sample_k_interval(builder_data, k_min: float, k_max: float, c_o: float, c_i: float, R=None, q: int = 2,
                      batch_size: int = 256, limit: int = 10, tol: Optional[float] = None, seed: int = 0, proj=True):
    """opt_subdiv over sigma_min(k) as examples/find_resonances.cpp:141 runs it
    (limit 10, tol 1e-4 (k_max - k_min)), each round one batched device sweep
    of batch_size wavenumbers.  Returns (n, ks, sigma, minima positions)."""
    from . import gaussian_test_matrix, sigma_sweep
    if R is None:
        R = gaussian_test_matrix(2 * builder_data.getNumPanels(), 2, 0)
    if tol is None:
        tol = 1e-4 * (k_max - k_min)
    return opt_subdiv(lambda ks: sigma_sweep(builder_data, ks, c_o, c_i, R, q, proj=proj), k_min, k_max, batch_size,
                      limit, tol, seed)
