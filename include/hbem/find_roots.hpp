// find_roots.hpp -- the root finders of the reference's resonance drivers for
// the C++ facade (header-only, host code): zbrent and rtsafe with the
// reference's signatures (find_roots.hpp:37-83, find_roots.cpp:19-215).
//
// zbrent: Brent's method (R. P. Brent, Algorithms for Minimization without
// Derivatives, 1973, ch. 4) on [x1, x2]; rtsafe: Newton steps on fct with
// fct_both(x) = [f(x), f'(x)] (a 1 x 2 matrix), bisecting whenever a step
// leaves the bracket or does not halve the previous step; brent_gsl /
// brent_gsl_with_values (find_roots.hpp:262-265, find_roots.cpp:1426-1512):
// Brent's minimiser (ch. 5 of the same book) in the form of the GSL
// fminimizer the reference wraps.  Step for step the same as
// helmholtzbem_amd/resonances.py, so both give the same iterates.
#pragma once

#include <cmath>
#include <functional>
#include <limits>
#include <vector>

#include "hbem.hpp"

namespace hbem_roots_detail {
constexpr int kMaxIt = 1000;
constexpr double kEps = std::numeric_limits<double>::epsilon();
}  // namespace hbem_roots_detail

// Root of f in [x1, x2]; without a sign change returns 0 with root_found false.
inline double zbrent(const std::function<double(double)> f, double x1, double x2, double tol, bool& root_found,
                     unsigned& num_iter) {
    using namespace hbem_roots_detail;
    double a = x1, b = x2, c = x2, d = 0.0, e = 0.0;
    double fa = f(a), fb = f(b), fc;
    root_found = false;
    if (fa * fb > 0.0) return 0.0;
    fc = fb;
    for (int it = 1; it <= kMaxIt; ++it) {
        if ((fb > 0.0 && fc > 0.0) || (fb < 0.0 && fc < 0.0)) {
            c = a;
            fc = fa;
            e = d = b - a;
        }
        if (std::fabs(fc) < std::fabs(fb)) {
            a = b; b = c; c = a;
            fa = fb; fb = fc; fc = fa;
        }
        const double tol1 = 2.0 * kEps * std::fabs(b) + 0.5 * tol;
        const double xm = 0.5 * (c - b);
        if (std::fabs(xm) <= tol1 || fb == 0.0) {
            root_found = true;
            num_iter = (unsigned)it;
            return b;
        }
        if (std::fabs(e) >= tol1 && std::fabs(fa) > std::fabs(fb)) {
            const double s = fb / fa;
            double p, q;
            if (a == c) {
                p = 2.0 * xm * s;
                q = 1.0 - s;
            } else {
                q = fa / fc;
                const double r = fb / fc;
                p = s * (2.0 * xm * q * (q - r) - (b - a) * (r - 1.0));
                q = (q - 1.0) * (r - 1.0) * (s - 1.0);
            }
            if (p > 0.0) q = -q;
            p = std::fabs(p);
            const double m1 = 3.0 * xm * q - std::fabs(tol1 * q), m2 = std::fabs(e * q);
            if (2.0 * p < (m1 < m2 ? m1 : m2)) {
                e = d;
                d = p / q;
            } else {
                d = xm;
                e = d;
            }
        } else {
            d = xm;
            e = d;
        }
        a = b;
        fa = fb;
        if (std::fabs(d) > tol1) b += d;
        else b += xm >= 0 ? std::fabs(tol1) : -std::fabs(tol1);
        fb = f(b);
    }
    num_iter = (unsigned)kMaxIt;
    return 0.0;
}

// Root of fct in [x1, x2] by safeguarded Newton with fct_both = [f, f'].
inline double rtsafe(std::function<double(double)> fct, std::function<hbem::MatrixXd(double)> fct_both, double x1,
                     double x2, double tol, bool& root_found, unsigned& num_iter) {
    using namespace hbem_roots_detail;
    root_found = false;
    num_iter = 0;
    const double fl = fct(x1), fh = fct(x2);
    double xl, xh, rts;
    if (fl == fl && fh == fh) {
        if (fl * fh > 0.0) return fl > fh ? x1 : x2;
        if (fl == 0.0) { root_found = true; return x1; }
        if (fh == 0.0) { root_found = true; return x2; }
        if (fl < 0.0) { xl = x1; xh = x2; } else { xh = x1; xl = x2; }
        rts = 0.5 * (x1 + x2);
    } else {
        // an end value is NaN: fct(NaN) supplies the first guess (find_roots.cpp:159-162)
        rts = fct(std::numeric_limits<double>::quiet_NaN());
        xl = x1;
        xh = x2;
    }
    double dxold = std::fabs(x2 - x1), dx = dxold;
    hbem::MatrixXd m = fct_both(rts);
    double f = m(0, 0), df = m(0, 1);
    for (int j = 1; j <= kMaxIt; ++j) {
        if ((((rts - xh) * df - f) * ((rts - xl) * df - f) > 0.0) || (std::fabs(2.0 * f) > std::fabs(dxold * df))) {
            dxold = dx;
            dx = 0.5 * (xh - xl);
            rts = xl + dx;
            if (std::fabs(xl - rts) < kEps) { root_found = true; num_iter = (unsigned)j; return rts; }
        } else {
            dxold = dx;
            dx = f / df;
            const double temp = rts;
            rts -= dx;
            if (temp == rts) { root_found = true; num_iter = (unsigned)j; return rts; }
        }
        if (std::fabs(dx) < tol) { root_found = true; num_iter = (unsigned)j; return rts; }
        m = fct_both(rts);
        f = m(0, 0);
        df = m(0, 1);
        if (f < 0.0) xl = rts;
        else xh = rts;
    }
    num_iter = (unsigned)kMaxIt;
    return rts;
}

namespace hbem_roots_detail {
constexpr unsigned kMaxItMin = 100;

// |b - a| < epsabs + epsrel min(|a|, |b|) for a bracket not holding 0
inline bool min_test_interval(double a, double b, double epsabs, double epsrel) {
    const double min_abs = ((a > 0 && b > 0) || (a < 0 && b < 0)) ? std::min(std::fabs(a), std::fabs(b)) : 0.0;
    return std::fabs(b - a) < epsabs + epsrel * min_abs;
}

// Brent's minimiser on [lo, hi] from z with f(z) below f(lo), f(hi).
inline double brent_min(double lo, double f_lo, double hi, double f_hi, double z, double f_z,
                        const std::function<double(double)>& f, double tol, unsigned& ic) {
    const double golden = 0.5 * (3.0 - std::sqrt(5.0)), sqrt_eps = std::sqrt(kEps);
    const double nan = std::numeric_limits<double>::quiet_NaN();
    ic = 0;
    if (!(f_z < f_lo && f_z < f_hi)) return nan;
    double v = lo + golden * (hi - lo), w = v;
    double f_v = f(v), f_w = f_v, d = 0.0, e = 0.0;
    for (unsigned it = 1; it <= kMaxItMin; ++it) {
        ic = it;
        if (!std::isfinite(f_z) || !std::isfinite(f_v) || !std::isfinite(f_w)) return nan;
        const double tolerance = sqrt_eps * std::fabs(z), mid = 0.5 * (lo + hi);
        double p = 0.0, q = 0.0, r = 0.0, dd = e, ee = d;
        if (std::fabs(ee) > tolerance) {
            r = (z - w) * (f_z - f_v);
            q = (z - v) * (f_z - f_w);
            p = (z - v) * q - (z - w) * r;
            q = 2.0 * (q - r);
            if (q > 0.0) p = -p;
            else q = -q;
            r = ee;
            ee = dd;
        }
        if (std::fabs(p) < std::fabs(0.5 * q * r) && p < q * (z - lo) && p < q * (hi - z)) {
            dd = p / q;
            const double u = z + dd;
            if ((u - lo) < 2.0 * tolerance || (hi - u) < 2.0 * tolerance) dd = z < mid ? tolerance : -tolerance;
        } else {
            ee = z < mid ? (hi - z) : -(z - lo);
            dd = golden * ee;
        }
        const double u = std::fabs(dd) >= tolerance ? z + dd : z + (dd > 0.0 ? tolerance : -tolerance);
        e = ee;
        d = dd;
        const double f_u = f(u);
        if (!std::isfinite(f_u)) return nan;
        if (f_u <= f_z) {
            if (u < z) hi = z;
            else lo = z;
            v = w; f_v = f_w;
            w = z; f_w = f_z;
            z = u; f_z = f_u;
        } else {
            if (u < z) lo = u;
            else hi = u;
            if (f_u <= f_w || w == z) {
                v = w; f_v = f_w;
                w = u; f_w = f_u;
            } else if (f_u <= f_v || v == z || v == w) {
                v = u; f_v = f_u;
            }
        }
        if (min_test_interval(lo, hi, 0.0, tol)) return z;
    }
    return z;
}
}  // namespace hbem_roots_detail

// Local minimum of f in [a, b] from guess; NaN when [a, b] does not enclose a
// minimum below f(guess) or f is not finite.  The progress and value
// callbacks of the reference (rp, rv) are accepted and not called.
inline double brent_gsl(double a, double b, double guess, std::function<double(double)>& f, double tol,
                        unsigned int& ic, std::function<void(double)>* rp = nullptr,
                        std::function<void(double, double)>* rv = nullptr) {
    (void)rp;
    (void)rv;
    const double fa = f(a), fb = f(b), fg = f(guess);
    return hbem_roots_detail::brent_min(a, fa, b, fb, guess, fg, f, tol, ic);
}

inline double brent_gsl_with_values(double a, double fa, double b, double fb, double guess, double fguess,
                                    std::function<double(double)>& f, double tol, unsigned int& ic) {
    return hbem_roots_detail::brent_min(a, fa, b, fb, guess, fguess, f, tol, ic);
}
