// find_roots.hpp -- the root finders of the reference's resonance drivers for
// the C++ facade (header-only, host code): zbrent and rtsafe with the
// reference's signatures (find_roots.hpp:37-83, find_roots.cpp:19-215).
//
// zbrent: Brent's method (R. P. Brent, Algorithms for Minimization without
// Derivatives, 1973, ch. 4) on [x1, x2]; rtsafe: Newton steps on fct with
// fct_both(x) = [f(x), f'(x)] (a 1 x 2 matrix), bisecting whenever a step
// leaves the bracket or does not halve the previous step.  Step for step the
// same as helmholtzbem_amd/resonances.py, so both give the same iterates.
#pragma once

#include <cmath>
#include <functional>
#include <limits>

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
    if (fl * fh > 0.0) return fl > fh ? x1 : x2;
    if (fl == 0.0) { root_found = true; return x1; }
    if (fh == 0.0) { root_found = true; return x2; }
    double xl, xh;
    if (fl < 0.0) { xl = x1; xh = x2; } else { xh = x1; xl = x2; }
    double rts = 0.5 * (x1 + x2), dxold = std::fabs(x2 - x1), dx = dxold;
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
