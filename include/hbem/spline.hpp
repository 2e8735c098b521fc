// spline.hpp -- complex Akima splines and spline panels for the C++ facade
// (header-only host code): ComplexSpline (cspline.hpp, cspline.cpp:12-76) and
// ParametrizedPolynomial (parametrized_polynomial.cpp:21-173), the panels of
// smooth scatterers.
//
// Akima's method (H. Akima, J. ACM 17, 1970): secants m_j, node slopes
//   t_i = (|m_{i+1} - m_i| m_{i-1} + |m_{i-1} - m_{i-2}| m_i) / (|m_{i+1} - m_i| + |m_{i-1} - m_{i-2}|)
// (the mean of m_{i-1}, m_i when the denominator is 0), end secants continued
// periodically or by linear extrapolation, and on [x_j, x_{j+1}] the cubic
//   y_j + t_j d + (3 m_j - 2 t_j - t_{j+1}) / h d^2 + (t_j + t_{j+1} - 2 m_j) / h^2 d^3.
// The arithmetic is in the order of helmholtzbem_amd/spline.py, so the
// coefficient rows are the same bit for bit (compile with -ffp-contract=off).
//
// Spline panels carry the chain rule of the parameter map in Derivative_01
// ((tmax - tmin) z'(t)), which the Galerkin integrals need; the reference
// returns z'(t) there (parametrized_polynomial.cpp:91-97).
#pragma once

#include <cmath>
#include <complex>
#include <memory>
#include <stdexcept>
#include <vector>

#include "hbem.hpp"

namespace hbem {

namespace detail {
// n-point Gauss-Legendre rule on [-1, 1] by Newton on the Legendre recurrence
inline void gauss_legendre(int n, std::vector<double>& x, std::vector<double>& w) {
    x.assign(n, 0.0);
    w.assign(n, 0.0);
    for (int i = 0; i < n; ++i) {
        double z = std::cos(M_PI * (i + 0.75) / (n + 0.5)), dp = 1.0;
        for (int it = 0; it < 100; ++it) {
            double p0 = 1.0, p1 = z;
            for (int j = 2; j <= n; ++j) {
                const double p2 = ((2.0 * j - 1.0) * z * p1 - (j - 1.0) * p0) / j;
                p0 = p1;
                p1 = p2;
            }
            dp = n * (z * p1 - p0) / (z * z - 1.0);
            const double dz = p1 / dp;
            z -= dz;
            if (std::fabs(dz) < 1e-16) break;
        }
        x[i] = -z;
        w[i] = 2.0 / ((1.0 - z * z) * dp * dp);
    }
}
}  // namespace detail

class ComplexSpline {
  public:
    ComplexSpline(const std::vector<double>& x, const std::vector<std::complex<double>>& y, bool periodic)
        : x_(x), y_(y) {
        const size_t n = x.size();
        if (n != y.size() || n < (periodic ? 4u : 5u)) throw std::invalid_argument("not enough points for an Akima spline");
        for (size_t i = 0; i + 1 < n; ++i)
            if (!(x[i] < x[i + 1])) throw std::invalid_argument("spline abscissae must increase");
        for (int part = 0; part < 2; ++part) {
            std::vector<double> v(n), m(n - 1), h(n - 1);
            for (size_t i = 0; i < n; ++i) v[i] = part ? y[i].imag() : y[i].real();
            for (size_t j = 0; j + 1 < n; ++j) {
                h[j] = x[j + 1] - x[j];
                m[j] = (v[j + 1] - v[j]) / h[j];
            }
            std::vector<double> ext(n + 3);   // m_{-2} .. m_n
            if (periodic) {
                ext[0] = m[n - 3];
                ext[1] = m[n - 2];
                ext[n + 1] = m[0];
                ext[n + 2] = m[1];
            } else {
                ext[0] = 3.0 * m[0] - 2.0 * m[1];
                ext[1] = 2.0 * m[0] - m[1];
                ext[n + 1] = 2.0 * m[n - 2] - m[n - 3];
                ext[n + 2] = 3.0 * m[n - 2] - 2.0 * m[n - 3];
            }
            for (size_t j = 0; j + 1 < n; ++j) ext[j + 2] = m[j];
            // node slopes; flat[i]: both Akima weights at node i vanish (two
            // straight runs meet there), where GSL's akima keeps each run exact
            std::vector<double> t(n);
            std::vector<char> flat(n);
            for (size_t i = 0; i < n; ++i) {
                const double mm2 = ext[i], mm1 = ext[i + 1], m0 = ext[i + 2], mp1 = ext[i + 3];
                const double w1 = std::fabs(mp1 - m0), w2 = std::fabs(mm1 - mm2);
                flat[i] = w1 + w2 == 0.0;
                t[i] = flat[i] ? 0.0 : (w1 * mm1 + w2 * m0) / (w1 + w2);
            }
            std::vector<double>& b = part ? by_ : bx_;
            std::vector<double>& c = part ? cy_ : cx_;
            std::vector<double>& d = part ? dy_ : dx_;
            b.resize(n - 1);
            c.resize(n - 1);
            d.resize(n - 1);
            for (size_t j = 0; j + 1 < n; ++j) {
                if (flat[j]) {              // a linear piece of slope m_j
                    b[j] = m[j];
                    c[j] = 0.0;
                    d[j] = 0.0;
                    continue;
                }
                const double tr = flat[j + 1] ? m[j] : t[j + 1];   // slope at the right end
                b[j] = t[j];
                c[j] = (3.0 * m[j] - 2.0 * t[j] - tr) / h[j];
                d[j] = (t[j] + tr - 2.0 * m[j]) / (h[j] * h[j]);
            }
        }
        rows_ = std::make_shared<std::vector<double>>();
        for (size_t j = 0; j + 1 < n; ++j) {
            const double r[9] = {x[j], y[j].real(), y[j].imag(), bx_[j], cx_[j], dx_[j], by_[j], cy_[j], dy_[j]};
            rows_->insert(rows_->end(), r, r + 9);
        }
        const double last[3] = {x[n - 1], y[n - 1].real(), y[n - 1].imag()};
        rows_->insert(rows_->end(), last, last + 3);
    }
    size_t pieces() const { return x_.size() - 1; }
    std::complex<double> eval(double t) const {
        const size_t j = piece(t);
        const double d = t - x_[j];
        return {y_[j].real() + d * (bx_[j] + d * (cx_[j] + d * dx_[j])),
                y_[j].imag() + d * (by_[j] + d * (cy_[j] + d * dy_[j]))};
    }
    // derivative of order 0, 1 or 2 (cspline.cpp:68-76)
    std::complex<double> eval_der(double t, int order = 1) const {
        if (order < 0 || order > 2) throw std::runtime_error("Order of the spline derivative must be 0, 1 or 2");
        if (order == 0) return eval(t);
        const size_t j = piece(t);
        const double d = t - x_[j];
        if (order == 1)
            return {bx_[j] + d * (2.0 * cx_[j] + 3.0 * d * dx_[j]), by_[j] + d * (2.0 * cy_[j] + 3.0 * d * dy_[j])};
        return {2.0 * cx_[j] + 6.0 * d * dx_[j], 2.0 * cy_[j] + 6.0 * d * dy_[j]};
    }
    // arc length over [a, b]: 16-point Gauss-Legendre on every piece meeting it
    double arc_length(double a, double b) const {
        if (b <= a) return 0.0;
        static thread_local std::vector<double> gx, gw;
        if (gx.empty()) detail::gauss_legendre(16, gx, gw);
        double total = 0.0;
        for (size_t j = 0; j + 1 < x_.size(); ++j) {
            const double u = std::max(a, x_[j]), v = std::min(b, x_[j + 1]);
            if (v <= u) continue;
            double s = 0.0;
            for (int q = 0; q < 16; ++q) s += gw[q] * std::abs(eval_der(0.5 * (v - u) * gx[q] + 0.5 * (u + v)));
            total += 0.5 * (v - u) * s;
        }
        return total;
    }
    // device table rows: 9 doubles per piece {t_j, x_j, y_j, bx, cx, dx, by, cy, dy}, then {t_K, x_K, y_K}
    std::shared_ptr<const std::vector<double>> rows() const { return rows_; }

  private:
    size_t piece(double t) const {
        size_t lo = 0, hi = x_.size() - 1;
        while (hi - lo > 1) {
            const size_t mid = (lo + hi) / 2;
            if (x_[mid] <= t) lo = mid;
            else hi = mid;
        }
        return lo;
    }
    std::vector<double> x_;
    std::vector<std::complex<double>> y_;
    std::vector<double> bx_, cx_, dx_, by_, cy_, dy_;
    std::shared_ptr<std::vector<double>> rows_;
};

class ParametrizedPolynomial : public AbstractParametrizedCurve {
  public:
    ParametrizedPolynomial(std::shared_ptr<const ComplexSpline> spline, double tmin, double tmax, double len = -1.0)
        : s_(std::move(spline)), t0_(tmin), t1_(tmax), len_(len >= 0.0 ? len : s_->arc_length(tmin, tmax)),
          rows_(s_->rows()) {}
    // operator() on [-1, 1] (parametrized_polynomial.cpp:25-32)
    Point2 operator()(double t) const {
        const std::complex<double> z = s_->eval(t * (t1_ - t0_) / 2 + (t1_ + t0_) / 2);
        return {z.real(), z.imag()};
    }
    Point2 operator[](double s) const override {
        const std::complex<double> z = s_->eval(s * (t1_ - t0_) + t0_);
        return {z.real(), z.imag()};
    }
    Point2 Derivative_01(double s) const override {
        const std::complex<double> z = (t1_ - t0_) * s_->eval_der(s * (t1_ - t0_) + t0_);
        return {z.real(), z.imag()};
    }
    PanelVector split(unsigned N) const {
        PanelVector out;
        for (unsigned i = 0; i < N; ++i) {
            double a = t0_ + i * (t1_ - t0_) / N, b = t0_ + (i + 1) * (t1_ - t0_) / N;
            if (i == N - 1) b = t1_;
            out.push_back(std::make_shared<ParametrizedPolynomial>(s_, a, b));
        }
        return out;
    }
    double length() const { return len_; }
    void pack(double* o, int off) const override {
        double v[8] = {3, t0_, t1_, (double)off, (double)s_->pieces(), 0, 0, 0};
        std::memcpy(o, v, sizeof(v));
    }
    bool isLineSegment() const override { return false; }
    const std::vector<double>* coefficients() const override { return rows_.get(); }
    double tmin() const { return t0_; }
    double tmax() const { return t1_; }

  private:
    std::shared_ptr<const ComplexSpline> s_;
    double t0_, t1_, len_;
    std::shared_ptr<const std::vector<double>> rows_;
};

}  // namespace hbem
