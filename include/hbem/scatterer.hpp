// scatterer.hpp -- scatterer input files and polygonal meshing for the C++
// facade (header-only, host code).
//
// Restates the reference's scatterer namespace (scatterer.hpp:39-92,
// scatterer.cpp:29-262) over std::vector<double> vertex lists:
//   read_array / read      the data/scatterer_*.dat format: comma-separated
//                          numbers per line, whitespace removed, empty lines
//                          skipped, any non-number makes the read fail
//   length                 polygon perimeter
//   auto_num_panels        round(perimeter / (shortest edge * f))
//   make_N_polygonal_panels  exactly N panels, uniform per edge
//   make_polygonal_panels    about N panels graded toward the corners
//   SmoothScatterer          periodic Akima spline through the points, split
//                            into panels of equal arc length (panelize)
// The arithmetic is sequential and in the same order as helmholtzbem_amd/
// scatterer.py, so both produce the same panels bit for bit (compile with
// -ffp-contract=off).
#pragma once

#include <algorithm>
#include <cmath>
#include <fstream>
#include <limits>
#include <sstream>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "hbem.hpp"
#include "spline.hpp"

namespace scatterer {

using hbem::PanelVector;
using hbem::ParametrizedLine;
using hbem::Point2;

// Numeric table of a scatterer or incoming-wave file (scatterer.cpp:29-72);
// rows shorter than the widest row are padded with zeros.
inline bool read_array(const std::string& fname, std::vector<std::vector<double>>& rows, size_t& ncols) {
    std::ifstream in(fname);
    if (!in) return false;
    rows.clear();
    ncols = 0;
    std::string line, tok;
    while (std::getline(in, line)) {
        std::string s;
        for (char ch : line)
            if (!std::isspace((unsigned char)ch)) s.push_back(ch);
        if (s.empty()) continue;
        std::vector<double> row;
        std::istringstream iss(s);
        while (std::getline(iss, tok, ',')) {
            double v;
            try {
                v = std::stod(tok);   // the longest numeric prefix, as the reference reads it
            } catch (const std::exception&) {
                return false;
            }
            row.push_back(v);
        }
        ncols = std::max(ncols, row.size());
        rows.push_back(row);
    }
    for (auto& r : rows) r.resize(ncols, 0.0);
    return !rows.empty();
}

// Polygon vertices of a scatterer file (two columns), scaled (scatterer.cpp:74-81).
inline bool read(const std::string& fname, std::vector<double>& x, std::vector<double>& y, double scale = 1.0) {
    std::vector<std::vector<double>> rows;
    size_t nc = 0;
    if (!read_array(fname, rows, nc) || nc != 2) return false;
    x.clear();
    y.clear();
    for (const auto& r : rows) {
        x.push_back(scale * r[0]);
        y.push_back(scale * r[1]);
    }
    return true;
}

namespace detail {
inline double cround(double v) {   // std::round for v >= 0
    const double r = std::floor(v);
    return v - r >= 0.5 ? r + 1.0 : r;
}
inline std::vector<double> edge_lengths(const std::vector<double>& x, const std::vector<double>& y) {
    const size_t n = x.size();
    std::vector<double> d(n);
    for (size_t i = 0; i < n; ++i) {
        const size_t j = i + 1 < n ? i + 1 : 0;
        const double dx = x[j] - x[i], dy = y[j] - y[i];
        d[i] = std::sqrt(dx * dx + dy * dy);
    }
    return d;
}
inline std::shared_ptr<ParametrizedLine> line(double ax, double ay, double bx, double by) {
    return std::make_shared<ParametrizedLine>(Point2{ax, ay}, Point2{bx, by});
}
}  // namespace detail

inline double length(const std::vector<double>& x, const std::vector<double>& y) {
    double L = 0.0;
    for (double v : detail::edge_lengths(x, y)) L += v;
    return L;
}

inline unsigned auto_num_panels(const std::vector<double>& x, const std::vector<double>& y, double f) {
    if (x.empty()) return 0;
    double tot = 0.0, mn = std::numeric_limits<double>::infinity();
    for (double v : detail::edge_lengths(x, y)) {
        tot += v;
        mn = std::min(mn, v);
    }
    if (mn == 0.0) throw std::invalid_argument("polygon has coincident consecutive vertices");
    return (unsigned)detail::cround(tot / (mn * f));
}

// Exactly N panels (scatterer.cpp:114-155): round(N d_i / L) per edge, then
// one panel at a time to / from the edge that keeps var(d_i / m_i) smallest.
inline PanelVector make_N_polygonal_panels(const std::vector<double>& x, const std::vector<double>& y, unsigned N) {
    const size_t n = x.size();
    if (n == 0 || y.size() != n) throw std::invalid_argument("empty polygon");
    const std::vector<double> d = detail::edge_lengths(x, y);
    const double alpha = N / length(x, y);
    std::vector<double> m(n);
    long cs = 0;
    for (size_t i = 0; i < n; ++i) {
        m[i] = detail::cround(alpha * d[i]);
        cs += (long)m[i];
    }
    while (cs != (long)N) {
        const int k = cs > (long)N ? -1 : 1;
        double best = std::numeric_limits<double>::infinity();
        size_t i_min = 0;
        for (size_t i = 0; i < n; ++i) {
            m[i] += k;
            double mean = 0.0;
            std::vector<double> a(n);
            for (size_t t = 0; t < n; ++t) {
                a[t] = m[t] != 0.0 ? d[t] / m[t] : std::numeric_limits<double>::infinity();
                mean += a[t];
            }
            mean /= n;
            double var = 0.0;
            for (size_t t = 0; t < n; ++t) var += (a[t] - mean) * (a[t] - mean);
            var /= n;
            if (var < best) {
                best = var;
                i_min = i;
            }
            m[i] -= k;
        }
        m[i_min] += k;
        cs += k;
    }
    PanelVector res;
    for (size_t i = 0; i < n; ++i) {
        const size_t j = i + 1 < n ? i + 1 : 0;
        PanelVector part = ParametrizedLine({x[i], y[i]}, {x[j], y[j]}).split((unsigned)m[i]);
        res.insert(res.end(), part.begin(), part.end());
    }
    return res;
}

// About N panels graded toward the corners (scatterer.cpp:157-262); the panel
// counts per edge are N d_i / L (the solution of the reference's linear
// system), rounded edge by edge in order of decreasing count.
inline PanelVector make_polygonal_panels(const std::vector<double>& x, const std::vector<double>& y, unsigned N,
                                         double refinement_factor = 0.5) {
    const int n = (int)x.size();
    const double eps = std::numeric_limits<double>::epsilon();
    const std::vector<double> d = detail::edge_lengths(x, y);
    double L = 0.0;
    for (double v : d) L += v;
    std::vector<double> k(n);
    for (int i = 0; i < n; ++i) k[i] = N * d[i] / L;
    std::vector<std::pair<double, int>> order(n);
    for (int i = 0; i < n; ++i) order[i] = {k[i], i};
    std::sort(order.begin(), order.end());
    std::reverse(order.begin(), order.end());
    auto obj = [&]() {
        std::vector<double> q(n);
        double mu = 0.0;
        for (int t = 0; t < n; ++t) {
            q[t] = d[t] / k[t];
            mu += q[t];
        }
        mu /= n;
        double ss = 0.0, ks = 0.0;
        for (int t = 0; t < n; ++t) ss += (q[t] - mu) * (q[t] - mu);
        for (int t = 0; t < n; ++t) ks += k[t];
        return std::sqrt(ss / n) / mu + std::fabs(N / ks - 1.0);
    };
    for (const auto& pr : order) {
        const int j = pr.second;
        k[j] = std::floor(k[j]);
        const double f1 = obj();
        k[j] += 1.0;
        const double f2 = obj();
        if (f1 == f1 && f1 < f2) k[j] -= 1.0;
    }
    auto dot = [](double ax, double ay, double bx, double by) { return ax * bx + ay * by; };
    auto norm = [&](double ax, double ay) { return std::sqrt(dot(ax, ay, ax, ay)); };
    PanelVector res;
    for (int i = 0; i < n; ++i) {
        const int i0 = i > 0 ? i - 1 : n - 1, i2 = i + 1 < n ? i + 1 : 0, i3 = i < n - 2 ? i + 2 : i - n + 2;
        const double p0x = x[i0], p0y = y[i0], p1x = x[i], p1y = y[i], p2x = x[i2], p2y = y[i2], p3x = x[i3],
                     p3y = y[i3];
        const double Dx = (p2x - p1x) / k[i], Dy = (p2y - p1y) / k[i];
        const int n_i = (int)k[i];
        double ux = p0x - p1x, uy = p0y - p1y, vx = p2x - p1x, vy = p2y - p1y;
        double f1 = (1.0 - dot(ux, uy, vx, vy) / (norm(ux, uy) * norm(vx, vy))) / 2.0;
        ux = p1x - p2x; uy = p1y - p2y; vx = p3x - p2x; vy = p3y - p2y;
        double f2 = (1.0 - dot(ux, uy, vx, vy) / (norm(ux, uy) * norm(vx, vy))) / 2.0;
        f1 = refinement_factor + (1.0 - refinement_factor) * f1;
        f2 = refinement_factor + (1.0 - refinement_factor) * f2;
        if (refinement_factor >= 1.0 || n_i < 3) {
            for (int j = 0; j < n_i; ++j) {
                const bool last = j + 1 >= n_i;
                res.push_back(detail::line(p1x + j * Dx, p1y + j * Dy, last ? p2x : p1x + (j + 1) * Dx,
                                           last ? p2y : p1y + (j + 1) * Dy));
            }
            continue;
        }
        int j_best = 0;
        double alpha_best = 0.0, d1_best = 0.0, d2_best = 0.0;
        for (int j = 1; j < n_i; ++j) {
            double alpha, alpha_next = refinement_factor;
            unsigned it = 0;
            do {
                alpha = alpha_next;
                const double fv = f1 * std::pow(alpha, double(1 - j)) + f2 * std::pow(alpha, double(j + 1 - n_i)) -
                                  alpha * (f1 + f2) - n_i * (1.0 - alpha);
                const double fd = n_i - f1 - f2 - f1 * (j - 1.0) * std::pow(alpha, double(-j)) -
                                  f2 * (n_i - j - 1.0) * std::pow(alpha, double(j - n_i));
                alpha_next = alpha - fv / fd;
                ++it;
            } while (std::fabs(alpha_next - alpha) > eps && it <= 5000);
            const double d1 = f1 / std::pow(alpha, double(j - 1)), d2 = f2 / std::pow(alpha, double(n_i - j - 1));
            if (!j_best || std::fabs(d1 - d2) < std::fabs(d1_best - d2_best)) {
                d1_best = d1;
                d2_best = d2;
                j_best = j;
                alpha_best = alpha;
            }
        }
        PanelVector graded;
        double px = p1x, py = p1y;
        for (int j = 0; j < n_i; ++j) {
            const double sc = j < j_best ? d1_best * std::pow(alpha_best, double(j_best - 1 - j))
                                         : d2_best * std::pow(alpha_best, double(j - j_best));
            const double nx = px + Dx * sc, ny = py + Dy * sc;
            const bool last = j + 1 >= n_i;
            graded.push_back(detail::line(px, py, last ? p2x : nx, last ? p2y : ny));
            px = nx;
            py = ny;
        }
        if (norm(px - p2x, py - p2y) / norm(Dx, Dy) > 1e-3) {
            graded.clear();
            px = p1x;
            py = p1y;
            for (int j = 0; j < n_i; ++j) {
                const double nx = px + Dx, ny = py + Dy;
                const bool last = j + 1 >= n_i;
                graded.push_back(detail::line(px, py, last ? p2x : nx, last ? p2y : ny));
                px = nx;
                py = ny;
            }
        }
        res.insert(res.end(), graded.begin(), graded.end());
    }
    return res;
}

// Smooth closed curve through the points of a scatterer file
// (scatterer.hpp:140-250, scatterer.cpp:385-533): periodic Akima spline z(t),
// t in [0, 1], through the points closed by the first one, point i at the
// polygon length up to it over the perimeter.
class SmoothScatterer {
  public:
    explicit SmoothScatterer(const std::string& fname, double scale = 1.0) {
        std::vector<double> x, y;
        if (!read(fname, x, y, scale) || x.size() < 3) throw std::runtime_error("Failed to read data for smooth scatterer");
        const size_t n = x.size();
        x.push_back(x[0]);
        y.push_back(y[0]);
        const double L = length(x, y);
        t_.assign(n + 1, 0.0);
        std::vector<std::complex<double>> z(n + 1);
        for (size_t i = 0; i < n; ++i) {
            const double dx = x[i] - x[i + 1], dy = y[i] - y[i + 1];
            t_[i + 1] = std::sqrt(dx * dx + dy * dy) / L + t_[i];
        }
        for (size_t i = 0; i <= n; ++i) z[i] = {x[i], y[i]};
        spline_ = std::make_shared<hbem::ComplexSpline>(t_, z, true);
    }
    hbem::Point2 curve_point(double t) const {
        const std::complex<double> z = spline_->eval(t);
        return {z.real(), z.imag()};
    }
    double total_length() {
        if (total_ < 0.0) total_ = spline_->arc_length(0.0, t_.back());
        return total_;
    }
    // npanels panels of equal arc length (scatterer.cpp:492-518)
    PanelVector panelize(unsigned npanels) {
        double len = total_length(), t0 = 0.0;
        const double panel_length = len / npanels, tol = 1e-8;
        PanelVector panels;
        for (unsigned i = 1; i < npanels; ++i) {
            double t1 = t0 + (1.0 - t0) / double(npanels - i + 1), L = 0.0;
            for (unsigned ic = 0; ic < 1000; ++ic) {
                L = spline_->arc_length(t0, t1);
                const double t2 = t1 + (panel_length - L) / std::abs(spline_->eval_der(t1));
                if (std::fabs(t2 - t1) / t1 < tol) break;
                t1 = t2;
            }
            panels.push_back(std::make_shared<hbem::ParametrizedPolynomial>(spline_, t0, t1, L));
            t0 = t1;
            len -= L;
        }
        panels.push_back(std::make_shared<hbem::ParametrizedPolynomial>(spline_, t0, 1.0, len));
        return panels;
    }
    std::shared_ptr<const hbem::ComplexSpline> spline() const { return spline_; }

  private:
    std::vector<double> t_;
    std::shared_ptr<hbem::ComplexSpline> spline_;
    double total_ = -1.0;
};

}  // namespace scatterer
