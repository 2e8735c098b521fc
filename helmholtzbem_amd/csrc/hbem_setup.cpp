// hbem_setup.cpp -- host-side, once-per-geometry setup of the sigma_min(k) path.
//
// Built with -ffp-contract=off: the mass matrix and the projection factor
// decide the LDLT pivot order by exact comparisons of diagonal entries, so
// their arithmetic follows a fixed operation order.
#include <cmath>
#include <complex>
#include <cstring>
#include <stdexcept>
#include <vector>

namespace hbem {

static const double kEpsH = 2.220446049250313080847e-16;
static const double kPiH = 3.14159265358979323846;

// Gauss-Legendre on [a,b]: Newton iteration on the Legendre three-term
// recurrence from the Chebyshev-like initial guess cos(pi (i + 3/4)/(n + 1/2)),
// stopping when a Newton step is <= DBL_EPSILON (gauleg.hpp:35-79, 133-146).
void gauss_rule(double a, double b, int n, double* x, double* w) {
    const int half = (n + 1) / 2;
    const double mid = 0.5 * (a + b), hw = 0.5 * (b - a);
    for (int i = 0; i < half; ++i) {
        double z = std::cos(kPiH * (i + 1 - 0.25) / (n + 0.5)), zold, pn, pn1, dpn;
        do {
            pn = 1.0;
            pn1 = 0.0;
            for (int j = 1; j <= n; ++j) {
                double pn2 = pn1;
                pn1 = pn;
                pn = ((2.0 * j - 1.0) * z * pn1 - (j - 1.0) * pn2) / j;
            }
            dpn = n * (z * pn - pn1) / (z * z - 1.0);
            zold = z;
            z = zold - pn / dpn;
        } while (std::fabs(z - zold) > kEpsH);
        x[i] = mid - hw * z;
        x[n - 1 - i] = mid + hw * z;
        double wi = 2.0 * hw / ((1.0 - z * z) * dpn * dpn);
        w[i] = wi;
        w[n - 1 - i] = wi;
    }
}

// Reduced geometrically graded composite rule on [sigma^(n-1), 1],
// sigma = (sqrt2 - 1)^2, n - k + 1 points on [sigma^k, sigma^(k-1)]
// (gauleg.hpp:89-121, 156-168).
void composite_rule(int n, double* x, double* w) {
    const int total = n * (n + 1) / 2 - 1;
    const double sig = (std::sqrt(2.0) - 1.0) * (std::sqrt(2.0) - 1.0);
    double lo = sig, hi = 1.0;
    int pos = total - 1;
    std::vector<double> gx(n), gw(n);
    for (int lev = 1, npt = n; lev < n; ++lev, --npt) {
        gauss_rule(0.0, 1.0, npt, gx.data(), gw.data());
        for (int j = npt - 1; j >= 0; --j, --pos) {
            w[pos] = (hi - lo) * gw[j];
            x[pos] = (hi - lo) * gx[j] + lo;
        }
        hi = lo;
        lo = lo * sig;
    }
}

// Spline piece at t (rows of 9 doubles per piece, see the device panel_eval):
// point z and dz/dt.
static void host_spline(const double* p, const double* coefs, double t, std::complex<double>& z,
                        std::complex<double>& dz) {
    const double* c = coefs + 3 * (int)p[3];
    int lo = 0, hi = (int)p[4];
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (c[9 * mid] <= t) lo = mid;
        else hi = mid;
    }
    const double* r = c + 9 * lo;
    const double d = t - r[0];
    z = std::complex<double>(r[1] + d * (r[3] + d * (r[4] + d * r[5])), r[2] + d * (r[6] + d * (r[7] + d * r[8])));
    dz = std::complex<double>(r[3] + d * (2.0 * r[4] + 3.0 * d * r[5]), r[6] + d * (2.0 * r[7] + 3.0 * d * r[8]));
}

// Host panel evaluation (same parametrizations as the device panel_eval).
static std::complex<double> host_tangent(const double* p, const double* coefs, double s) {
    typedef std::complex<double> cd;
    switch ((int)p[0]) {
    case 3: {
        cd z, dz;
        host_spline(p, coefs, p[1] + s * (p[2] - p[1]), z, dz);
        return (p[2] - p[1]) * dz;
    }
    case 0: return cd(p[3] - p[1], p[4] - p[2]);
    case 1: {
        double d = p[5] - p[4];
        return cd(0.0, p[3] * d) * std::exp(cd(0.0, s * d + p[4]));
    }
    default: {
        double t = p[1] + s * (p[2] - p[1]);
        cd dz = 0.0;
        int c0 = (int)p[3], nc = (int)p[4];
        for (int k = 0; k < nc; ++k) {
            const double* c = coefs + 3 * (c0 + k);
            dz += cd(0.0, c[0]) * cd(c[1], c[2]) * std::polar(1.0, c[0] * t);
        }
        return (p[2] - p[1]) * dz;
    }
    }
}

static std::complex<double> host_point(const double* p, const double* coefs, double s) {
    typedef std::complex<double> cd;
    switch ((int)p[0]) {
    case 3: {
        cd z, dz;
        host_spline(p, coefs, p[1] + s * (p[2] - p[1]), z, dz);
        return z;
    }
    case 0: return cd(p[3] - p[1], p[4] - p[2]) * s + cd(p[1], p[2]);
    case 1: {
        double d = p[5] - p[4];
        return cd(p[1], p[2]) + p[3] * std::exp(cd(0.0, s * d + p[4]));
    }
    default: {
        double t = p[1] + s * (p[2] - p[1]);
        cd z = 0.0;
        int c0 = (int)p[3], nc = (int)p[4];
        for (int k = 0; k < nc; ++k) {
            const double* c = coefs + 3 * (c0 + k);
            z += cd(c[1], c[2]) * std::polar(1.0, c[0] * t);
        }
        return z;
    }
    }
}

// Panel start / end as BuilderData::is_adjacent reads them through
// operator()(-1 / +1) (galerkin_builder.cpp:189-201).
static std::complex<double> host_endpoint(const double* p, const double* coefs, bool end) {
    if ((int)p[0] == 0) {
        double t = end ? 1.0 : -1.0;
        return 0.5 * std::complex<double>(t * (p[3] - p[1]) + (p[3] + p[1]), t * (p[4] - p[2]) + (p[4] + p[2]));
    }
    return host_point(p, coefs, end ? 1.0 : 0.0);
}

// ends[4i..4i+3] = start(i).x, start(i).y, end(i).x, end(i).y
void panel_endpoints(const double* panels, const double* coefs, int N, double* ends) {
    for (int i = 0; i < N; ++i) {
        std::complex<double> a = host_endpoint(panels + 8 * i, coefs, false);
        std::complex<double> b = host_endpoint(panels + 8 * i, coefs, true);
        ends[4 * i] = a.real(); ends[4 * i + 1] = a.imag(); ends[4 * i + 2] = b.real(); ends[4 * i + 3] = b.imag();
    }
}

void panel_bounds(const double* panels, const double* coefs, int N, double* cr) {
    for (int i = 0; i < N; ++i) {
        const double* p = panels + 8 * i;
        const std::complex<double> c = host_point(p, coefs, 0.5);
        double rad = 0.0;
        if ((int)p[0] == 0) {
            rad = 0.5 * std::hypot(p[3] - p[1], p[4] - p[2]);
        } else {
            for (int s = 0; s <= 64; ++s) rad = std::max(rad, std::abs(host_point(p, coefs, s / 64.0) - c));
            rad *= 1.02;
        }
        cr[3 * i] = c.real();
        cr[3 * i + 1] = c.imag();
        cr[3 * i + 2] = rad;
    }
}

// mass_matrix::GalerkinMatrix with the rule chosen by SolutionsOperator
// (mass_matrix.cpp:6-47, gen_sol_op.cpp:30-33).  P1: dof(0,p) = (p+1)%N, dof(1,p) = p;
// shapes b0 = t, b1 = 1 - t; |gamma'| = sqrt(x^2 + y^2).
void mass_matrix_host(const double* panels, int N, const double* coefs, bool lines, int order, double* M) {
    const int n = lines ? 4 : order;
    std::vector<double> x(n), w(n);
    gauss_rule(0.0, 1.0, n, x.data(), w.data());
    std::memset(M, 0, sizeof(double) * (size_t)N * N);
    for (int p = 0; p < N; ++p)
        for (int ii = 0; ii < 2; ++ii)
            for (int jj = 0; jj < 2; ++jj) {
                double sum = 0.0;
                for (int k = 0; k < n; ++k) {
                    std::complex<double> tg = host_tangent(panels + 8 * p, coefs, x[k]);
                    double nrm = std::sqrt(tg.real() * tg.real() + tg.imag() * tg.imag());
                    double F = (jj == 0 ? x[k] : 1.0 - x[k]) * nrm;
                    double G = (ii == 0 ? x[k] : 1.0 - x[k]);
                    double wf = w[k] * F;
                    double term = wf * G;
                    sum = sum + term;
                }
                int r = ii == 0 ? (p + 1) % N : p, c = jj == 0 ? (p + 1) % N : p;
                M[r + (size_t)c * N] += sum;
            }
}

// LDLT with diagonal pivoting in Eigen's left-looking form (the pivot at step
// k is the largest |A(i,i)|, i >= k, of the not-yet-updated diagonal; first
// index on ties), as SolutionsOperator's ctor uses it (gen_sol_op.cpp:34-41).
// Exact zeros of the row being eliminated are skipped, which leaves every sum
// unchanged and makes the cyclic-tridiagonal mass matrices O(n^2).
// Returns F = P^T L sqrt(D) (column-major) and the transpositions.
void ldlt_projection_factor(const double* Ain, int n, double* F, int* tr) {
    std::vector<double> A(Ain, Ain + (size_t)n * n);
    auto a = [&](int i, int j) -> double& { return A[i + (size_t)j * n]; };
    std::vector<double> temp(n);
    std::vector<int> J;
    J.reserve(n);
    for (int k = 0; k < n; ++k) {
        int big = k;
        double bv = std::fabs(a(k, k));
        for (int i = k + 1; i < n; ++i)
            if (std::fabs(a(i, i)) > bv) { bv = std::fabs(a(i, i)); big = i; }
        tr[k] = big;
        if (big != k) {
            for (int j = 0; j < k; ++j) std::swap(a(k, j), a(big, j));
            for (int i = big + 1; i < n; ++i) std::swap(a(i, k), a(i, big));
            std::swap(a(k, k), a(big, big));
            for (int i = k + 1; i < big; ++i) {
                double tmp = a(i, k);
                a(i, k) = a(big, i);
                a(big, i) = tmp;
            }
        }
        if (k > 0) {
            J.clear();
            for (int j = 0; j < k; ++j)
                if (a(k, j) != 0.0) J.push_back(j);
            for (int j : J) temp[j] = a(j, j) * a(k, j);
            double s = 0.0;
            for (int j : J) s += a(k, j) * temp[j];
            a(k, k) -= s;
            if (!J.empty())
                for (int i = k + 1; i < n; ++i) {
                    double u = 0.0;
                    for (int j : J) u += a(i, j) * temp[j];
                    a(i, k) -= u;
                }
        }
        double akk = a(k, k);
        if (k + 1 < n && std::fabs(akk) > 0.0)
            for (int i = k + 1; i < n; ++i) a(i, k) /= akk;
    }
    std::memset(F, 0, sizeof(double) * (size_t)n * n);
    for (int j = 0; j < n; ++j) {
        double sd = std::sqrt(a(j, j));
        F[j + (size_t)j * n] = sd;
        for (int i = j + 1; i < n; ++i) {
            double l = a(i, j);
            if (l != 0.0) F[i + (size_t)j * n] = l * sd;
        }
    }
    for (int k = n - 1; k >= 0; --k)
        if (tr[k] != k)
            for (int j = 0; j < n; ++j) std::swap(F[k + (size_t)j * n], F[tr[k] + (size_t)j * n]);
}

// Dense inverse by column-wise substitution through the factor structure:
// F = P^T L S with S = sqrt(D) diag.  F^{-1} e_c = S^{-1} L^{-1} P e_c.
// L is recovered from F (rows permuted back, columns scaled back); its
// columns are sparse, so each column solve costs O(nnz(L)).
void invert_projection_factor(const double* F, const int* tr, int n, double* Finv) {
    // Undo P^T on rows: L S = P F  (apply T_0 .. T_{n-1} to rows of F)
    std::vector<double> LS(F, F + (size_t)n * n);
    for (int k = 0; k < n; ++k)
        if (tr[k] != k)
            for (int j = 0; j < n; ++j) std::swap(LS[k + (size_t)j * n], LS[tr[k] + (size_t)j * n]);
    std::vector<double> sd(n);
    std::vector<std::vector<std::pair<int, double>>> col(n);
    for (int j = 0; j < n; ++j) {
        sd[j] = LS[j + (size_t)j * n];
        for (int i = j + 1; i < n; ++i) {
            double v = LS[i + (size_t)j * n];
            if (v != 0.0) col[j].push_back({i, v / sd[j]});
        }
    }
    std::memset(Finv, 0, sizeof(double) * (size_t)n * n);
    std::vector<double> x(n);
    for (int c = 0; c < n; ++c) {
        int pos = c;
        for (int k = 0; k < n; ++k) {
            if (pos == k) pos = tr[k];
            else if (pos == tr[k]) pos = k;
        }
        std::fill(x.begin(), x.end(), 0.0);
        x[pos] = 1.0;
        for (int k = pos; k < n; ++k) {
            double xk = x[k];
            if (xk == 0.0) continue;
            for (auto& e : col[k]) x[e.first] -= e.second * xk;
        }
        for (int i = 0; i < n; ++i) Finv[i + (size_t)c * n] = x[i] / sd[i];
    }
}

}  // namespace hbem
