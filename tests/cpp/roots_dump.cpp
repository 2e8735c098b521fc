// Runs zbrent / rtsafe of include/hbem/find_roots.hpp on fixed test functions
// and prints root, root_found, num_iter per case (17 significant digits), then
// brent_gsl on fixed minimisation cases: x_minimum and iterations per line.
#include <cmath>
#include <cstdio>

#include "hbem/find_roots.hpp"

int main() {
    struct Case { double (*f)(double); double (*df)(double); double a, b, tol; };
    const Case cases[] = {
        {[](double x) { return std::cos(x) - x; }, [](double x) { return -std::sin(x) - 1.0; }, 0.0, 1.0, 1e-12},
        {[](double x) { return x * x * x - 2.0 * x - 5.0; }, [](double x) { return 3.0 * x * x - 2.0; }, 2.0, 3.0, 1e-14},
        {[](double x) { return std::tanh(3.0 * (x - 0.3)); }, [](double x) { double t = std::tanh(3.0 * (x - 0.3)); return 3.0 * (1.0 - t * t); }, -4.0, 5.0, 1e-10},
        {[](double x) { return std::exp(x) - 10.0; }, [](double x) { return std::exp(x); }, 0.0, 4.0, 1e-9},
        {[](double x) { return x * x + 1.0; }, [](double x) { return 2.0 * x; }, -1.0, 2.0, 1e-9},
    };
    for (const Case& c : cases) {
        bool rf;
        unsigned it = 0;
        double r = zbrent(c.f, c.a, c.b, c.tol, rf, it);
        std::printf("%.17g %d %u\n", r, (int)rf, it);
        auto both = [&](double x) { hbem::MatrixXd m(1, 2); m(0, 0) = c.f(x); m(0, 1) = c.df(x); return m; };
        it = 0;
        r = rtsafe(c.f, both, c.a, c.b, c.tol, rf, it);
        std::printf("%.17g %d %u\n", r, (int)rf, it);
    }
    struct MinCase { double (*f)(double); double a, b, g; };
    const MinCase mins[] = {
        {[](double x) { return (x - 2.0) * (x - 2.0) + 1.0; }, 1.0, 3.5, 2.4},
        {[](double x) { return std::cos(2.0 * x); }, 0.0, 2.0, 1.5},
        {[](double x) { return std::fabs(x - 0.1234) + 1.0; }, -1.0, 1.0, 0.3},
        {[](double x) { return std::exp(x) - 3.0 * x; }, 0.0, 3.0, 1.0},
    };
    for (const MinCase& c : mins) {
        std::function<double(double)> f = c.f;
        unsigned ic = 0;
        const double m = brent_gsl(c.a, c.b, c.g, f, 1e-7, ic);
        std::printf("%.17g %u\n", m, ic);
    }
    // rtsafe's NaN-start mode: f(x1) is NaN, so fct(NaN) supplies the first guess
    {
        auto f = [](double x) { return x != x ? 1.25 : (x == 1.0 ? std::nan("") : std::sin(x) - 0.9); };
        auto both = [&](double x) { hbem::MatrixXd m(1, 2); m(0, 0) = f(x); m(0, 1) = std::cos(x); return m; };
        bool rf;
        unsigned it = 0;
        const double r = rtsafe(f, both, 1.0, 2.0, 1e-13, rf, it);
        std::printf("%.17g %d %u\n", r, (int)rf, it);
    }
    return 0;
}
