// Reference-style driver of the transmission solve on the C++ facade
// (transmission_problem_verification.cpp pattern: a circle split into
// ParametrizedCircularArc panels, incoming waves from a wave file,
// tp::direct_second_kind::solve).
//   transmission_driver waves <file> <k>            wave values and gradients
//   transmission_driver solve <file> <N> <k> <c_o> <c_i>   the 2N traces
// Prints one complex number per line as "re im".
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "hbem/incoming.hpp"
#include "hbem/solvers.hpp"

using namespace hbem;

int main(int argc, char** argv) {
    if (argc < 4) return 2;
    incoming::wave w;
    if (!incoming::load(argv[2], w)) {
        std::printf("load failed\n");
        return 1;
    }
    try {
        if (!std::strcmp(argv[1], "waves")) {
            const double k = std::atof(argv[3]);
            const Point2 pts[3] = {{0.31, -0.22}, {-0.13, 0.47}, {0.05, 0.06}};
            for (const Point2& p : pts) {
                const std::complex<double> u = incoming::compute(w, p, k);
                const Vector2cd g = incoming::compute_del(w, p, k);
                std::printf("%.17g %.17g\n%.17g %.17g\n%.17g %.17g\n", u.real(), u.imag(), g(0).real(), g(0).imag(),
                            g(1).real(), g(1).imag());
            }
            return 0;
        }
        const unsigned N = (unsigned)std::atoi(argv[3]);
        const double k = std::atof(argv[4]), c_o = std::atof(argv[5]), c_i = std::atof(argv[6]);
        ParametrizedCircularArc curve({0.0, 0.0}, 0.25, 0.0, 2 * M_PI);
        ParametrizedMesh mesh(curve.split(N));
        const double ko = std::sqrt(c_o) * k;
        auto u_inc = [&](double x1, double x2) { return incoming::compute(w, {x1, x2}, ko); };
        auto u_inc_del = [&](double x1, double x2) { return incoming::compute_del(w, {x1, x2}, ko); };
        VectorXcd sol = tp::direct_second_kind::solve(mesh, u_inc, u_inc_del, 11, k, c_o, c_i);
        for (long i = 0; i < (long)sol.size(); ++i) std::printf("%.17g %.17g\n", sol(i).real(), sol(i).imag());
    } catch (const std::exception& e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
    return 0;
}
