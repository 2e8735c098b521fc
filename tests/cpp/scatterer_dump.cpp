// Prints the panels of scatterer::make_N_polygonal_panels /
// make_polygonal_panels for a scatterer file (one panel per line, endpoints
// with 17 significant digits), for comparison with helmholtzbem_amd.scatterer.
//   scatterer_dump FILE N MODE [RF]     MODE 0: make_N, 1: graded with factor RF,
//                                       2: SmoothScatterer::panelize (tmin, tmax, length per
//                                          panel, then the spline's table rows)
#include <cstdio>
#include <cstdlib>

#include "hbem/scatterer.hpp"

int main(int argc, char** argv) {
    if (argc < 4) return 2;
    std::vector<double> x, y;
    if (!scatterer::read(argv[1], x, y)) {
        std::printf("READ_FAILED\n");
        return 0;
    }
    const unsigned N = (unsigned)std::atoi(argv[2]);
    const int mode = std::atoi(argv[3]);
    if (mode == 2) {
        scatterer::SmoothScatterer sc(argv[1]);
        hbem::PanelVector p = sc.panelize(N);
        std::printf("%zu %.17g\n", p.size(), sc.total_length());
        for (const auto& q : p) {
            const auto& sp = static_cast<const hbem::ParametrizedPolynomial&>(*q);
            std::printf("%.17g %.17g %.17g\n", sp.tmin(), sp.tmax(), sp.length());
        }
        const std::vector<double>& rows = *sc.spline()->rows();
        for (size_t r = 0; r + 2 < rows.size(); r += 3) std::printf("%.17g %.17g %.17g\n", rows[r], rows[r + 1], rows[r + 2]);
        return 0;
    }
    const double rf = argc > 4 ? std::atof(argv[4]) : 0.5;
    hbem::PanelVector p = mode == 0 ? scatterer::make_N_polygonal_panels(x, y, N)
                                    : scatterer::make_polygonal_panels(x, y, N, rf);
    std::printf("%zu %.17g\n", p.size(), scatterer::length(x, y));
    double v[8];
    for (const auto& q : p) {
        q->pack(v, 0);
        std::printf("%.17g %.17g %.17g %.17g\n", v[1], v[2], v[3], v[4]);
    }
    return 0;
}
