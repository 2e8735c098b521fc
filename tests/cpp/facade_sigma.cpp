// Reference-style driver on the C++ facade: the 48-panel square of the reference's small CPU case
// (roots_brent_square_rsvd.cpp:79-171 pattern) -- mesh from ParametrizedLine::split,
// BuilderData, SolutionsOperator::gen_sol_op, randomized_svd::sv; also
// GalerkinBuilder getters and gen_sol_op_1st_der + sv_der.
// Prints: sigma, sigma', and |V(0,1)| for the Python test to compare.
#include <cstdio>
#include <cstdlib>

#include "hbem/hbem.hpp"

using namespace hbem;

int main(int argc, char** argv) {
    const double eps = 0.25;
    const unsigned per_side = argc > 1 ? (unsigned)std::atoi(argv[1]) : 12;
    const unsigned order = argc > 2 ? (unsigned)std::atoi(argv[2]) : 11;
    ParametrizedLine l1({-eps, -eps}, {eps, -eps}), l2({eps, -eps}, {eps, eps}), l3({eps, eps}, {-eps, eps}),
        l4({-eps, eps}, {-eps, -eps});
    PanelVector panels;
    for (auto* l : {&l1, &l2, &l3, &l4}) {
        PanelVector p = l->split(per_side);
        panels.insert(panels.end(), p.begin(), p.end());
    }
    ParametrizedMesh mesh(panels);
    ContinuousSpace<1> cont_space;
    try {
        BuilderData bd(mesh, cont_space, cont_space, order);
        SolutionsOperator so(bd);
        GalerkinBuilder builder(bd);
        MatrixXcd T, Td;
        so.gen_sol_op(2.0, 1.0, 20.0, T);
        MatrixXcd R = randomized_svd::randGaussian((int)(2 * mesh.getNumPanels()), 2, 0);
        double s = randomized_svd::sv(T, R, 2);
        so.gen_sol_op_1st_der(2.0, 1.0, 20.0, T, Td);
        Vector2d der = randomized_svd::sv_der(T, Td, R, 30);
        builder.assembleDense(2.0, 20.0, LayerType::ALL, 1);
        std::printf("%.17g %.17g %.17g\n", s, der(1), std::abs(builder.getSingleLayer(0)(0, 1)));
        try {
            builder.getDoubleLayer(2);
        } catch (const std::runtime_error& e) {
            std::printf("%s\n", e.what());
        }
    } catch (const std::exception& e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
    return 0;
}
