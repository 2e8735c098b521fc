// Reference-style driver on the C++ facade: the 48-panel square of the reference's small CPU case
// (roots_brent_square_rsvd.cpp:79-171 pattern) -- mesh from ParametrizedLine::split,
// BuilderData, SolutionsOperator::gen_sol_op, randomized_svd::sv; also
// GalerkinBuilder getters and gen_sol_op_1st_der + sv_der; then the batched
// sweep of 32 k in FP64 and in mixed precision on a context with the
// separated-pair rules (HBEM_FLAG_FAR_FIELD).
// Prints: sigma, sigma', and |V(0,1)|; the getter's error message; the
// largest relative sigma difference of the two sweeps and the number of k
// the mixed sweep recomputed in FP64.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

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
        const int nk = 32;
        std::vector<double> ks(nk), s64(nk), smx(nk);
        std::vector<int> used(nk);
        for (int i = 0; i < nk; ++i) ks[i] = 2.0 + 8.0 * i / (nk - 1);
        BuilderData bdf(mesh, cont_space, cont_space, order, HBEM_FLAG_FAR_FIELD);
        check(hbem_sigma_sweep(bd.handle(), ks.data(), nk, 1.0, 20.0, R.data(), 2, 2, 0, s64.data(), nullptr));
        check(hbem_sigma_sweep_mixed(bdf.handle(), ks.data(), nk, 1.0, 20.0, R.data(), 2, 2, 2, 1, smx.data(),
                                     used.data()));
        double err = 0.0;
        int nused = 0;
        for (int i = 0; i < nk; ++i) {
            err = std::max(err, std::abs(smx[i] - s64[i]) / s64[i]);
            nused += used[i];
        }
        std::printf("mixed_far %.3e %d\n", err, nused);
    } catch (const std::exception& e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
    return 0;
}
