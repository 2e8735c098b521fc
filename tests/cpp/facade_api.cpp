// Host-only use of the facade's API surface that needs no device: the vector
// randGaussian overload, assembleSparse's declaration, and the
// randomized_svd timers before any device work.  Prints one line per check.
#include <cmath>
#include <cstdio>
#include <map>
#include <vector>

#include "hbem/hbem.hpp"

int main() {
    const hbem::VectorXcd v = randomized_svd::randGaussian(7);
    const hbem::MatrixXcd M = randomized_svd::randGaussian(7, 1, 0);
    double err = 0;
    for (int i = 0; i < 7; ++i) err = std::max(err, std::abs(v(i) - M(i, 0) * std::sqrt(2.0)));
    std::printf("randGaussian_vector %ld %.3e\n", (long)v.rows(), err);
    // assembleSparse is declared with the reference's signature (it computes nothing there either)
    using SM = hbem::SparseMatrix<std::complex<double>>;
    void (hbem::GalerkinBuilder::*fn)(const std::complex<double>&, double, std::map<hbem::LayerType, SM>&, size_t,
                                       size_t, size_t, size_t, std::vector<size_t>&, std::vector<size_t>&, size_t,
                                       size_t, hbem::GalerkinBuilder::workspace_t&) = &hbem::GalerkinBuilder::assembleSparse;
    std::printf("assembleSparse %d\n", fn != nullptr ? 1 : 0);
    randomized_svd::benchmarking(true);
    randomized_svd::reset_timer();
    std::printf("timers %g %g %g %g\n", randomized_svd::get_lu_time(), randomized_svd::get_qr_time(),
                randomized_svd::get_sub_iter_time(), randomized_svd::get_svd_time());
    randomized_svd::benchmarking(false);
    return 0;
}
