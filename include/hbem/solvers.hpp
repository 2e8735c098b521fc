// solvers.hpp -- the transmission solve of the C++ facade (header-only host
// code over hbem_tp_A_rhs / hbem_tp_solve): tp::direct_second_kind::A_rhs and
// ::solve with the reference's signatures (solvers.hpp:73-113,
// solvers.cpp:104-152).
//
//   A_rhs  A = sol_op.gen_sol_op(k, c_o, c_i) (projected when sol_op projects)
//          and rhs = A_o u_inc_N, A_o = [[-K_o + M/2, V_o], [W_o, K_o^T + M/2]],
//          u_inc_N = [Interpolate_helmholtz(u_inc); Interpolate_helmholtz_neu(u_inc_del)]
//   solve  the traces (Dirichlet, then Neumann coefficients of the total field)
//          of A x = rhs with the unprojected operator; throws the reference's
//          std::runtime_error messages for non-finite data and a singular LU.
//
// The interpolation runs on the host through the caller's functions; the
// assembly, A_o u_inc_N and the LU solve run on the device.
#pragma once

#include <complex>
#include <functional>
#include <vector>

#include "hbem.hpp"

namespace tp {
namespace direct_second_kind {

namespace detail {
// [Interpolate_helmholtz(u_inc); Interpolate_helmholtz_neu(u_inc_del)] (solvers.cpp:124-127)
inline std::vector<std::complex<double>> u_inc_N(const hbem::ParametrizedMesh& mesh,
                                                 const std::function<std::complex<double>(double, double)>& u_inc,
                                                 const std::function<hbem::Vector2cd(double, double)>& u_inc_del) {
    const hbem::ContinuousSpace<1> space;
    const hbem::VectorXcd d = space.Interpolate_helmholtz(u_inc, mesh), n = space.Interpolate_helmholtz_neu(u_inc_del, mesh);
    std::vector<std::complex<double>> U((size_t)(d.size() + n.size()));
    for (long i = 0; i < (long)d.size(); ++i) U[(size_t)i] = d(i);
    for (long i = 0; i < (long)n.size(); ++i) U[(size_t)(d.size() + i)] = n(i);
    return U;
}
}  // namespace detail

inline void A_rhs(const hbem::ParametrizedMesh& mesh, const std::function<std::complex<double>(double, double)> u_inc,
                  const std::function<hbem::Vector2cd(double, double)> u_inc_del, const std::complex<double>& k,
                  const double c_o, const double c_i, hbem::SolutionsOperator& sol_op, hbem::MatrixXcd& A,
                  hbem::VectorXcd& rhs) {
    const std::vector<std::complex<double>> U = detail::u_inc_N(mesh, u_inc, u_inc_del);
    const long n = (long)U.size();
    A.resize(n, n);
    rhs.resize(n, 1);
    sol_op.transmission_system(k, c_o, c_i, U.data(), 1, A.data(), rhs.data());
}

inline hbem::VectorXcd solve(const hbem::ParametrizedMesh& mesh,
                             const std::function<std::complex<double>(double, double)> u_inc,
                             const std::function<hbem::Vector2cd(double, double)> u_inc_del, const unsigned order,
                             const std::complex<double>& k, const double c_o, const double c_i) {
    hbem::ContinuousSpace<1> test_space;
    hbem::BuilderData builder_data(mesh, test_space, test_space, order);
    hbem::SolutionsOperator sol_op(builder_data, false);
    const std::vector<std::complex<double>> U = detail::u_inc_N(mesh, u_inc, u_inc_del);
    hbem::VectorXcd sol((long)U.size(), 1);
    sol_op.transmission_solve(k, c_o, c_i, U.data(), 1, sol.data());
    return sol;
}

}  // namespace direct_second_kind
}  // namespace tp
