// hbem.hpp -- C++ facade of the MI355X sigma_min(k) path with the reference's
// class names, argument order, defaults and error messages (HelmholtzBEM's
// galerkin_builder.hpp, gen_sol_op.hpp, randsvd.hpp, parametrized_*.hpp,
// continuous_space.hpp), implemented over the C-ABI of hbem.h.
//
// Matrix types: Eigen::MatrixXcd / MatrixXd / Vector2d / Vector3d when
// <Eigen/Dense> is available (the reference's own types, column-major, so no
// copies beyond the device transfer), otherwise the minimal column-major
// hbem::Matrix below with the members the path uses.
#pragma once

#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <functional>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../hbem.h"

#if defined(__has_include)
#if __has_include(<Eigen/Dense>)
#include <Eigen/Dense>
#include <Eigen/Sparse>
#define HBEM_HAVE_EIGEN 1
#endif
#endif

namespace hbem {

#ifdef HBEM_HAVE_EIGEN
using MatrixXcd = Eigen::MatrixXcd;
using MatrixXd = Eigen::MatrixXd;
using VectorXcd = Eigen::VectorXcd;
using Vector2d = Eigen::Vector2d;
using Vector3d = Eigen::Vector3d;
using Vector2cd = Eigen::Vector2cd;
template <class T>
using SparseMatrix = Eigen::SparseMatrix<T>;
#else
template <class T>
class Matrix {
  public:
    Matrix() = default;
    Matrix(long r, long c) { resize(r, c); }
    void resize(long r, long c) { r_ = r; c_ = c; d_.assign((size_t)(r * c), T(0)); }
    void setZero(long r, long c) { resize(r, c); }
    long rows() const { return r_; }
    long cols() const { return c_; }
    long size() const { return r_ * c_; }
    T* data() { return d_.data(); }
    const T* data() const { return d_.data(); }
    T& operator()(long i, long j) { return d_[(size_t)(i + j * r_)]; }
    const T& operator()(long i, long j) const { return d_[(size_t)(i + j * r_)]; }
    T& operator()(long i) { return d_[(size_t)i]; }
    const T& operator()(long i) const { return d_[(size_t)i]; }

  private:
    long r_ = 0, c_ = 0;
    std::vector<T> d_;
};
using MatrixXcd = Matrix<std::complex<double>>;
using MatrixXd = Matrix<double>;
using VectorXcd = Matrix<std::complex<double>>;
// the element type of GalerkinBuilder::assembleSparse's result map
template <class T>
struct SparseMatrix {
    long rows = 0, cols = 0;
};
struct Vector2d { double v[2] = {0, 0}; double& operator()(int i) { return v[i]; } double operator()(int i) const { return v[i]; } };
struct Vector3d { double v[3] = {0, 0, 0}; double& operator()(int i) { return v[i]; } double operator()(int i) const { return v[i]; } };
struct Vector2cd {
    std::complex<double> v[2] = {0.0, 0.0};
    Vector2cd() = default;
    Vector2cd(std::complex<double> a, std::complex<double> b) : v{a, b} {}
    std::complex<double>& operator()(int i) { return v[i]; }
    std::complex<double> operator()(int i) const { return v[i]; }
};
#endif

inline void check(int rc) {
    if (rc != HBEM_OK) {
        char buf[1024];
        hbem_last_error(buf, sizeof(buf));
        throw std::runtime_error(buf);
    }
}

// ---------------------------------------------------------------------------
// Panels and mesh (abstract_parametrized_curve.hpp, parametrized_line.cpp,
// parametrized_circular_arc.cpp, parametrized_mesh.hpp)
// ---------------------------------------------------------------------------
struct Point2 { double x, y; };

class AbstractParametrizedCurve {
  public:
    virtual ~AbstractParametrizedCurve() = default;
    // point at s in [0, 1] and its derivative in s (abstract_parametrized_curve.hpp:72, 92)
    virtual Point2 operator[](double s) const = 0;
    virtual Point2 Derivative_01(double s) const = 0;
    virtual void pack(double* out8, int coef_offset) const = 0;
    virtual bool isLineSegment() const = 0;
    virtual const std::vector<double>* coefficients() const { return nullptr; }
};
using PanelVector = std::vector<std::shared_ptr<AbstractParametrizedCurve>>;

class ParametrizedLine : public AbstractParametrizedCurve {
  public:
    ParametrizedLine(Point2 first, Point2 second) : a_(first), b_(second) {}
    // operator() on [-1, 1] (parametrized_line.cpp:23-29)
    Point2 operator()(double t) const {
        return {0.5 * (t * (b_.x - a_.x) + (b_.x + a_.x)), 0.5 * (t * (b_.y - a_.y) + (b_.y + a_.y))};
    }
    // parametrized_line.cpp:36-42, 78-84
    Point2 operator[](double s) const override { return {s * (b_.x - a_.x) + a_.x, s * (b_.y - a_.y) + a_.y}; }
    Point2 Derivative_01(double) const override { return {b_.x - a_.x, b_.y - a_.y}; }
    // parametrized_line.cpp:116-137
    PanelVector split(unsigned N) const {
        PanelVector out;
        for (unsigned i = 0; i < N; ++i) {
            double t1 = -1.0 + i * 2.0 / N, t2 = -1.0 + (i + 1) * 2.0 / N;
            if (i == N - 1) t2 = 1.0;
            out.push_back(std::make_shared<ParametrizedLine>((*this)(t1), (*this)(t2)));
        }
        return out;
    }
    void pack(double* o, int) const override {
        double v[8] = {0, a_.x, a_.y, b_.x, b_.y, 0, 0, 0};
        std::memcpy(o, v, sizeof(v));
    }
    bool isLineSegment() const override { return true; }

  private:
    Point2 a_, b_;
};

class ParametrizedCircularArc : public AbstractParametrizedCurve {
  public:
    ParametrizedCircularArc(Point2 center, double r, double phi_start, double phi_end)
        : c_(center), r_(r), p0_(phi_start), p1_(phi_end) {}
    // parametrized_circular_arc.cpp:46-53, 92-99
    Point2 operator[](double s) const override {
        const double a = s * (p1_ - p0_) + p0_;
        return {c_.x + r_ * std::cos(a), c_.y + r_ * std::sin(a)};
    }
    Point2 Derivative_01(double s) const override {
        const double d = p1_ - p0_, a = s * d + p0_;
        return {-r_ * d * std::sin(a), r_ * d * std::cos(a)};
    }
    // parametrized_circular_arc.cpp:134-148
    PanelVector split(unsigned N) const {
        PanelVector out;
        for (unsigned i = 0; i < N; ++i) {
            double a = p0_ + i * (p1_ - p0_) / N, b = p0_ + (i + 1) * (p1_ - p0_) / N;
            if (i == N - 1) b = p1_;
            out.push_back(std::make_shared<ParametrizedCircularArc>(c_, r_, a, b));
        }
        return out;
    }
    void pack(double* o, int) const override {
        double v[8] = {1, c_.x, c_.y, r_, p0_, p1_, 0, 0};
        std::memcpy(o, v, sizeof(v));
    }
    bool isLineSegment() const override { return false; }

  private:
    Point2 c_;
    double r_, p0_, p1_;
};

// z(t) = sum_m c_m e^{imt}, t in [t0, t1]; coefficients as {m, re, im} triples.
class ParametrizedTrigCurve : public AbstractParametrizedCurve {
  public:
    ParametrizedTrigCurve(std::shared_ptr<const std::vector<double>> coefs, double t0, double t1)
        : co_(std::move(coefs)), t0_(t0), t1_(t1) {}
    // z(t0 + s (t1 - t0)) and its derivative in s
    Point2 operator[](double s) const override {
        const double t = t0_ + s * (t1_ - t0_);
        std::complex<double> z = 0.0;
        for (size_t i = 0; i + 2 < co_->size(); i += 3)
            z += std::complex<double>((*co_)[i + 1], (*co_)[i + 2]) * std::polar(1.0, (*co_)[i] * t);
        return {z.real(), z.imag()};
    }
    Point2 Derivative_01(double s) const override {
        const double t = t0_ + s * (t1_ - t0_);
        std::complex<double> z = 0.0;
        for (size_t i = 0; i + 2 < co_->size(); i += 3)
            z += std::complex<double>(0.0, (*co_)[i]) * std::complex<double>((*co_)[i + 1], (*co_)[i + 2]) *
                 std::polar(1.0, (*co_)[i] * t);
        z *= (t1_ - t0_);
        return {z.real(), z.imag()};
    }
    PanelVector split(unsigned N) const {
        PanelVector out;
        for (unsigned i = 0; i < N; ++i) {
            double a = t0_ + i * (t1_ - t0_) / N, b = t0_ + (i + 1) * (t1_ - t0_) / N;
            if (i == N - 1) b = t1_;
            out.push_back(std::make_shared<ParametrizedTrigCurve>(co_, a, b));
        }
        return out;
    }
    void pack(double* o, int off) const override {
        double v[8] = {2, t0_, t1_, (double)off, (double)(co_->size() / 3), 0, 0, 0};
        std::memcpy(o, v, sizeof(v));
    }
    bool isLineSegment() const override { return false; }
    const std::vector<double>* coefficients() const override { return co_.get(); }

  private:
    std::shared_ptr<const std::vector<double>> co_;
    double t0_, t1_;
};

typedef enum MeshType { MESH_TYPE_GENERAL, MESH_TYPE_POLYGONAL, MESH_TYPE_SMOOTH } mesh_type;

class ParametrizedMesh {
  public:
    ParametrizedMesh(PanelVector panels, mesh_type type = MESH_TYPE_GENERAL, void* data = nullptr)
        : panels_(std::move(panels)), type_(type), data_(data) {}
    const PanelVector& getPanels() const { return panels_; }
    unsigned getNumPanels() const { return (unsigned)panels_.size(); }
    bool isPolygonal() const {
        if (type_ == MESH_TYPE_POLYGONAL) return true;
        if (type_ != MESH_TYPE_GENERAL) return false;
        for (auto& p : panels_) if (!p->isLineSegment()) return false;
        return true;
    }
    void* getData() const { return data_; }
    // packed form for the C-ABI
    void pack(std::vector<double>& pan, std::vector<double>& coefs) const {
        pan.assign(8 * panels_.size(), 0.0);
        coefs.clear();
        std::vector<const std::vector<double>*> seen;
        std::vector<int> offs;
        for (size_t i = 0; i < panels_.size(); ++i) {
            int off = 0;
            if (const std::vector<double>* c = panels_[i]->coefficients()) {
                size_t k = 0;
                for (; k < seen.size(); ++k) if (seen[k] == c) break;
                if (k == seen.size()) {
                    seen.push_back(c);
                    offs.push_back((int)(coefs.size() / 3));
                    coefs.insert(coefs.end(), c->begin(), c->end());
                }
                off = offs[k];
            }
            panels_[i]->pack(&pan[8 * i], off);
        }
    }

  private:
    PanelVector panels_;
    mesh_type type_;
    void* data_;
};

// ContinuousSpace<1> (continuous_space.hpp:33-152)
class AbstractBEMSpace {
  public:
    virtual ~AbstractBEMSpace() = default;
    virtual unsigned LocGlobMap(unsigned q, unsigned n, unsigned N) const = 0;
    virtual unsigned getSpaceDim(unsigned numpanels) const = 0;
    int getQ() const { return q_; }

  protected:
    int q_ = 0;
};

template <unsigned p> class ContinuousSpace;
template <> class ContinuousSpace<1> : public AbstractBEMSpace {
  public:
    ContinuousSpace() { q_ = 2; }
    unsigned LocGlobMap(unsigned q, unsigned n, unsigned N) const override {
        if (!(q <= (unsigned)q_ && n <= N)) throw std::out_of_range("Panel/RSF index out of range!");
        return q == 2 ? n : ((n % N == 0) ? 1 : (n + 1));
    }
    unsigned getSpaceDim(unsigned numpanels) const override { return numpanels * (q_ - 1); }
    // values of func at the panel start points (continuous_space.hpp:67-82)
    VectorXcd Interpolate_helmholtz(const std::function<std::complex<double>(double, double)>& func,
                                    const ParametrizedMesh& mesh) const {
        const unsigned N = mesh.getNumPanels();
        VectorXcd c(getSpaceDim(N), 1);
        for (unsigned i = 0; i < N; ++i) {
            const Point2 x = (*mesh.getPanels()[i])[0.0];
            c(i) = func(x.x, x.y);
        }
        return c;
    }
    // n . func at the panel start points, n = (t_y, -t_x) / |t| (continuous_space.hpp:84-101)
    VectorXcd Interpolate_helmholtz_neu(const std::function<Vector2cd(double, double)>& func,
                                        const ParametrizedMesh& mesh) const {
        const unsigned N = mesh.getNumPanels();
        VectorXcd c(getSpaceDim(N), 1);
        for (unsigned i = 0; i < N; ++i) {
            const AbstractParametrizedCurve& p = *mesh.getPanels()[i];
            const Point2 x = p[0.0], t = p.Derivative_01(0.0);
            const double nrm = std::sqrt(t.y * t.y + t.x * t.x);
            const Vector2cd g = func(x.x, x.y);
            c(i) = (t.y / nrm) * g(0) + (-t.x / nrm) * g(1);
        }
        return c;
    }
};

// ---------------------------------------------------------------------------
// BuilderData / GalerkinBuilder (galerkin_builder.hpp:24-277)
// ---------------------------------------------------------------------------
enum class LayerType : uint8_t { NONE = 0, DOUBLE = 1, SINGLE = 2, HYPERSINGULAR = 4, ALL = 7 };
inline LayerType operator|(LayerType a, LayerType b) { return static_cast<LayerType>(static_cast<uint8_t>(a) | static_cast<uint8_t>(b)); }
inline LayerType operator~(LayerType a) { return static_cast<LayerType>(~static_cast<uint8_t>(a) & 0x07); }
inline uint8_t operator&(LayerType a, LayerType b) { return static_cast<uint8_t>(a) & static_cast<uint8_t>(b); }

class BuilderData {
  public:
    const ParametrizedMesh& mesh;
    const AbstractBEMSpace& test_space;
    const AbstractBEMSpace& trial_space;

    BuilderData(const ParametrizedMesh& mesh_in, const AbstractBEMSpace& test_space_in,
                const AbstractBEMSpace& trial_space_in, unsigned order, int flags = 0)
        : mesh(mesh_in), test_space(test_space_in), trial_space(trial_space_in), order_(order), flags_(flags) {
        if (&test_space != &trial_space || test_space.getQ() != 2)
            throw std::invalid_argument("the device path supports equal ContinuousSpace<1> test and trial spaces");
        mesh.pack(pan_, coefs_);
        if (coefs_.empty()) coefs_.assign(3, 0.0);
        hbem_geometry_t g = nullptr;
        check(hbem_geometry_create(pan_.data(), (int)mesh.getNumPanels(), coefs_.data(),
                                   mesh_has_coefs() ? (int)(coefs_.size() / 3) : 0, (int)order, 1, flags, &g));
        geo_.reset(g, [](hbem_geometry_t h) { hbem_geometry_destroy(h); });
    }
    size_t getTestSpaceDimension(size_t n_panels = 0) const { return test_space.getSpaceDim(n_panels ? (unsigned)n_panels : mesh.getNumPanels()); }
    size_t getTrialSpaceDimension(size_t n_panels = 0) const { return trial_space.getSpaceDim(n_panels ? (unsigned)n_panels : mesh.getNumPanels()); }
    bool testTrialSpacesAreEqual() const { return &test_space == &trial_space; }
    size_t getQtest() const { return 2; }
    size_t getQtrial() const { return 2; }
    size_t getGaussQROrder() const { return order_; }
    size_t getCGaussQROrder() const { return order_ * (order_ + 1) / 2 - 1; }
    size_t getNumPanels() const { return mesh.getNumPanels(); }
    bool getPanelsAreLines() const { return mesh.isPolygonal(); }
    // device contexts (projection on / off)
    hbem_geometry_t handle(bool proj = true) const {
        if (proj) return geo_.get();
        if (!geo_noproj_) {
            hbem_geometry_t g = nullptr;
            check(hbem_geometry_create(pan_.data(), (int)mesh.getNumPanels(), coefs_.data(),
                                       mesh_has_coefs() ? (int)(coefs_.size() / 3) : 0, (int)order_, 0, flags_, &g));
            geo_noproj_.reset(g, [](hbem_geometry_t h) { hbem_geometry_destroy(h); });
        }
        return geo_noproj_.get();
    }

  private:
    bool mesh_has_coefs() const {
        for (auto& p : mesh.getPanels()) if (p->coefficients()) return true;
        return false;
    }
    unsigned order_;
    int flags_;
    std::vector<double> pan_, coefs_;
    std::shared_ptr<hbem_geometry> geo_;
    mutable std::shared_ptr<hbem_geometry> geo_noproj_;
};

class GalerkinBuilder {
  public:
    explicit GalerkinBuilder(const BuilderData& builder_data) : data(builder_data) {}
    const BuilderData& getData() const { return data; }

    void assembleDense(const std::complex<double>& k_in, double c_in, LayerType layers = LayerType::ALL, size_t d = 0,
                       size_t row = 0, size_t col = 0, size_t rows = 0, size_t cols = 0) {
        if (row || col || rows || cols)
            throw std::invalid_argument("submatrix assembly is not provided by the device path");
        const long N = (long)data.getNumPanels();
        for (int l = 0; l < 3; ++l) has_[l] = false;
        MatrixXcd* out[3] = {nullptr, nullptr, nullptr};
        std::vector<std::complex<double>> buf[3];
        const LayerType lt[3] = {LayerType::DOUBLE, LayerType::HYPERSINGULAR, LayerType::SINGLE};
        for (int l = 0; l < 3; ++l)
            if (layers & lt[l]) buf[l].resize((size_t)N * N * (d + 1));
        check(hbem_assemble_dense(data.handle(), k_in.real(), k_in.imag(), c_in, (int)layers, (int)d,
                                  buf[0].empty() ? nullptr : buf[0].data(), buf[1].empty() ? nullptr : buf[1].data(),
                                  buf[2].empty() ? nullptr : buf[2].data()));
        for (int l = 0; l < 3; ++l) {
            if (buf[l].empty()) continue;
            has_[l] = true;
            for (size_t t = 0; t <= d; ++t) {
                mats_[l][t].resize(N, N);
                std::memcpy(mats_[l][t].data(), buf[l].data() + t * (size_t)N * N, sizeof(std::complex<double>) * N * N);
            }
            d_[l] = d;
        }
        (void)out;
    }
    const MatrixXcd& getDoubleLayer(size_t d = 0) const { return get(0, d); }
    const MatrixXcd& getHypersingular(size_t d = 0) const { return get(1, d); }
    const MatrixXcd& getSingleLayer(size_t d = 0) const { return get(2, d); }

    // Per-call scratch of the reference's pair routines (galerkin_builder.hpp:117-137);
    // the device path keeps its work space in the geometry context.
    struct workspace_t {};
    // GalerkinBuilder::assembleSparse (galerkin_builder.hpp:250-255).  The
    // reference's body computes nothing and leaves `layers` unchanged
    // (galerkin_builder.cpp:1020-1048); it is declared here with the same
    // signature so that code naming it compiles, and it likewise leaves
    // `layers` unchanged.
    void assembleSparse(const std::complex<double>& k_in, double c_in,
                        std::map<LayerType, SparseMatrix<std::complex<double>>>& layers, size_t row, size_t col,
                        size_t rows, size_t cols, std::vector<size_t>& ai, std::vector<size_t>& aj, size_t i, size_t j,
                        workspace_t& ws) {
        (void)k_in; (void)c_in; (void)layers; (void)row; (void)col; (void)rows; (void)cols;
        (void)ai; (void)aj; (void)i; (void)j; (void)ws;
    }

  private:
    const MatrixXcd& get(int l, size_t d) const {
        if (d > 2) throw std::runtime_error("Invalid order of derivative");
        if (!has_[l] || d > d_[l]) throw std::runtime_error("Requested layer is not computed, so cannot be retrieved");
        return mats_[l][d];
    }
    const BuilderData& data;
    MatrixXcd mats_[3][3];
    bool has_[3] = {false, false, false};
    size_t d_[3] = {0, 0, 0};
};

// ---------------------------------------------------------------------------
// SolutionsOperator (gen_sol_op.hpp:8-57).  The derivative variants fill T,
// T_der, T_der2 with the operator and its derivatives (the intended semantics).
// ---------------------------------------------------------------------------
class SolutionsOperator {
  public:
    SolutionsOperator(const BuilderData& builder_data_in, bool proj = true)
        : builder_data(builder_data_in), do_projection(proj), g_(builder_data_in.handle(proj)) {
        const long N = (long)builder_data.getNumPanels();
        M_.resize(N, N);
        check(hbem_mass_matrix(g_, M_.data()));
    }
    void gen_sol_op(const std::complex<double>& k, double c_o, double c_i, MatrixXcd& T) { gen(k, c_o, c_i, 0, &T, nullptr, nullptr); }
    void gen_sol_op(GalerkinBuilder&, const std::complex<double>& k, double c_o, double c_i, MatrixXcd& T) { gen_sol_op(k, c_o, c_i, T); }
    void gen_sol_op_1st_der(const std::complex<double>& k, double c_o, double c_i, MatrixXcd& T, MatrixXcd& T_der) {
        gen(k, c_o, c_i, 1, &T, &T_der, nullptr);
    }
    void gen_sol_op_1st_der(GalerkinBuilder&, const std::complex<double>& k, double c_o, double c_i, MatrixXcd& T, MatrixXcd& T_der) {
        gen_sol_op_1st_der(k, c_o, c_i, T, T_der);
    }
    void gen_sol_op_2nd_der(const std::complex<double>& k, double c_o, double c_i, MatrixXcd& T, MatrixXcd& T_der, MatrixXcd& T_der2) {
        gen(k, c_o, c_i, 2, &T, &T_der, &T_der2);
    }
    void gen_sol_op_2nd_der(GalerkinBuilder&, const std::complex<double>& k, double c_o, double c_i, MatrixXcd& T,
                            MatrixXcd& T_der, MatrixXcd& T_der2) {
        gen_sol_op_2nd_der(k, c_o, c_i, T, T_der, T_der2);
    }
    const MatrixXd& mass_matrix() const { return M_; }
    // operators of the last gen_sol_op* call, assembled on request
    const MatrixXcd& K_int() { return layer(0, true); }
    const MatrixXcd& K_ext() { return layer(0, false); }
    const MatrixXcd& W_int() { return layer(1, true); }
    const MatrixXcd& W_ext() { return layer(1, false); }
    const MatrixXcd& V_int() { return layer(2, true); }
    const MatrixXcd& V_ext() { return layer(2, false); }
    MatrixXcd projection(const MatrixXcd& T) const {
        MatrixXcd out(T.rows(), T.cols());
        check(hbem_projection(g_, T.data(), out.data()));
        return out;
    }
    const BuilderData& getBuilderData() const { return builder_data; }
    bool projects() const { return do_projection; }
    // device calls behind tp::direct_second_kind (hbem/solvers.hpp): U holds
    // nrhs columns of 2N incoming-wave coefficients
    void transmission_system(const std::complex<double>& k, double c_o, double c_i, const std::complex<double>* U,
                             int nrhs, std::complex<double>* A, std::complex<double>* rhs) {
        check(hbem_tp_A_rhs(g_, k.real(), k.imag(), c_o, c_i, U, nrhs, A, rhs));
        k_ = k; c_o_ = c_o; c_i_ = c_i; have_k_ = true;
        for (auto& f : fresh_) f = false;
    }
    void transmission_solve(const std::complex<double>& k, double c_o, double c_i, const std::complex<double>* U,
                            int nrhs, std::complex<double>* sol) {
        check(hbem_tp_solve(g_, k.real(), k.imag(), c_o, c_i, U, nrhs, sol));
    }

  private:
    void gen(const std::complex<double>& k, double c_o, double c_i, int d, MatrixXcd* T, MatrixXcd* Td, MatrixXcd* Tdd) {
        const long n = 2 * (long)builder_data.getNumPanels();
        T->resize(n, n);
        if (Td) Td->resize(n, n);
        if (Tdd) Tdd->resize(n, n);
        check(hbem_gen_sol_op(g_, k.real(), k.imag(), c_o, c_i, d, T->data(), Td ? Td->data() : nullptr,
                              Tdd ? Tdd->data() : nullptr));
        k_ = k; c_o_ = c_o; c_i_ = c_i; have_k_ = true;
        for (auto& f : fresh_) f = false;
    }
    const MatrixXcd& layer(int l, bool inner) {
        if (!have_k_) throw std::runtime_error("no operator has been generated yet");
        const int slot = l * 2 + (inner ? 0 : 1);
        if (!fresh_[slot]) {
            GalerkinBuilder b(builder_data);
            b.assembleDense(k_, inner ? c_i_ : c_o_);
            ops_[slot] = l == 0 ? b.getDoubleLayer() : l == 1 ? b.getHypersingular() : b.getSingleLayer();
            fresh_[slot] = true;
        }
        return ops_[slot];
    }
    const BuilderData& builder_data;
    bool do_projection;
    hbem_geometry_t g_;
    MatrixXd M_;
    std::complex<double> k_ = 0;
    double c_o_ = 1, c_i_ = 1;
    bool have_k_ = false;
    MatrixXcd ops_[6];
    bool fresh_[6] = {false, false, false, false, false, false};
};

}  // namespace hbem

// ---------------------------------------------------------------------------
// randomized_svd (randsvd.hpp:23-77)
// ---------------------------------------------------------------------------
namespace randomized_svd {

// randsvd.cpp:27-47: while benchmarking is on, sv / sv_der / sv_der2 record
// device events around the LU, the first solve + orthonormalisation, the
// subspace iterations and the last solve + SVD; the getters return the
// accumulated seconds of each phase.
inline void benchmarking(bool yes) { hbem::check(hbem_rsvd_benchmarking(yes ? 1 : 0)); }
inline double timer_seconds(int which) {
    double t[4];
    hbem::check(hbem_rsvd_timers(t));
    return t[which];
}
inline double get_lu_time() { return timer_seconds(0); }
inline double get_qr_time() { return timer_seconds(1); }
inline double get_sub_iter_time() { return timer_seconds(2); }
inline double get_svd_time() { return timer_seconds(3); }
inline void reset_timer() { hbem::check(hbem_rsvd_reset_timer()); }

// Complex Gaussian n x nc test matrix (g1 + i g2)/sqrt2 from a seeded
// splitmix64 / xoshiro256** stream (the reference's draws are unseeded).
inline hbem::MatrixXcd randGaussian(int nr, int nc, uint64_t seed = 0) {
    uint64_t x = seed, s[4];
    for (int i = 0; i < 4; ++i) {
        x += 0x9E3779B97F4A7C15ULL;
        uint64_t z = x;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
        s[i] = z ^ (z >> 31);
    }
    auto rotl = [](uint64_t v, int k) { return (v << k) | (v >> (64 - k)); };
    auto next = [&]() {
        uint64_t r = rotl(s[1] * 5, 7) * 9, t = s[1] << 17;
        s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3]; s[2] ^= t; s[3] = rotl(s[3], 45);
        return r;
    };
    auto unif = [&]() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); };
    auto gauss = [&]() {
        double u1 = unif(), u2 = unif();
        if (u1 < 1e-300) u1 = 1e-300;
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * 3.14159265358979323846 * u2);
    };
    hbem::MatrixXcd R(nr, nc);
    for (int j = 0; j < nc; ++j)
        for (int i = 0; i < nr; ++i) {
            double a = gauss(), b = gauss();
            R(i, j) = std::complex<double>(a, b) * std::sqrt(0.5);
        }
    return R;
}

// randGaussian(int) (randsvd.cpp:60-69): a vector of n entries g1 + i g2,
// without the 1/sqrt2 of the matrix form, as the reference's vector overload;
// the draws are those of randGaussian(n, 1) (seed 0).
inline hbem::VectorXcd randGaussian(int n) {
    const hbem::MatrixXcd R = randGaussian(n, 1, 0);
    hbem::VectorXcd v(n, 1);
    for (int i = 0; i < n; ++i) v(i) = R(i, 0) * std::sqrt(2.0);
    return v;
}

inline double sv(const hbem::MatrixXcd& T, const hbem::MatrixXcd& R, int q = 0) {
    double s = 0;
    hbem::check(hbem_rsvd_sv(T.data(), (int)T.rows(), R.data(), (int)R.cols(), q, &s));
    return s;
}

inline hbem::Vector2d sv_der(const hbem::MatrixXcd& T, const hbem::MatrixXcd& T_der, const hbem::MatrixXcd& R, int q = 0) {
    double o[2];
    hbem::check(hbem_rsvd_sv_der(T.data(), T_der.data(), (int)T.rows(), R.data(), (int)R.cols(), q, o));
    hbem::Vector2d v;
    v(0) = o[0]; v(1) = o[1];
    return v;
}

inline hbem::Vector3d sv_der2(const hbem::MatrixXcd& T, const hbem::MatrixXcd& T_der, const hbem::MatrixXcd& T_der2,
                              const hbem::MatrixXcd& R, int q) {
    double o[3];
    hbem::check(hbem_rsvd_sv_der2(T.data(), T_der.data(), T_der2.data(), (int)T.rows(), R.data(), (int)R.cols(), q, o));
    hbem::Vector3d v;
    v(0) = o[0]; v(1) = o[1]; v(2) = o[2];
    return v;
}

}  // namespace randomized_svd
