// Host C++ API mirroring polatory's public interface for the RBF fit/evaluate
// hot path, implemented over the C ABI of libpolatory_amd.so (polatory_amd.h).
//
// Names, argument meaning and error behaviour follow the reference:
//   polatory::rbf::{Biharmonic3D, Triharmonic3D, CovExponential}<3>
//       include/polatory/rbf/polyharmonic_odd.hpp, cov_exponential.hpp, rbf_base.hpp
//   polatory::Model<3>            include/polatory/model.hpp:32-183
//   polatory::Interpolant<3>      include/polatory/interpolant.hpp:34-367
//   polatory::fmm::FmmGenericEvaluatorBase<3>, make_fmm_evaluator,
//   polatory::fmm::FmmGenericSymmetricEvaluatorBase<3>, make_fmm_symmetric_evaluator
//       include/polatory/fmm/fmm_evaluator.hpp:17-106, fmm_symmetric_evaluator.hpp:15-86
// std::invalid_argument / std::runtime_error are thrown where polatory throws them.
// Eigen is not required: VecX and Points are small owning containers with the
// same row-major layout (types.hpp:9-22).
#pragma once

#include <polatory_amd.h>

#include <array>
#include <cmath>
#include <cstdint>
#include <limits>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace polatory {

using Index = std::int64_t;

namespace detail {
inline void check(int rc) {
  if (rc == PAM_OK) return;
  char buf[1024];
  pam_last_error(buf, sizeof buf);
  if (rc == PAM_ERR_INVALID_ARGUMENT) throw std::invalid_argument(buf);
  throw std::runtime_error(buf);
}
}  // namespace detail

// Column vector of doubles (Eigen::VectorXd stand-in).
class VecX {
 public:
  VecX() = default;
  explicit VecX(Index n, double v = 0.0) : d_(static_cast<size_t>(n), v) {}
  explicit VecX(std::vector<double> v) : d_(std::move(v)) {}
  Index rows() const { return static_cast<Index>(d_.size()); }
  Index size() const { return rows(); }
  double& operator()(Index i) { return d_[static_cast<size_t>(i)]; }
  double operator()(Index i) const { return d_[static_cast<size_t>(i)]; }
  double* data() { return d_.data(); }
  const double* data() const { return d_.data(); }
  static VecX Zero(Index n) { return VecX(n); }

 private:
  std::vector<double> d_;
};

namespace geometry {

// Row-major n x Dim point matrix (geometry::Points<Dim>).
template <int Dim>
class Points {
 public:
  Points() = default;
  Points(Index rows, Index cols) : n_(rows), d_(static_cast<size_t>(rows) * Dim) {
    if (cols != Dim) throw std::invalid_argument("Points has a fixed number of columns");
  }
  Index rows() const { return n_; }
  static constexpr Index cols() { return Dim; }
  double& operator()(Index i, Index j) { return d_[static_cast<size_t>(i * Dim + j)]; }
  double operator()(Index i, Index j) const { return d_[static_cast<size_t>(i * Dim + j)]; }
  double* data() { return d_.data(); }
  const double* data() const { return d_.data(); }

 private:
  Index n_ = 0;
  std::vector<double> d_;
};

using Points3 = Points<3>;

template <int Dim>
struct Bbox {
  std::array<double, Dim> min_, max_;
  Bbox() {
    min_.fill(std::numeric_limits<double>::infinity());
    max_.fill(-std::numeric_limits<double>::infinity());
  }
  Bbox(std::array<double, Dim> lo, std::array<double, Dim> hi) : min_(lo), max_(hi) {}
  static Bbox from_points(const Points<Dim>& p) {
    Bbox b;
    for (Index i = 0; i < p.rows(); i++)
      for (int k = 0; k < Dim; k++) {
        b.min_[k] = std::min(b.min_[k], p(i, k));
        b.max_[k] = std::max(b.max_[k], p(i, k));
      }
    return b;
  }
  Bbox convex_hull(const Bbox& o) const {
    Bbox b;
    for (int k = 0; k < Dim; k++) {
      b.min_[k] = std::min(min_[k], o.min_[k]);
      b.max_[k] = std::max(max_[k], o.max_[k]);
    }
    return b;
  }
  const std::array<double, Dim>& min() const { return min_; }
  const std::array<double, Dim>& max() const { return max_; }
};

using Bbox3 = Bbox<3>;

}  // namespace geometry

template <int Dim>
using Mat = std::array<double, Dim * Dim>;  // row-major

namespace rbf {

// Rbf handle (rbf/rbf.hpp:11-118) for the kernels of this library.
template <int Dim>
class Rbf {
  static_assert(Dim == 3, "only 3-D kernels are provided");

 public:
  const Mat<Dim>& anisotropy() const { return aniso_; }
  void set_anisotropy(const Mat<Dim>& a) {
    double det = a[0] * (a[4] * a[8] - a[5] * a[7]) - a[1] * (a[3] * a[8] - a[5] * a[6]) +
                 a[2] * (a[3] * a[7] - a[4] * a[6]);
    if (!(det > 0.0)) throw std::invalid_argument("aniso must have a positive determinant");
    aniso_ = a;
  }
  const std::vector<double>& parameters() const { return params_; }
  int cpd_order() const { return kind_ == PAM_RBF_BH3 ? 1 : (kind_ == PAM_RBF_TH3 ? 2 : 0); }
  std::string short_name() const { return kind_ == PAM_RBF_BH3 ? "bh3" : (kind_ == PAM_RBF_TH3 ? "th3" : "exp"); }
  static Rbf from_c(const pam_rbf& c) {
    Rbf r(c.kind, {c.params[0], c.params[1]});
    Mat<Dim> a;
    for (int i = 0; i < 9; i++) a[i] = c.aniso[i];
    r.set_anisotropy(a);
    return r;
  }
  pam_rbf to_c() const {
    pam_rbf r{};
    r.kind = kind_;
    r.params[0] = params_.at(0);
    r.params[1] = params_.at(1);
    for (int i = 0; i < 9; i++) r.aniso[i] = aniso_[i];
    return r;
  }

 protected:
  Rbf(int kind, std::vector<double> params) : kind_(kind), params_(std::move(params)) {
    aniso_ = {1, 0, 0, 0, 1, 0, 0, 0, 1};
  }

 private:
  int kind_;
  std::vector<double> params_;
  Mat<Dim> aniso_;
};

// polyharmonic_odd.hpp:87-99: {} -> {1, 0}, {slope} -> {slope, 0}
inline std::vector<double> polyharmonic_params(std::vector<double> p) {
  if (p.empty()) return {1.0, 0.0};
  if (p.size() == 1) return {p[0], 0.0};
  if (p.size() != 2) throw std::invalid_argument("params.size() must be 2");
  return p;
}

template <int Dim>
class Biharmonic3D : public Rbf<Dim> {
 public:
  explicit Biharmonic3D(const std::vector<double>& p = {}) : Rbf<Dim>(PAM_RBF_BH3, polyharmonic_params(p)) {}
};

template <int Dim>
class Triharmonic3D : public Rbf<Dim> {
 public:
  explicit Triharmonic3D(const std::vector<double>& p = {}) : Rbf<Dim>(PAM_RBF_TH3, polyharmonic_params(p)) {}
};

template <int Dim>
class CovExponential : public Rbf<Dim> {
 public:
  explicit CovExponential(const std::vector<double>& p) : Rbf<Dim>(PAM_RBF_EXP, check(p)) {}

 private:
  static std::vector<double> check(const std::vector<double>& p) {
    if (p.size() != 2) throw std::invalid_argument("params.size() must be 2");
    return p;
  }
};

}  // namespace rbf

template <int Dim>
class Model {
 public:
  static inline const int kMinRequiredPolyDegree = PAM_POLY_DEGREE_MIN;

  explicit Model(const rbf::Rbf<Dim>& r, int poly_degree = kMinRequiredPolyDegree)
      : Model(std::vector<rbf::Rbf<Dim>>{r}, poly_degree) {}
  explicit Model(std::vector<rbf::Rbf<Dim>> rbfs, int poly_degree = kMinRequiredPolyDegree)
      : rbfs_(std::move(rbfs)), degree_(poly_degree) {
    rebuild();
  }
  Model(const Model& o) : rbfs_(o.rbfs_), degree_(o.degree_), nugget_(o.nugget_) { rebuild(); }
  Model& operator=(const Model&) = delete;
  ~Model() { pam_model_destroy(h_); }

  double nugget() const { return nugget_; }
  void set_nugget(double nugget) {
    if (!(nugget >= 0.0)) throw std::invalid_argument("nugget must be non-negative");
    nugget_ = nugget;
    rebuild();
  }
  int poly_degree() const {
    int d = 0;
    detail::check(pam_model_poly_degree(h_, &d));
    return d;
  }
  Index poly_basis_size() const {
    int l = 0;
    detail::check(pam_model_poly_basis_size(h_, &l));
    return l;
  }
  const std::vector<rbf::Rbf<Dim>>& rbfs() const { return rbfs_; }
  const pam_model* handle() const { return h_; }

  // Model::save / Model::load (model.hpp:169): polatory's binary model file.
  void save(const std::string& filename) const { detail::check(pam_model_save(h_, filename.c_str())); }
  static Model load(const std::string& filename) {
    pam_model* h = nullptr;
    detail::check(pam_model_load(filename.c_str(), &h));
    std::unique_ptr<pam_model, int (*)(pam_model*)> g(h, pam_model_destroy);
    return from_handle(h);
  }
  // A copy of the model behind a C handle.
  static Model from_handle(const pam_model* h) {
    int n = 0, degree = 0;
    double nugget = 0.0;
    detail::check(pam_model_num_rbfs(h, &n));
    std::vector<rbf::Rbf<Dim>> rbfs;
    for (int i = 0; i < n; i++) {
      pam_rbf c{};
      detail::check(pam_model_rbf(h, i, &c));
      rbfs.push_back(rbf::Rbf<Dim>::from_c(c));
    }
    detail::check(pam_model_poly_degree(h, &degree));
    detail::check(pam_model_nugget(h, &nugget));
    Model m(std::move(rbfs), degree);
    m.set_nugget(nugget);
    return m;
  }

 private:
  void rebuild() {
    if (h_) pam_model_destroy(h_);
    h_ = nullptr;
    std::vector<pam_rbf> c;
    for (const auto& r : rbfs_) c.push_back(r.to_c());
    detail::check(pam_model_create(c.data(), static_cast<int>(c.size()), degree_, nugget_, &h_));
  }

  std::vector<rbf::Rbf<Dim>> rbfs_;
  int degree_;
  double nugget_ = 0.0;
  pam_model* h_ = nullptr;
};

template <int Dim>
class Interpolant {
  static constexpr double kInfinity = std::numeric_limits<double>::infinity();
  using Points = geometry::Points<Dim>;

 public:
  explicit Interpolant(const Model<Dim>& model) : model_(model) {
    detail::check(pam_interpolant_create(model_.handle(), &h_));
  }
  ~Interpolant() { pam_interpolant_destroy(h_); }
  Interpolant(const Interpolant&) = delete;
  Interpolant& operator=(const Interpolant&) = delete;

  // interpolant.hpp:77-81
  void fit(const Points& points, const VecX& values, double tolerance, int max_iter = 100,
           double accuracy = kInfinity, const Interpolant* initial = nullptr) {
    if (values.rows() != points.rows())
      throw std::invalid_argument("values.rows() must be equal to " + std::to_string(points.rows()));
    detail::check(pam_interpolant_fit(h_, points.data(), points.rows(), values.data(), tolerance,
                                      max_iter, accuracy, nullptr, initial ? initial->h_ : nullptr));
  }

  // interpolant.hpp:83-113: with gradient data, values = [values at points;
  // gradients at grad_points, x y z each]; the weights become [w; g; c].
  void fit(const Points& points, const Points& grad_points, const VecX& values, double tolerance,
           double grad_tolerance, int max_iter = 100, double accuracy = kInfinity,
           double grad_accuracy = kInfinity, const Interpolant* initial = nullptr) {
    if (values.rows() != points.rows() + 3 * grad_points.rows())
      throw std::invalid_argument("values.rows() must be equal to " +
                                  std::to_string(points.rows() + 3 * grad_points.rows()));
    detail::check(pam_interpolant_fit_hermite(h_, points.data(), points.rows(), grad_points.data(),
                                              grad_points.rows(), values.data(), tolerance, grad_tolerance,
                                              max_iter, accuracy, grad_accuracy, nullptr,
                                              initial ? initial->h_ : nullptr));
  }

  // interpolant.hpp:48
  Points grad_centers() const {
    int64_t n = 0;
    detail::check(pam_interpolant_num_grad_centers(h_, &n));
    Points out(n, Dim);
    detail::check(pam_interpolant_grad_centers(h_, out.data()));
    return out;
  }

  // interpolant.hpp:54-56
  VecX evaluate(const Points& points, double accuracy = kInfinity) {
    VecX out(points.rows());
    detail::check(pam_interpolant_evaluate(h_, points.data(), points.rows(), accuracy, out.data()));
    return out;
  }

  // interpolant.hpp:58-67: values at points, then gradients at grad_points (x y z each)
  VecX evaluate(const Points& points, const Points& grad_points, double accuracy = kInfinity,
                double grad_accuracy = kInfinity) {
    VecX out(points.rows() + 3 * grad_points.rows());
    int64_t n_grad_centers = 0;
    detail::check(pam_interpolant_num_grad_centers(h_, &n_grad_centers));
    if (n_grad_centers > 0) {
      detail::check(pam_interpolant_evaluate_hermite(h_, points.data(), points.rows(), grad_points.data(),
                                                     grad_points.rows(), out.data()));
      return out;
    }
    detail::check(pam_interpolant_evaluate(h_, points.data(), points.rows(), accuracy, out.data()));
    detail::check(pam_interpolant_evaluate_gradient(h_, grad_points.data(), grad_points.rows(), grad_accuracy,
                                                    out.data() + points.rows()));
    return out;
  }

  // interpolant.hpp:69-75
  VecX evaluate_impl(const Points& points) const {
    VecX out(points.rows());
    detail::check(pam_interpolant_evaluate_impl(h_, points.data(), points.rows(), out.data()));
    return out;
  }

  // interpolant.hpp:201-210
  void set_evaluation_bbox_impl(const geometry::Bbox<Dim>& bbox, double accuracy = kInfinity,
                                double grad_accuracy = kInfinity) {
    detail::check(pam_interpolant_set_evaluation_bbox(h_, bbox.min().data(), bbox.max().data(), accuracy,
                                                      grad_accuracy));
  }

  // interpolant.hpp:115-147 (value data)
  void fit_incrementally(const Points& points, const VecX& values, double tolerance, int max_iter = 100,
                         double accuracy = kInfinity) {
    if (values.rows() != points.rows())
      throw std::invalid_argument("values.rows() must be equal to " + std::to_string(points.rows()));
    detail::check(pam_interpolant_fit_incrementally(h_, points.data(), points.rows(), values.data(), tolerance,
                                                    max_iter, accuracy, nullptr));
  }

  // interpolant.hpp:149-190 (NaN: no value / bound at that point)
  void fit_inequality(const Points& points, const VecX& values, const VecX& values_lb, const VecX& values_ub,
                      double tolerance, int max_iter = 100, double accuracy = kInfinity,
                      const Interpolant* initial = nullptr) {
    if (values.rows() != points.rows() || values_lb.rows() != points.rows() || values_ub.rows() != points.rows())
      throw std::invalid_argument("values, values_lb and values_ub must have points.rows() rows");
    detail::check(pam_interpolant_fit_inequality(h_, points.data(), points.rows(), values.data(), values_lb.data(),
                                                 values_ub.data(), tolerance, max_iter, accuracy, nullptr,
                                                 initial ? initial->h_ : nullptr));
  }

  Points centers() const {
    Index n = 0;
    detail::check(pam_interpolant_num_centers(h_, &n));
    Points p(n, Dim);
    detail::check(pam_interpolant_centers(h_, p.data()));
    return p;
  }

  VecX weights() const {
    Index n = 0;
    detail::check(pam_interpolant_num_centers(h_, &n));
    int64_t n_grad = 0;
    detail::check(pam_interpolant_num_grad_centers(h_, &n_grad));
    VecX w(n + 3 * n_grad + model_.poly_basis_size());
    detail::check(pam_interpolant_weights(h_, w.data()));
    return w;
  }

  const Model<Dim>& model() const { return model_; }

  bool fitted() const {
    int f = 0;
    detail::check(pam_interpolant_fitted(h_, &f));
    return f != 0;
  }

  // Interpolant::save / Interpolant::load (interpolant.hpp:218): polatory's
  // binary interpolant file.
  void save(const std::string& filename) const { detail::check(pam_interpolant_save(h_, filename.c_str())); }
  static Interpolant load(const std::string& filename) {
    pam_interpolant* h = nullptr;
    detail::check(pam_interpolant_load(filename.c_str(), &h));
    std::unique_ptr<pam_interpolant, int (*)(pam_interpolant*)> ih(h, pam_interpolant_destroy);
    pam_model* mh = nullptr;
    detail::check(pam_interpolant_model(h, &mh));
    std::unique_ptr<pam_model, int (*)(pam_model*)> mg(mh, pam_model_destroy);
    Model<Dim> m = Model<Dim>::from_handle(mh);
    return Interpolant(ih.release(), m);
  }

 private:
  Interpolant(pam_interpolant* h, const Model<Dim>& model) : model_(model), h_(h) {}

  Model<Dim> model_;
  pam_interpolant* h_ = nullptr;
};

namespace kriging {

// kriging/cross_validate.hpp:14-47 (set_ids as a std::vector instead of Eigen::VectorXi)
template <int Dim>
VecX cross_validate(const Model<Dim>& model, const geometry::Points<Dim>& points, const VecX& values,
                    const std::vector<int32_t>& set_ids, double tolerance, int max_iter, double accuracy) {
  if (values.rows() != points.rows() || static_cast<Index>(set_ids.size()) != points.rows())
    throw std::invalid_argument("values and set_ids must have points.rows() rows");
  VecX out(points.rows());
  detail::check(pam_cross_validate(model.handle(), points.data(), points.rows(), values.data(), set_ids.data(),
                                   tolerance, max_iter, accuracy, nullptr, out.data()));
  return out;
}

}  // namespace kriging

namespace isosurface {

// isosurface/rbf_field_function.hpp:12-36: the field an isosurface extractor
// samples (the FieldFunction base is not part of this library).
class RbfFieldFunction {
  static constexpr double kInfinity = std::numeric_limits<double>::infinity();

 public:
  explicit RbfFieldFunction(Interpolant<3>& interpolant, double accuracy = kInfinity,
                            double grad_accuracy = kInfinity)
      : interpolant_(interpolant), accuracy_(accuracy), grad_accuracy_(grad_accuracy) {}

  VecX operator()(const geometry::Points<3>& points) const { return interpolant_.evaluate_impl(points); }

  void set_evaluation_bbox(const geometry::Bbox<3>& bbox) {
    interpolant_.set_evaluation_bbox_impl(bbox, accuracy_, grad_accuracy_);
  }

 private:
  Interpolant<3>& interpolant_;
  double accuracy_;
  double grad_accuracy_;
};

}  // namespace isosurface

namespace fmm {

template <int Dim>
class FmmGenericEvaluatorBase {
 public:
  virtual ~FmmGenericEvaluatorBase() = default;
  virtual VecX evaluate() const = 0;
  virtual void set_accuracy(double accuracy) = 0;
  virtual void set_source_points(const geometry::Points<Dim>& points) = 0;
  virtual void set_target_points(const geometry::Points<Dim>& points) = 0;
  virtual void set_weights(const VecX& weights) = 0;
};

template <int Dim>
class FmmGenericSymmetricEvaluatorBase {
 public:
  virtual ~FmmGenericSymmetricEvaluatorBase() = default;
  virtual VecX evaluate() const = 0;
  virtual void set_accuracy(double accuracy) = 0;
  virtual void set_points(const geometry::Points<Dim>& points) = 0;
  virtual void set_weights(const VecX& weights) = 0;
};

template <int Dim>
using FmmGenericEvaluatorPtr = std::unique_ptr<FmmGenericEvaluatorBase<Dim>>;
template <int Dim>
using FmmGenericSymmetricEvaluatorPtr = std::unique_ptr<FmmGenericSymmetricEvaluatorBase<Dim>>;

namespace detail {

using CreateFn = int (*)(const pam_rbf*, const double*, const double*, pam_fmm_evaluator**);

// One handle per kernel kind: kn outputs per target (1 for values, 3 for
// gradients), km weights per source.
template <int Dim>
class DeviceEvaluator final : public FmmGenericEvaluatorBase<Dim> {
 public:
  DeviceEvaluator(const rbf::Rbf<Dim>& r, const geometry::Bbox<Dim>& b, CreateFn create = pam_fmm_evaluator_create,
                  int kn = 1)
      : kn_(kn) {
    pam_rbf c = r.to_c();
    polatory::detail::check(create(&c, b.min().data(), b.max().data(), &h_));
  }
  ~DeviceEvaluator() override { pam_fmm_evaluator_destroy(h_); }
  VecX evaluate() const override {
    VecX out(kn_ * n_trg_);
    polatory::detail::check(pam_fmm_evaluator_evaluate(h_, out.data(), n_trg_));
    return out;
  }
  void set_accuracy(double a) override { polatory::detail::check(pam_fmm_evaluator_set_accuracy(h_, a)); }
  void set_source_points(const geometry::Points<Dim>& p) override {
    polatory::detail::check(pam_fmm_evaluator_set_source_points(h_, p.data(), p.rows()));
  }
  void set_target_points(const geometry::Points<Dim>& p) override {
    n_trg_ = p.rows();
    polatory::detail::check(pam_fmm_evaluator_set_target_points(h_, p.data(), p.rows()));
  }
  void set_weights(const VecX& w) override {
    polatory::detail::check(pam_fmm_evaluator_set_weights(h_, w.data(), w.rows()));
  }

 private:
  pam_fmm_evaluator* h_ = nullptr;
  Index n_trg_ = 0;
  int kn_ = 1;
};

template <int Dim>
class DeviceSymmetricEvaluator final : public FmmGenericSymmetricEvaluatorBase<Dim> {
 public:
  DeviceSymmetricEvaluator(const rbf::Rbf<Dim>& r, const geometry::Bbox<Dim>& b,
                           CreateFn create = pam_fmm_symmetric_evaluator_create, int kn = 1)
      : kn_(kn) {
    pam_rbf c = r.to_c();
    polatory::detail::check(create(&c, b.min().data(), b.max().data(), &h_));
  }
  ~DeviceSymmetricEvaluator() override { pam_fmm_evaluator_destroy(h_); }
  VecX evaluate() const override {
    VecX out(kn_ * n_);
    polatory::detail::check(pam_fmm_evaluator_evaluate(h_, out.data(), n_));
    return out;
  }
  void set_accuracy(double a) override { polatory::detail::check(pam_fmm_evaluator_set_accuracy(h_, a)); }
  void set_points(const geometry::Points<Dim>& p) override {
    n_ = p.rows();
    polatory::detail::check(pam_fmm_evaluator_set_points(h_, p.data(), p.rows()));
  }
  void set_weights(const VecX& w) override {
    polatory::detail::check(pam_fmm_evaluator_set_weights(h_, w.data(), w.rows()));
  }

 private:
  pam_fmm_evaluator* h_ = nullptr;
  Index n_ = 0;
  int kn_ = 1;
};

}  // namespace detail

// fmm_evaluator.hpp:92-93
template <int Dim>
FmmGenericEvaluatorPtr<Dim> make_fmm_evaluator(const rbf::Rbf<Dim>& r, const geometry::Bbox<Dim>& bbox) {
  return std::make_unique<detail::DeviceEvaluator<Dim>>(r, bbox);
}

// fmm_symmetric_evaluator.hpp:80-82
template <int Dim>
FmmGenericSymmetricEvaluatorPtr<Dim> make_fmm_symmetric_evaluator(const rbf::Rbf<Dim>& r,
                                                                  const geometry::Bbox<Dim>& bbox) {
  return std::make_unique<detail::DeviceSymmetricEvaluator<Dim>>(r, bbox);
}

// fmm_evaluator.hpp:94-106 (make_fmm_evaluator.cpp:78-188): gradient-weighted
// sources to values, scalar weights to gradients, gradient-weighted sources to
// gradients.
template <int Dim>
FmmGenericEvaluatorPtr<Dim> make_fmm_gradient_evaluator(const rbf::Rbf<Dim>& r, const geometry::Bbox<Dim>& bbox) {
  return std::make_unique<detail::DeviceEvaluator<Dim>>(r, bbox, pam_fmm_gradient_evaluator_create, 1);
}

template <int Dim>
FmmGenericEvaluatorPtr<Dim> make_fmm_gradient_transpose_evaluator(const rbf::Rbf<Dim>& r,
                                                                  const geometry::Bbox<Dim>& bbox) {
  return std::make_unique<detail::DeviceEvaluator<Dim>>(r, bbox, pam_fmm_gradient_transpose_evaluator_create, 3);
}

template <int Dim>
FmmGenericEvaluatorPtr<Dim> make_fmm_hessian_evaluator(const rbf::Rbf<Dim>& r, const geometry::Bbox<Dim>& bbox) {
  return std::make_unique<detail::DeviceEvaluator<Dim>>(r, bbox, pam_fmm_hessian_evaluator_create, 3);
}

// fmm_symmetric_evaluator.hpp:85-86: 3 weights and 3 results per point
template <int Dim>
FmmGenericSymmetricEvaluatorPtr<Dim> make_fmm_hessian_symmetric_evaluator(const rbf::Rbf<Dim>& r,
                                                                          const geometry::Bbox<Dim>& bbox) {
  return std::make_unique<detail::DeviceSymmetricEvaluator<Dim>>(r, bbox, pam_fmm_hessian_symmetric_evaluator_create,
                                                                 3);
}

}  // namespace fmm

}  // namespace polatory
