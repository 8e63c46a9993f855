// densecrf_ho_amd C++ interface: the reference's classes (Permutohedral,
// PairwisePotential, LabelCompatibility/PottsCompatibility, DenseCRF,
// DenseCRF2D, LP_inf_params) with the same names, argument meaning and
// exceptions, forwarding to the C ABI of include/densecrf_ho_amd.h.
//
// Matrices are column-major float matrices exactly as the reference's
// Eigen::MatrixXf (fields L x N, features d x N).  Define DCRF_USE_EIGEN
// before including this header to use Eigen::MatrixXf / VectorXf; otherwise
// the minimal dcrf_ho::MatrixXf below provides the members the API touches
// (rows, cols, data, resize, operator()).
//
// Reference interfaces: include/permutohedral.h:50-88, include/pairwise.h:33-95,
// include/labelcompatibility.h:36-63, include/densecrf.h:43-194.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "densecrf_ho_amd.h"

#ifdef DCRF_USE_EIGEN
#include <Eigen/Core>
#endif

namespace dcrf_ho {

#ifdef DCRF_USE_EIGEN
using MatrixXf = Eigen::MatrixXf;
using VectorXf = Eigen::VectorXf;
using MatrixXi = Eigen::MatrixXi;
using VectorXs = Eigen::Matrix<short, Eigen::Dynamic, 1>;
#else
// column-major matrix (the subset of Eigen::MatrixXf / MatrixXi the API uses)
template <class T>
class DenseMatrix {
 public:
  DenseMatrix() = default;
  DenseMatrix(long r, long c) : r_(r), c_(c), v_((size_t)r * c, T(0)) {}
  long rows() const { return r_; }
  long cols() const { return c_; }
  long size() const { return r_ * c_; }
  T* data() { return v_.data(); }
  const T* data() const { return v_.data(); }
  void resize(long r, long c) {
    r_ = r;
    c_ = c;
    v_.assign((size_t)r * c, T(0));
  }
  T& operator()(long i, long j) { return v_[(size_t)j * r_ + i]; }
  T operator()(long i, long j) const { return v_[(size_t)j * r_ + i]; }
  void fill(T x) { std::fill(v_.begin(), v_.end(), x); }

 private:
  long r_ = 0, c_ = 0;
  std::vector<T> v_;
};
using MatrixXf = DenseMatrix<float>;
using MatrixXi = DenseMatrix<int>;
// column vector (the subset of Eigen::VectorXf the API uses)
class VectorXf : public DenseMatrix<float> {
 public:
  VectorXf() = default;
  explicit VectorXf(long n) : DenseMatrix<float>(n, 1) {}
  float& operator()(long i) { return data()[i]; }
  float operator()(long i) const { return data()[i]; }
  float& operator[](long i) { return data()[i]; }
  float operator[](long i) const { return data()[i]; }
};
class VectorXs {
 public:
  VectorXs() = default;
  explicit VectorXs(long n) : v_((size_t)n, 0) {}
  long rows() const { return (long)v_.size(); }
  long size() const { return (long)v_.size(); }
  short& operator[](long i) { return v_[(size_t)i]; }
  short operator[](long i) const { return v_[(size_t)i]; }
  short* data() { return v_.data(); }
  const short* data() const { return v_.data(); }

 private:
  std::vector<short> v_;
};
#endif

// the reference's exceptions are std::runtime_error with these messages
inline void check(int rc) {
  if (rc != DCRF_OK) throw std::runtime_error(dcrf_last_error());
}

// Follow the reference's portable build (scalar Permutohedral::init and
// seqCompute, src/permutohedral.cpp:334-540) instead of its SSE2 build for
// lattices built from now on.
inline void set_reference_path(bool portable) { check(dcrf_set_reference_path(portable ? 1 : 0)); }
inline bool reference_path() { return dcrf_reference_path() != 0; }

enum NormalizationType { NO_NORMALIZATION, NORMALIZE_BEFORE, NORMALIZE_AFTER, NORMALIZE_SYMMETRIC };
enum KernelType { CONST_KERNEL, DIAG_KERNEL, FULL_KERNEL };

// LabelCompatibility (include/labelcompatibility.h:36-92): Potts
// (out = -w Q), Diagonal (out = diag(v) Q) and Matrix (out = W Q with
// W = (m + m^T)/2)
class LabelCompatibility {
 public:
  virtual ~LabelCompatibility() = default;
  virtual int type() const = 0;                      // DCRF_COMPAT_*
  virtual std::vector<float> c_params() const = 0;  // parameters as the C ABI takes them
  virtual float potts_weight() const { return 1.f; }
};
class PottsCompatibility : public LabelCompatibility {
 public:
  explicit PottsCompatibility(float weight = 1.0f) : w_(weight) {}
  int type() const override { return DCRF_COMPAT_POTTS; }
  std::vector<float> c_params() const override { return {w_}; }
  float potts_weight() const override { return w_; }

 protected:
  float w_;
};
class DiagonalCompatibility : public LabelCompatibility {
 public:
  explicit DiagonalCompatibility(const std::vector<float>& v) : v_(v) {}
#ifdef DCRF_USE_EIGEN
  explicit DiagonalCompatibility(const Eigen::VectorXf& v) : v_(v.data(), v.data() + v.size()) {}
#endif
  int type() const override { return DCRF_COMPAT_DIAGONAL; }
  std::vector<float> c_params() const override { return v_; }

 protected:
  std::vector<float> v_;
};
class MatrixCompatibility : public LabelCompatibility {
 public:
  explicit MatrixCompatibility(const MatrixXf& m) : m_(m.data(), m.data() + m.rows() * m.cols()) {
    if (m.rows() != m.cols()) throw std::runtime_error("MatrixCompatibility: the matrix must be square");
  }
  int type() const override { return DCRF_COMPAT_MATRIX; }
  std::vector<float> c_params() const override { return m_; }  // column-major

 protected:
  std::vector<float> m_;
};

inline VectorXf vector_of(const float* p, long n) {
  VectorXf v(n);
  std::copy(p, p + n, v.data());
  return v;
}

// UnaryEnergy (include/unary.h:31-39): get() is the L x N unary; parameters,
// setParameters and gradient are empty unless a subclass overrides them.  The
// engine reads the unary through upload(): the default copies get() to the
// device; the classes below put theirs there directly.  A DenseCRF uploads its
// unary again before each computation unless versioned() says that version()
// tells every change (then only when it moved), so a subclass whose get()
// changes on its own is read afresh each time, as the reference calls get().
class UnaryEnergy {
 public:
  virtual ~UnaryEnergy() = default;
  virtual MatrixXf get() const = 0;
  virtual VectorXf parameters() const { return VectorXf(); }
  virtual void setParameters(const VectorXf&) {}
  virtual VectorXf gradient(const MatrixXf&) const { return VectorXf(); }
  virtual void upload(dcrf_crf_t h) const {
    const MatrixXf u = get();
    check(dcrf_crf_set_unary(h, u.data(), 0));
  }
  virtual bool versioned() const { return false; }
  unsigned long long version() const { return version_; }

 protected:
  void touch() { ++version_; }
  unsigned long long version_ = 0;
};
// ConstUnaryEnergy (include/unary.h:40-46, src/unary.cpp:46-51)
class ConstUnaryEnergy : public UnaryEnergy {
 public:
  explicit ConstUnaryEnergy(const MatrixXf& unary) : unary_(unary) {}
  MatrixXf get() const override { return unary_; }
  void upload(dcrf_crf_t h) const override { check(dcrf_crf_set_unary(h, unary_.data(), 0)); }
  bool versioned() const override { return true; }

 protected:
  MatrixXf unary_;
};
// LogisticUnaryEnergy (include/unary.h:47-57, src/unary.cpp:53-68): get() = L f on
// the GPU's matrix cores; parameters() is L resized to L*F x 1 (column-major);
// gradient(b) = b f^T resized likewise
class LogisticUnaryEnergy : public UnaryEnergy {
 public:
  LogisticUnaryEnergy(const MatrixXf& L, const MatrixXf& feature) : L_(L), f_(feature) {
    if (L.cols() != feature.rows()) throw std::runtime_error("LogisticUnaryEnergy: L.cols() != f.rows()");
  }
  MatrixXf get() const override {
    MatrixXf u(L_.rows(), f_.cols());
    check(dcrf_logistic_unary(L_.data(), f_.data(), (int)L_.rows(), (int)L_.cols(), (int)f_.cols(), u.data(), 0));
    return u;
  }
  VectorXf parameters() const override { return vector_of(L_.data(), (long)L_.rows() * L_.cols()); }
  void setParameters(const VectorXf& v) override {
    if ((long)v.rows() != (long)L_.rows() * L_.cols())
      throw std::runtime_error("LogisticUnaryEnergy::setParameters: wrong size");
    std::copy(v.data(), v.data() + v.rows(), L_.data());
    touch();
  }
  VectorXf gradient(const MatrixXf& b) const override {
    VectorXf g((long)L_.rows() * L_.cols());
    check(dcrf_logistic_unary_gradient(b.data(), f_.data(), (int)L_.rows(), (int)L_.cols(), (int)f_.cols(), g.data(),
                                       0));
    return g;
  }
  void upload(dcrf_crf_t h) const override {
    check(dcrf_crf_set_unary_logistic(h, L_.data(), f_.data(), (int)L_.cols(), 0));
  }
  bool versioned() const override { return true; }

 protected:
  MatrixXf L_, f_;
};

// LP_inf_params (include/densecrf.h:43-69), defaults of src/densecrf.cpp:287-302
class LP_inf_params {
 public:
  float prox_reg_const, dual_gap_tol, prox_energy_tol;
  int prox_max_iter, fw_max_iter, qp_max_iter;
  float qp_tol, qp_const;
  bool best_int, accel_prox;
  int work_set_size, approx_fw_iter;
  float less_confident_percent, confidence_tol;
  LP_inf_params() {
    dcrf_lp_params p;
    dcrf_lp_params_default(&p);
    from_c(p);
  }
  LP_inf_params(float prox_reg_const_, float dual_gap_tol_, float prox_energy_tol_, int prox_max_iter_,
                int fw_max_iter_, int qp_max_iter_, float qp_tol_, float qp_const_, bool best_int_, bool accel_prox_,
                int work_set_size_, int approx_fw_iter_, float less_confident_percent_, float confidence_tol_)
      : prox_reg_const(prox_reg_const_), dual_gap_tol(dual_gap_tol_), prox_energy_tol(prox_energy_tol_),
        prox_max_iter(prox_max_iter_), fw_max_iter(fw_max_iter_), qp_max_iter(qp_max_iter_), qp_tol(qp_tol_),
        qp_const(qp_const_), best_int(best_int_), accel_prox(accel_prox_), work_set_size(work_set_size_),
        approx_fw_iter(approx_fw_iter_), less_confident_percent(less_confident_percent_),
        confidence_tol(confidence_tol_) {}
  dcrf_lp_params to_c() const {
    dcrf_lp_params p;
    p.prox_reg_const = prox_reg_const;
    p.dual_gap_tol = dual_gap_tol;
    p.prox_energy_tol = prox_energy_tol;
    p.prox_max_iter = prox_max_iter;
    p.fw_max_iter = fw_max_iter;
    p.qp_max_iter = qp_max_iter;
    p.qp_tol = qp_tol;
    p.qp_const = qp_const;
    p.best_int = best_int;
    p.accel_prox = accel_prox;
    p.work_set_size = work_set_size;
    p.approx_fw_iter = approx_fw_iter;
    p.less_confident_percent = less_confident_percent;
    p.confidence_tol = confidence_tol;
    return p;
  }

 private:
  void from_c(const dcrf_lp_params& p) {
    prox_reg_const = p.prox_reg_const;
    dual_gap_tol = p.dual_gap_tol;
    prox_energy_tol = p.prox_energy_tol;
    prox_max_iter = p.prox_max_iter;
    fw_max_iter = p.fw_max_iter;
    qp_max_iter = p.qp_max_iter;
    qp_tol = p.qp_tol;
    qp_const = p.qp_const;
    best_int = p.best_int != 0;
    accel_prox = p.accel_prox != 0;
    work_set_size = p.work_set_size;
    approx_fw_iter = p.approx_fw_iter;
    less_confident_percent = p.less_confident_percent;
    confidence_tol = p.confidence_tol;
  }
};

// Permutohedral (include/permutohedral.h:50-88)
class Permutohedral {
 public:
  Permutohedral() = default;
  ~Permutohedral() {
    if (h_) dcrf_lattice_destroy(h_);
  }
  Permutohedral(const Permutohedral&) = delete;
  Permutohedral& operator=(const Permutohedral&) = delete;
  void init(const MatrixXf& features) {
    if (h_) dcrf_lattice_destroy(h_);
    h_ = nullptr;
    check(dcrf_lattice_create(features.data(), (int)features.cols(), (int)features.rows(), 0, &h_));
    N_ = (int)features.cols();
  }
  void compute(MatrixXf& out, const MatrixXf& in, bool reverse = false) const {
    if (out.rows() != in.rows() || out.cols() != in.cols()) out.resize(in.rows(), in.cols());
    check(dcrf_lattice_compute(h_, in.data(), out.data(), (int)in.rows(), reverse ? 1 : 0, 0));
  }
  MatrixXf compute(const MatrixXf& in, bool reverse = false) const {
    MatrixXf r;
    compute(r, in, reverse);
    return r;
  }
  void compute_upper_minus_lower_ord(MatrixXf& out, const MatrixXf& Q) {
    if (out.rows() != Q.rows() || out.cols() != Q.cols()) out.resize(Q.rows(), Q.cols());
    check(dcrf_lattice_compute_upper_minus_lower_ord(h_, Q.data(), out.data(), (int)Q.rows(), 0));
  }
  void compute_upper_minus_lower_dc(MatrixXf& out, int low, int middle_low, int middle_high, int high) const {
    check(dcrf_lattice_compute_upper_minus_lower_dc(h_, out.data(), low, middle_low, middle_high, high, 0));
  }
  // df (d x N, column-major like the reference's features) of a^T K b
  void gradient(float* df, const float* a, const float* b, int value_size) const {
    check(dcrf_lattice_gradient(h_, df, a, b, value_size, 0));
  }
  // rout (V x |pI|) is added to, as the reference does
  void compute_upper_minus_lower_ord_restricted(MatrixXf& rout, const MatrixXf& rQ, const std::vector<int>& pI,
                                                const MatrixXf& Q, const bool store) {
    if (rout.rows() != rQ.rows() || rout.cols() != rQ.cols()) {
      rout.resize(rQ.rows(), rQ.cols());
      std::fill(rout.data(), rout.data() + (size_t)rout.rows() * rout.cols(), 0.f);
    }
    check(dcrf_lattice_compute_upper_minus_lower_ord_restricted(h_, rout.data(), rQ.data(), pI.data(), (int)pI.size(),
                                                                store ? Q.data() : nullptr, (int)rQ.rows(),
                                                                store ? 1 : 0, 0));
  }
  int latticeSize() const {
    int M = 0;
    check(dcrf_lattice_info(h_, &M, nullptr));
    return M;
  }

 private:
  dcrf_lattice_t h_ = nullptr;
  int N_ = 0;
};

// PairwisePotential (include/pairwise.h:63-95) with a Potts compatibility;
// takes ownership of `compatibility` as the reference does
class PairwisePotential {
 public:
  PairwisePotential(const MatrixXf& features, LabelCompatibility* compatibility, KernelType ktype = CONST_KERNEL,
                    NormalizationType ntype = NORMALIZE_SYMMETRIC)
      : compat_(compatibility), f_(features), ktype_(ktype), ntype_(ntype) {
    check(dcrf_pairwise_create(features.data(), (int)features.cols(), (int)features.rows(),
                               compatibility->potts_weight(), (int)ntype, 0, &h_));
    if (compatibility->type() != DCRF_COMPAT_POTTS) {
      const std::vector<float> p = compatibility->c_params();
      const int L = compatibility->type() == DCRF_COMPAT_DIAGONAL ? (int)p.size()
                                                                   : (int)std::lround(std::sqrt((double)p.size()));
      check(dcrf_pairwise_set_compatibility(h_, compatibility->type(), p.data(), L));
    }
    check(dcrf_pairwise_set_kernel_type(h_, (int)ktype));
  }
  dcrf_pairwise_t handle() const { return h_; }
  ~PairwisePotential() {
    if (h_) dcrf_pairwise_destroy(h_);
  }
  PairwisePotential(const PairwisePotential&) = delete;
  void apply(MatrixXf& out, const MatrixXf& Q) const {
    if (out.rows() != Q.rows() || out.cols() != Q.cols()) out.resize(Q.rows(), Q.cols());
    check(dcrf_pairwise_apply(h_, Q.data(), out.data(), (int)Q.rows(), 0, 0));
  }
  void applyTranspose(MatrixXf& out, const MatrixXf& Q) const {
    if (out.rows() != Q.rows() || out.cols() != Q.cols()) out.resize(Q.rows(), Q.cols());
    check(dcrf_pairwise_apply(h_, Q.data(), out.data(), (int)Q.rows(), 1, 0));
  }
  void apply_upper_minus_lower_ord(MatrixXf& out, const MatrixXf& Q) {
    if (out.rows() != Q.rows() || out.cols() != Q.cols()) out.resize(Q.rows(), Q.cols());
    check(dcrf_pairwise_apply_upper_minus_lower_ord(h_, Q.data(), out.data(), (int)Q.rows(), 0));
  }
  void apply_upper_minus_lower_ord_restricted(MatrixXf& rout, const MatrixXf& rQ, const std::vector<int>& pI,
                                              const MatrixXf& Q, const bool store) {
    if (rout.rows() != rQ.rows() || rout.cols() != rQ.cols()) rout.resize(rQ.rows(), rQ.cols());
    check(dcrf_pairwise_apply_upper_minus_lower_ord_restricted(h_, rout.data(), rQ.data(), pI.data(), (int)pI.size(),
                                                               store ? Q.data() : nullptr, (int)rQ.rows(),
                                                               store ? 1 : 0, 0));
  }
  // kernelParameters / setKernelParameters / kernelGradient (src/pairwise.cpp:396-407)
  std::vector<float> kernelParameters() const {
    std::vector<float> v((size_t)std::max(dcrf_pairwise_kernel_nparams(h_), 0));
    check(dcrf_pairwise_kernel_parameters(h_, v.data()));
    return v;
  }
  void setKernelParameters(const std::vector<float>& v) {
    check(dcrf_pairwise_set_kernel_parameters(h_, v.data(), (int)v.size()));
  }
  std::vector<float> kernelGradient(const MatrixXf& b, const MatrixXf& Q) const {
    std::vector<float> g((size_t)std::max(dcrf_pairwise_kernel_nparams(h_), 0));
    check(dcrf_pairwise_kernel_gradient(h_, b.data(), Q.data(), (int)Q.rows(), g.data(), 0));
    return g;
  }
  void apply_bf(MatrixXf& out, const MatrixXf& Q) const {
    if (out.rows() != Q.rows() || out.cols() != Q.cols()) out.resize(Q.rows(), Q.cols());
    check(dcrf_pairwise_apply_bf(h_, Q.data(), out.data(), (int)Q.rows(), 0));
  }
  void apply_upper_minus_lower_bf(MatrixXf& out, const MatrixXi& ind) const {
    if (out.rows() != ind.rows() || out.cols() != ind.cols()) out.resize(ind.rows(), ind.cols());
    check(dcrf_pairwise_apply_upper_minus_lower_bf(h_, ind.data(), out.data(), (int)ind.rows(), 0));
  }
  void apply_upper_minus_lower_dc(MatrixXf& out, const MatrixXi& ind) const {
    if (out.rows() != ind.rows() || out.cols() != ind.cols()) out.resize(ind.rows(), ind.cols());
    check(dcrf_pairwise_apply_upper_minus_lower_dc(h_, ind.data(), out.data(), (int)ind.rows(), 0));
  }
  void apply_upper_minus_lower_bf_ord(MatrixXf& out, const MatrixXi& ind, const MatrixXf& Q) const {
    if (out.rows() != Q.rows() || out.cols() != Q.cols()) out.resize(Q.rows(), Q.cols());
    check(dcrf_pairwise_apply_upper_minus_lower_bf_ord(h_, ind.data(), Q.data(), out.data(), (int)Q.rows(), 0));
  }
  MatrixXf features() const { return f_; }
  KernelType ktype() const { return ktype_; }
  NormalizationType ntype() const { return ntype_; }

 private:
  std::unique_ptr<LabelCompatibility> compat_;
  MatrixXf f_;
  KernelType ktype_;
  NormalizationType ntype_;
  dcrf_pairwise_t h_ = nullptr;
};

// DenseCRF (include/densecrf.h:73-170)
class DenseCRF {
 public:
  DenseCRF(int N, int M) : N_(N), M_(M) { check(dcrf_crf_create(N, M, &h_)); }
  virtual ~DenseCRF() {
    if (h_) dcrf_crf_destroy(h_);
  }
  DenseCRF(const DenseCRF&) = delete;
  DenseCRF& operator=(const DenseCRF&) = delete;

  // ownership of `function` is transferred, as in the reference
  virtual void addPairwiseEnergy(const MatrixXf& features, LabelCompatibility* function,
                                 KernelType kernel_type = DIAG_KERNEL,
                                 NormalizationType normalization_type = NORMALIZE_SYMMETRIC) {
    std::unique_ptr<LabelCompatibility> f(function);
    if (f->type() == DCRF_COMPAT_POTTS) {
      check(dcrf_crf_add_pairwise(h_, features.data(), (int)features.rows(), f->potts_weight(),
                                  (int)normalization_type, 0));
    } else {
      const std::vector<float> p = f->c_params();
      check(dcrf_crf_add_pairwise_compat(h_, features.data(), (int)features.rows(), f->type(), p.data(),
                                         (int)normalization_type, 0));
    }
    check(dcrf_crf_pairwise_set_kernel_type(h_, npairwise_, (int)kernel_type));
    ++npairwise_;
  }
  // DenseCRF::addPairwiseEnergy(PairwisePotential*) (src/densecrf.cpp:161-168):
  // ownership of `potential` is transferred, as in the reference
  virtual void addPairwiseEnergy(PairwisePotential* potential) {
    std::unique_ptr<PairwisePotential> p(potential);
    check(dcrf_crf_add_pairwise_potential(h_, p->handle()));
    ++npairwise_;
  }
  // PairwisePotential::parameters / setParameters / gradient(b, Q) of term k
  // (src/pairwise.cpp:382-395)
  std::vector<float> pairwise_parameters(int k) const {
    int n = 0;
    check(dcrf_crf_pairwise_nparams(h_, k, &n));
    std::vector<float> v(n);
    check(dcrf_crf_pairwise_parameters(h_, k, v.data()));
    return v;
  }
  void set_pairwise_parameters(int k, const std::vector<float>& v) {
    if (v.size() != pairwise_parameters(k).size()) throw std::runtime_error("setParameters: wrong size");
    check(dcrf_crf_pairwise_set_parameters(h_, k, v.data()));
  }
  std::vector<float> pairwise_gradient(int k, const MatrixXf& b, const MatrixXf& Q) const {
    std::vector<float> g(pairwise_parameters(k).size());
    check(dcrf_crf_pairwise_gradient(h_, k, b.data(), Q.data(), g.data(), 0));
    return g;
  }
  // unrolled mean field from softmax(-U) and its adjoint: dL/dQ_T -> dL/dU
  // and the parameter gradient of every term (back to back, FP64)
  MatrixXf meanfield(int iterations) {
    sync_unary();
    MatrixXf Q(M_, N_);
    check(dcrf_crf_meanfield(h_, iterations, Q.data(), 0));
    return Q;
  }
  MatrixXf meanfield_backward(const MatrixXf& dQ, std::vector<double>* dparams = nullptr) {
    MatrixXf dU(M_, N_);
    size_t n = 0;
    for (int k = 0; k < npairwise_; ++k) n += pairwise_parameters(k).size();
    std::vector<double> g(n ? n : 1);
    check(dcrf_crf_meanfield_backward(h_, dQ.data(), dU.data(), g.data(), 0));
    if (dparams) dparams->assign(g.begin(), g.begin() + n);
    return dU;
  }
  // setUnaryEnergy / getUnaryEnergy (src/densecrf.cpp:255-270): the CRF owns the unary
  void setUnaryEnergy(UnaryEnergy* unary) {
    unary_.reset(unary);
    uploaded_ = nullptr;
    sync_unary();
  }
  void setUnaryEnergy(const MatrixXf& unary) { setUnaryEnergy(new ConstUnaryEnergy(unary)); }
  void setUnaryEnergy(const MatrixXf& L, const MatrixXf& f) { setUnaryEnergy(new LogisticUnaryEnergy(L, f)); }
  UnaryEnergy* getUnaryEnergy() { return unary_.get(); }
  // superpixel layer from a region map (the state addSuperPixel derives from
  // mean-shift): region[i] in [0,R), exp_of_superpixels_ e and constants_ c
  void addSuperPixelLayer(const std::vector<int>& region, const std::vector<float>& e,
                          const std::vector<float>& c) {
    check(dcrf_crf_add_superpixel_layer(h_, region.data(), (int)e.size(), e.data(), c.data()));
  }
  MatrixXf unary_init() const {
    MatrixXf Q(M_, N_);
    sync_unary();
    check(dcrf_crf_unary_init(h_, Q.data(), 0));
    return Q;
  }
  // reference stopping rule (src/densecrf.cpp:390); max_iter > 0 runs a fixed count
  const MatrixXf qp_inference_super_pixels_non_convex(const MatrixXf& init, int max_iter = 0) {
    MatrixXf Q(M_, N_);
    sync_unary();
    check(dcrf_crf_qp(h_, init.data(), Q.data(), max_iter, max_iter > 0 ? 0 : 1, 0, nullptr));
    return Q;
  }
  const MatrixXf lp_inference_prox_super_pixels(const MatrixXf& init, const LP_inf_params& params) const {
    MatrixXf Q(M_, N_);
    sync_unary();
    const dcrf_lp_params p = params.to_c();
    check(dcrf_crf_lp(h_, init.data(), Q.data(), &p, 0, nullptr));
    return Q;
  }
  MatrixXf max_rounding(const MatrixXf& estimates) const {
    VectorXs l = currentMap(estimates);
    MatrixXf r(estimates.rows(), estimates.cols());
    r.fill(0.f);
    for (long i = 0; i < estimates.cols(); ++i) r(l[i], i) = 1.f;
    return r;
  }
  VectorXs currentMap(const MatrixXf& Q) const {
    std::vector<int> lab((size_t)N_);
    check(dcrf_crf_current_map(h_, Q.data(), lab.data(), 0));
    VectorXs r(N_);
    for (int i = 0; i < N_; ++i) r[i] = (short)lab[(size_t)i];
    return r;
  }
  double assignment_energy_true(const VectorXs& l) const {
    sync_unary();
    std::vector<int> lab = labels(l);
    double e = 0;
    check(dcrf_crf_assignment_energy_true(h_, lab.data(), 0, &e));
    return e;
  }
  // public debugging functions of the reference (include/densecrf.h:148-153)
  std::vector<float> unaryEnergy(const VectorXs& l) const {
    sync_unary();
    std::vector<int> lab(l.data(), l.data() + l.size());
    std::vector<float> r((size_t)N_);
    check(dcrf_crf_unary_energy(h_, lab.data(), r.data(), 0));
    return r;
  }
  std::vector<float> pairwise_energy_true(const VectorXs& l, int term = -1) const {
    std::vector<int> lab(l.data(), l.data() + l.size());
    std::vector<float> r((size_t)N_);
    check(dcrf_crf_pairwise_energy_true(h_, lab.data(), term, r.data(), 0));
    return r;
  }
  double compute_energy_higher_order(const MatrixXf& Q) const {
    sync_unary();
    double e = 0;
    check(dcrf_crf_energy_higher_order(h_, Q.data(), 0, &e));
    return e;
  }
  double assignment_energy_higher_order(const VectorXs& l) const {
    sync_unary();
    std::vector<int> lab = labels(l);
    double e = 0;
    check(dcrf_crf_assignment_energy_higher_order(h_, lab.data(), 0, &e));
    return e;
  }
  double compute_energy_LP_higher_order(const MatrixXf& Q) const {
    sync_unary();
    double e = 0;
    check(dcrf_crf_lp_energy(h_, Q.data(), 0, &e));
    return e;
  }
  // compute_LR_QP_value(Q, diag_dom): the engine evaluates diag_dom == 0
  double compute_LR_QP_value(const MatrixXf& Q, const MatrixXf& diag_dom) const {
    sync_unary();
    for (long i = 0; i < diag_dom.size(); ++i)
      if (diag_dom.data()[i] != 0.f) throw std::runtime_error("compute_LR_QP_value: diag_dom must be zero");
    double e = 0;
    check(dcrf_crf_qp_value(h_, Q.data(), 0, &e));
    return e;
  }

 protected:
  std::vector<int> labels(const VectorXs& l) const {
    std::vector<int> lab((size_t)N_);
    for (int i = 0; i < N_; ++i) lab[(size_t)i] = l[i];
    return lab;
  }
  // the device copy of the unary follows unary_ (see UnaryEnergy)
  void sync_unary() const {
    if (!unary_) return;
    if (uploaded_ == unary_.get() && unary_->versioned() && uploaded_version_ == unary_->version()) return;
    unary_->upload(h_);
    uploaded_ = unary_.get();
    uploaded_version_ = unary_->version();
  }
  int N_, M_;
  int npairwise_ = 0;
  dcrf_crf_t h_ = nullptr;
  std::unique_ptr<UnaryEnergy> unary_;
  mutable const UnaryEnergy* uploaded_ = nullptr;
  mutable unsigned long long uploaded_version_ = 0;
};

// DenseCRF2D (include/densecrf.h:172-194)
// The per-region terms of DenseCRF2D::addSuperPixel (src/densecrf.cpp:212-246)
// for a segmentation given as region ids in [0, R): e[r] = constant *
// exp(-sd_r / normaliser) with sd_r the summed squared deviation of the grey
// value from the region mean over the region's size, c[r] = constant.  The
// grey value of pixel i is read from img[i], img[i+1], img[i+2] (not
// img[3i..3i+2]), as the reference reads it; float sums of double terms as in
// the reference.
inline void superpixel_terms(const unsigned char* img, int N, const int* region, int R, float constant,
                             float normaliser, float* e, float* c) {
  std::vector<float> count(R, 0.f), mean(R, 0.f), sd(R, 0.f);
  for (int i = 0; i < N; i++) {
    const int r = region[i];
    if (r < 0 || r >= R) throw std::runtime_error("addSuperPixel: region id out of range");
    mean[r] = (float)(mean[r] + 0.2989 * (double)img[i]);
    mean[r] = (float)(mean[r] + 0.5870 * (double)img[i + 1]);
    mean[r] = (float)(mean[r] + 0.1140 * (double)img[i + 2]);
    count[r] += 1;
  }
  for (int r = 0; r < R; ++r) mean[r] = mean[r] / count[r];
  for (int i = 0; i < N; i++) {
    const int r = region[i];
    const double g = 0.2989 * (double)img[i] + 0.5870 * (double)img[i + 1] + 0.1140 * (double)img[i + 2];
    sd[r] = (float)(sd[r] + (mean[r] - g) * (mean[r] - g));
  }
  for (int r = 0; r < R; ++r) {
    e[r] = constant * std::exp(-1 * (sd[r] / (normaliser * count[r])));
    c[r] = constant;
  }
}

// less_confident_pixels (include/densecrf.h:41, src/densecrf.cpp:457-462)
inline void less_confident_pixels(std::vector<int>& indices, const MatrixXf& Q, float tol = 0.99f) {
  indices.assign((size_t)Q.cols(), 0);
  int n = 0;
  check(dcrf_less_confident_pixels(Q.data(), (int)Q.cols(), (int)Q.rows(), tol, indices.data(), &n, 0));
  indices.resize((size_t)n);
}

class DenseCRF2D : public DenseCRF {
 public:
  DenseCRF2D(int W, int H, int M) : DenseCRF(W * H, M), W_(W), H_(H) {}
  void addPairwiseGaussian(float sx, float sy, LabelCompatibility* function = nullptr,
                           KernelType kernel_type = DIAG_KERNEL,
                           NormalizationType normalization_type = NORMALIZE_SYMMETRIC) {
    std::unique_ptr<LabelCompatibility> f(function ? function : new PottsCompatibility(1.f));
    if (normalization_type == NORMALIZE_SYMMETRIC && f->type() == DCRF_COMPAT_POTTS) {
      check(dcrf_crf_add_pairwise_gaussian(h_, W_, H_, sx, sy, f->potts_weight()));
      check(dcrf_crf_pairwise_set_kernel_type(h_, npairwise_, (int)kernel_type));
      ++npairwise_;
      return;
    }
    MatrixXf feature(2, N_);
    for (int j = 0; j < H_; j++)
      for (int i = 0; i < W_; i++) {
        feature(0, j * W_ + i) = i / sx;
        feature(1, j * W_ + i) = j / sy;
      }
    addPairwiseEnergy(feature, f.release(), kernel_type, normalization_type);
  }
  void addPairwiseBilateral(float sx, float sy, float sr, float sg, float sb, const unsigned char* im,
                            LabelCompatibility* function = nullptr, KernelType kernel_type = DIAG_KERNEL,
                            NormalizationType normalization_type = NORMALIZE_SYMMETRIC) {
    std::unique_ptr<LabelCompatibility> f(function ? function : new PottsCompatibility(1.f));
    if (normalization_type == NORMALIZE_SYMMETRIC && f->type() == DCRF_COMPAT_POTTS) {
      check(dcrf_crf_add_pairwise_bilateral(h_, W_, H_, sx, sy, sr, sg, sb, im, f->potts_weight()));
      check(dcrf_crf_pairwise_set_kernel_type(h_, npairwise_, (int)kernel_type));
      ++npairwise_;
      return;
    }
    MatrixXf feature(5, N_);
    for (int j = 0; j < H_; j++)
      for (int i = 0; i < W_; i++) {
        feature(0, j * W_ + i) = i / sx;
        feature(1, j * W_ + i) = j / sy;
        feature(2, j * W_ + i) = im[(i + j * W_) * 3 + 0] / sr;
        feature(3, j * W_ + i) = im[(i + j * W_) * 3 + 1] / sg;
        feature(4, j * W_ + i) = im[(i + j * W_) * 3 + 2] / sb;
      }
    addPairwiseEnergy(feature, f.release(), kernel_type, normalization_type);
  }
  // DenseCRF2D::addSuperPixel (src/densecrf.cpp:193-251) with the mean-shift
  // segmentation given as its result, a region id per pixel in [0, R): the
  // layer's exp_of_superpixels_ = constant * exp(-sd / normaliser) from the
  // grey value of each region, constants_ = constant.  The grey value of pixel
  // i is read from img[i], img[i+1], img[i+2] (not img[3i..3i+2]), as the
  // reference reads it.  Accumulation types follow the reference: float
  // vectors with double terms.
  void addSuperPixel(const unsigned char* img, const std::vector<int>& region, int R, float constant,
                     float normaliser) {
    if ((long)region.size() != (long)W_ * H_) throw std::runtime_error("addSuperPixel: one region id per pixel");
    std::vector<float> e(R), c(R);
    superpixel_terms(img, W_ * H_, region.data(), R, constant, normaliser, e.data(), c.data());
    addSuperPixelLayer(region, e, c);
  }

 protected:
  int W_, H_;
};

// Many images of uneven sizes solved together (no reference counterpart;
// each image follows the per-image algorithm).  Fields are the images'
// fields concatenated: image b owns columns [offsets()[b], offsets()[b+1]).
class DenseCRFBatch {
 public:
  DenseCRFBatch(int M, const std::vector<int>& W, const std::vector<int>& H) : M_(M), B_((int)W.size()) {
    if (W.size() != H.size()) throw std::runtime_error("DenseCRFBatch: W and H differ in length");
    check(dcrf_batch_create_2d(M, B_, W.data(), H.data(), &h_));
    off_.resize(B_ + 1);
    check(dcrf_batch_pixel_offsets(h_, off_.data()));
  }
  ~DenseCRFBatch() {
    if (h_) dcrf_batch_destroy(h_);
  }
  DenseCRFBatch(const DenseCRFBatch&) = delete;
  DenseCRFBatch& operator=(const DenseCRFBatch&) = delete;
  const std::vector<long long>& offsets() const { return off_; }
  long long N() const { return off_.back(); }
  void setUnaryEnergy(int b, const MatrixXf& unary) { check(dcrf_batch_set_unary(h_, b, unary.data(), 0)); }
  void addPairwiseGaussian(float sx, float sy, float w) { check(dcrf_batch_add_pairwise_gaussian(h_, sx, sy, w)); }
  void addPairwiseBilateral(float sx, float sy, float sr, float sg, float sb,
                            const std::vector<const unsigned char*>& images, float w) {
    check(dcrf_batch_add_pairwise_bilateral(h_, sx, sy, sr, sg, sb, images.data(), w));
  }
  MatrixXf qp(int max_iter) {
    MatrixXf Q(M_, (long)N());
    check(dcrf_batch_qp(h_, nullptr, Q.data(), max_iter, max_iter > 0 ? 0 : 1, 0, nullptr));
    return Q;
  }
  MatrixXf lp(const MatrixXf& init, const LP_inf_params& params) {
    MatrixXf Q(M_, (long)N());
    dcrf_lp_params p = params.to_c();
    check(dcrf_batch_lp(h_, init.data(), Q.data(), &p, 0, nullptr, 0));
    return Q;
  }

 private:
  int M_, B_;
  std::vector<long long> off_;
  dcrf_batch_t h_ = nullptr;
};

}  // namespace dcrf_ho
