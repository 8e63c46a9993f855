// Dense CRF on the device: unary, Potts pairwise potentials over permutohedral
// lattices, Pn superpixel layers, and the QP / LP solvers.
//
// Reference: DenseCRF (include/densecrf.h:73-194, src/densecrf.cpp).
// Fields are L x N stored label-contiguous per pixel (Eigen column-major),
// index pixel*L + label, exactly as the reference's MatrixXf.
#pragma once

#include <functional>
#include <memory>
#include <vector>

#include "lattice.hpp"

namespace dcrf {

// throws Error(kOutOfBounds, "Value out of bounds") unless all of d[0..n) is in [0,1]
void check_unit_interval(const float* d, size_t n, hipStream_t s);

enum NormType : int { kNoNorm = 0, kNormBefore = 1, kNormAfter = 2, kNormSymmetric = 3 };

// LabelCompatibility kinds (include/labelcompatibility.h:51-92)
enum CompatType : int { kPotts = 0, kDiagonal = 1, kMatrix = 2 };
// KernelType (include/pairwise.h:40-44)
enum KernelKind : int { kConstKernel = 0, kDiagKernel = 1, kFullKernel = 2 };

struct Potential {
  std::unique_ptr<Lattice> lat;
  DBuf<float> norm;  // per pixel (DenseKernel::norm_)
  int d = 0;
  float w = 1.f;     // Potts weight
  int ntype = kNormSymmetric;
  // Diagonal / Matrix compatibility: cL labels, host and device copies of
  // v[L] or W[L*L] = (m + m^T)/2 (column-major, symmetric)
  int ctype = kPotts;
  int cL = 0;
  std::vector<float> cparam;
  DBuf<float> cdev;
  DBuf<float> ctmp;  // N x L filter output before the compatibility pass
  DBuf<float> feat;  // N x d features as given (brute-force filters, kernel gradient)
  // DenseKernel's kernel type and parameters (ones(d) for DIAG, identity for
  // FULL until set); the lattice is built on the features they transform
  int ktype = kDiagKernel;
  std::vector<float> kparams;
  // the NO_NORMALIZATION copy used by the true energies is always
  // PottsCompatibility(parameters()(0)), taken when the term is added
  // (src/densecrf.cpp:161-168)
  float nn_w = 1.f;
};

// Gives a freshly added potential a Diagonal (params v[L]) or Matrix (params
// m[L*L] column-major, stored as (m + m^T)/2) compatibility, or a Potts weight
// (params[0]); also fixes the no-norm copy's Potts weight to parameters()(0).
void potential_set_compat(Potential& p, int ctype, const float* params, int L, hipStream_t s);
// LabelCompatibility::parameters / setParameters (src/labelcompatibility.cpp:61-68, :88-96, :109-121)
int potential_nparams(const Potential& p, int L);
void potential_parameters(const Potential& p, int L, float* out);
void potential_set_parameters(Potential& p, int L, const float* v, hipStream_t s);
// Runs `filt` with a SliceArgs that writes the filtered field, then applies the
// compatibility of p: Potts folds into the slice (x -> -w x); Diagonal and
// Matrix filter into p.ctmp and a per-pixel pass writes (or adds) C x to o.out.
void compat_slice(Potential& p, const SliceArgs& o, int V, hipStream_t s,
                  const std::function<void(const SliceArgs&)>& filt);
// PairwisePotential::gradient(b, Q) = compatibility.gradient(b, K Q)
// (src/pairwise.cpp:390-395), device fields, result in the layout of
// parameters(), accumulated (FP64) into the device array acc.
void potential_gradient_dev(Potential& p, int L, const float* b, const float* Q, double* acc, double* parts,
                            hipStream_t s);

// DenseKernel::initLattice: build the lattice and the normalisation
void potential_init(Potential& p, const float* d_feat, int N, int d, float w, int ntype, hipStream_t s);
// the same on an already built lattice (e.g. the union of several images')
void potential_init_lattice(Potential& p, std::unique_ptr<Lattice> lat, float w, int ntype, hipStream_t s);
// PairwisePotential::apply / applyTranspose on V values per pixel
void potential_apply(Potential& p, const float* in, float* out, int V, bool transpose, bool no_norm);
// PairwisePotential::apply_upper_minus_lower_ord (the caller checks [0,1])
void potential_apply_ord(Potential& p, const float* q01, const float* d_zero_minmax, float* out, int V);
// PairwisePotential::apply_upper_minus_lower_ord_restricted (src/pairwise.cpp:229-239):
// rout (nI x V) = compatibility(restricted ordered filter); the caller checks
// pI and the [0,1] ranges
void potential_apply_ord_restricted(Potential& p, float* rout, const float* rq, const int* pI, int nI,
                                    const float* ext, int V, bool store);
// PairwisePotential::apply_bf (src/pairwise.cpp:368-383): dense Gaussian of
// the features, then the compatibility; Q, out N x L
void potential_apply_bf(Potential& p, const float* Q, float* out, int L);
// apply_upper_minus_lower_bf (ord = false, src/pairwise.cpp:294-320) and
// apply_upper_minus_lower_bf_ord (ord = true, :321-353); ind is N x L (the
// pixel at sorted position c of label l is ind[c*L + l], a valid index: the
// caller checks), Q is read only when ord; out N x L in sorted order
void potential_apply_uml_bf(Potential& p, const int* ind, const float* Q, float* out, int L, bool ord);
// DenseKernel::featureGradient (src/pairwise.cpp:95-122): fg (N x d, device)
// = d/df a^T K_norm b for a, b N x V (device)
void potential_feature_gradient(Potential& p, const float* a, const float* b, int V, float* fg);
// kernel parameter count of p (0 / d / d*d) and DenseKernel::gradient
// (:170-181) into host_out; PairwisePotential::kernelGradient passes
// b and compatibility(Q)
int potential_kernel_nparams(const Potential& p);
void potential_kernel_gradient(Potential& p, const float* a, const float* b, int V, float* host_out);
// DenseKernel::setParameters (:149-161): rebuild the lattice (and norm) on
// diag(v) f (DIAG) or V f (FULL, v column-major d x d)
void potential_set_kernel_parameters(Potential& p, const float* v, int n, hipStream_t s);
std::vector<float> potential_kernel_parameters(const Potential& p);
// PairwisePotential::kernelGradient(b, Q) (src/pairwise.cpp:402-407); host out
void potential_pairwise_kernel_gradient(Potential& p, const float* b, const float* Q, int L, float* host_out);
// PairwisePotential::apply_upper_minus_lower_dc (src/pairwise.cpp:241-293);
// ind N x L (checked by the caller), out N x L in sorted order
void potential_apply_uml_dc(Potential& p, const int* ind, float* out, int L);
// LabelCompatibility::apply of a Diagonal or Matrix compatibility on nrows
// label-contiguous rows: out = C in (the Potts form is a scale by -w)
void compat_apply_rows(Potential& p, const float* in, float* out, long nrows, int V, hipStream_t s);

struct QPResult {
  int iterations = 0, updates = 0;
  std::vector<double> energy, num, denom, step;
};

struct LPParams {
  float prox_reg_const = 1e-2f, dual_gap_tol = 1e3f, prox_energy_tol = 1e3f;
  int prox_max_iter = 10, fw_max_iter = 10, qp_max_iter = 1000;
  float qp_tol = 1e3f, qp_const = 1e-16f;
  bool best_int = true;
  // stop once more than this percentage of pixels has no label above
  // confidence_tol in the best iterate (0: no check)
  float less_confident_percent = 0.f, confidence_tol = 0.95f;
};

struct LPResult {
  int outer = 0;
  std::vector<int> qp_sweeps;
  std::vector<double> gap, delta, int_energy;
  double init_lp_energy = 0, init_int_energy = 0;
};

class CRF {
 public:
  CRF(int N, int L);
  ~CRF();
  int N() const { return N_; }
  int L() const { return L_; }
  hipStream_t stream() const { return s_; }

  // ---- model (device pointers) -------------------------------------------
  void set_unary(const float* dU);                                         // setUnaryEnergy
  void add_pairwise(const float* d_feat, int d, float w, int ntype);       // addPairwiseEnergy
  void add_layer(const int* h_region, int R, const float* h_e, const float* h_c);  // superpixel layer
  int num_pairwise() const { return (int)pw_.size(); }
  Potential& pairwise(int k) { return *pw_.at(k); }
  int num_regions() const { return R_; }

  // ---- potentials (device pointers, N x L) ---------------------------------
  void apply(int k, const float* in, float* out, bool transpose, bool no_norm);    // PairwisePotential::apply
  void apply_ord(int k, const float* q01, float* out, bool check);                 // apply_upper_minus_lower_ord
  void unary_init(float* Q);                                                       // DenseCRF::unary_init

  // ---- solvers (device pointers) ---------------------------------------------
  QPResult qp(const float* init, float* out, int max_iter, bool ref_stop, int stats_cap);
  LPResult lp(const float* init, float* out, const LPParams& p, int stats_cap);

  // ---- energies ---------------------------------------------------------------
  void current_map(const float* Q, int* lab);                 // currentMap (device)
  double assignment_energy_true(const int* lab);              // device labels
  // unaryEnergy / pairwise_energy_true (src/densecrf.cpp:887-927): per-pixel
  // float vectors into out (device); term -1 sums every pairwise term
  void unary_energy(const int* lab, float* out);
  void pairwise_energy_true(const int* lab, int term, float* out);
  double assignment_energy_ho(const int* lab);
  double lp_energy(const float* Q);                           // compute_energy_LP_higher_order
  double qp_value(const float* Q);                            // compute_LR_QP_value(Q, 0)
  // compute_energy_higher_order (src/densecrf.cpp:960-980): qp_value(Q) plus
  // the region term with the last QP's z_labels_
  double energy_higher_order(const float* Q) { return qp_value(Q) + ho_energy_term(Q); }
  double ho_energy_term(const float* Q);
  void ucondgrad(const float* Q, float* Us);                  // computeUCondGrad

  // z_labels_ after the last QP (L x R, region-major), host pointer
  void z_labels(float* host);

  // ---- unrolled mean field and its adjoint ------------------------------------
  // T updates from Q0 = softmax(-U): Q <- softmax(-U - sum_k apply_k(Q));
  // keeps Q_0..Q_T and the filtered fields K_k Q_t for meanfield_backward.
  void meanfield(int T, float* Qout);
  // dL/dQ_T -> dL/dU (device N x L) and every potential's parameter gradient,
  // back to back, into the host array dparams (FP64)
  void meanfield_backward(const float* dQ, float* dU, double* dparams);
  int meanfield_iterations() const;
  // The model's generation: it changes whenever the unary, a pairwise term or
  // its parameters change.  meanfield records it, and meanfield_backward
  // refuses to run when it has changed since (the checkpoints would belong to
  // another model).  meanfield_id() numbers the forward runs (0: none yet), so
  // a caller holding an earlier forward can tell that it was overwritten.
  void touch_model() { ++model_gen_; }
  unsigned long long model_generation() const { return model_gen_; }
  long long meanfield_id() const;
  // -sum_i log Q(y_i, i) over labelled pixels (y_i in [0,L)) and its gradient
  double cross_entropy(const float* Q, const int* lab, float* dQ);
  // PairwisePotential::gradient for potential k, host result (parameters() layout)
  void pairwise_gradient(int k, const float* b, const float* Q, float* host_out);
  struct MFState;

  struct Impl;
  struct HOState;
  struct LPState;

 private:
  // device-side building blocks of the LP and the energies
  void int_energy_dev(const int* lab, double* dst, double* scratch, hipStream_t s);
  void lp_energy_dev(const float* Q, double* dst, double* scratch, hipStream_t s);
  void minmax_dev(const float* Q, size_t n, float* mm, int* fl, hipStream_t s);
  void ucondgrad_dev(const float* Q, float* Us, hipStream_t s);
  void ho_maxmin_dev(const float* Q, double* out, hipStream_t s);
  void ho_assign_dev(const int* lab, double* out, int* scratch_int, hipStream_t s);
  std::unique_ptr<LPState> lp_;
  HOState* ho_state();
  void ho_qp_init(hipStream_t s, const float* Q, float* gy, double* sho);
  void ho_qp_direction(hipStream_t s, const float* Q, const int* lab, double* sho, int* fl);
  void ho_qp_zstep(hipStream_t s, double* sc, int* fl);
  void ho_qp_after(hipStream_t s, const float* Q, double* sho, int* fl);
  void ho_qp_finish(hipStream_t s);
  std::unique_ptr<HOState> ho_;
  std::unique_ptr<MFState> mf_;

  int N_, L_;
  int R_ = 0;
  unsigned long long model_gen_ = 0;
  long long mf_count_ = 0;
  hipStream_t s_ = nullptr;
  DBuf<float> U_;
  OrderedBins obins_;  // bins of the field the ordered filters of all potentials share
  std::vector<std::unique_ptr<Potential>> pw_;
  std::unique_ptr<Impl> im_;
};

}  // namespace dcrf
