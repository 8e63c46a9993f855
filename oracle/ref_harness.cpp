// Glue that exposes the compiled reference core through a C ABI so that the
// fixture script (tests/golden/make_golden.py) and the oracle tests can drive
// it with ctypes.  TEST INFRASTRUCTURE ONLY: it is built from the reference
// sources where they lie (see oracle/Makefile) into oracle/_ref/, and only in
// the development container where those sources exist.  The reference is
// never needed at run time; only the fixtures written from it are.
//
// Protected members of the reference classes are reached by deriving from
// them; no reference code is restated here.

#include <stdexcept>
#include <string>
#include <vector>

#include "densecrf.h"
#include "eigen_utils.hpp"
#include "permutohedral.h"
#include "probimage.h"

// free functions of src/densecrf.cpp that the header does not declare
void renormalize(MatrixXf& Q);
void feasible_Q(MatrixXf& tmp, MatrixXi& ind, VectorXd& sum, VectorXi& K, const MatrixXf& Q);

namespace {

struct RefLattice : public Permutohedral {
  int M() const { return M_; }
  int d() const { return d_; }
  void get(int* offset, int* rank, float* bary, int* nbr, int npad) const {
    const int d1 = d_ + 1;
    for (int i = 0; i < npad * d1; ++i) {
      if (offset) offset[i] = offset_[i];
      if (rank) rank[i] = rank_[i];
      if (bary) bary[i] = barycentric_[i];
    }
    if (nbr)
      for (int i = 0; i < d1 * M_; ++i) {
        nbr[2 * i] = blur_neighbors_[i].n1;
        nbr[2 * i + 1] = blur_neighbors_[i].n2;
      }
  }
};

struct RefCRF : public DenseCRF2D {
  RefCRF(int W, int H, int L) : DenseCRF2D(W, H, L) {}
  int N() const { return N_; }
  int L() const { return M_; }
  PairwisePotential* pw(int k) { return pairwise_[k]; }
  PairwisePotential* nn(int k) { return no_norm_pairwise_[k]; }
  void add_layer(const int* region, int R, const float* e, const float* c) {
    std::vector<std::vector<double>> mem(R);
    for (int i = 0; i < N_; ++i) mem[region[i]].push_back(i);
    for (int r = 0; r < R; ++r) {
      if (mem[r].empty()) continue;
      super_pixel_container_.push_back(mem[r]);
      constants_.push_back(c[r]);
      exp_of_superpixels_.conservativeResize(R_ + 1);
      exp_of_superpixels_(R_) = e[r];
      mean_of_superpixels_.conservativeResize(R_ + 1);
      mean_of_superpixels_(R_) = 0;
      ++R_;
    }
  }
  void ucond(MatrixXf& Us, const MatrixXf& Q) const { computeUCondGrad(Us, Q); }
  // region id per pixel (layer order) and exp_of_superpixels_ / constants_
  void layers(int* region, float* e, float* c) const {
    for (int r = 0; r < R_; ++r) {
      for (double m : super_pixel_container_[r]) region[(int)m] = r;
      if (e) e[r] = exp_of_superpixels_(r);
      if (c) c[r] = constants_[r];
    }
  }
  int R() const { return R_; }
  const MatrixXf& z() const { return z_labels_; }
};

thread_local std::string g_err;

MatrixXf as_mat(const float* p, int rows, int cols) {
  MatrixXf m(rows, cols);
  std::copy(p, p + (size_t)rows * cols, m.data());
  return m;
}
VectorXs as_lab(const int* l, int n) {
  VectorXs v(n);
  for (int i = 0; i < n; ++i) v[i] = (short)l[i];
  return v;
}

}  // namespace

extern "C" {

const char* ref_last_error() { return g_err.c_str(); }

// ---- Permutohedral ---------------------------------------------------------
void* ref_lattice_new(const float* feat, int N, int d) {
  RefLattice* L = new RefLattice();
  L->init(as_mat(feat, d, N));
  return L;
}
void ref_lattice_free(void* h) { delete (RefLattice*)h; }
int ref_lattice_M(void* h) { return ((RefLattice*)h)->M(); }
void ref_lattice_get(void* h, int* offset, int* rank, float* bary, int* nbr, int npad) {
  ((RefLattice*)h)->get(offset, rank, bary, nbr, npad);
}
void ref_lattice_compute(void* h, const float* in, float* out, int V, int N, int reverse) {
  MatrixXf i = as_mat(in, V, N), o;
  ((RefLattice*)h)->compute(o, i, reverse != 0);
  std::copy(o.data(), o.data() + (size_t)V * N, out);
}
void ref_lattice_ordered(void* h, const float* q, float* out, int V, int N) {
  MatrixXf i = as_mat(q, V, N), o(V, N);
  o.fill(0);
  ((RefLattice*)h)->compute_upper_minus_lower_ord(o, i);
  std::copy(o.data(), o.data() + (size_t)V * N, out);
}

// Permutohedral::compute_upper_minus_lower_ord_restricted: out (V x nI) is
// passed in and added to, as the reference does
int ref_lattice_ordered_restricted(void* h, const float* rq, const int* pI, int nI, const float* q, int V, int N,
                                   int store, float* out) {
  MatrixXf r = as_mat(rq, V, nI), full = as_mat(q, V, N), o = as_mat(out, V, nI);
  std::vector<int> idx(pI, pI + nI);
  ((RefLattice*)h)->compute_upper_minus_lower_ord_restricted(o, r, idx, full, store != 0);
  std::copy(o.data(), o.data() + (size_t)V * nI, out);
  return 0;
}

// ProbImage::decompress then boostToProb (external/ProbImage/probimage.cpp):
// dims = {width, height, depth}; probs (width*height*depth) may be null
void ref_probimage(const char* path, float* probs, int* dims) {
  ProbImage p;
  p.decompress(path);
  p.boostToProb();
  dims[0] = p.width();
  dims[1] = p.height();
  dims[2] = p.depth();
  if (probs)
    for (int y = 0; y < p.height(); ++y)
      for (int x = 0; x < p.width(); ++x)
        for (int k = 0; k < p.depth(); ++k) probs[((size_t)y * p.width() + x) * p.depth() + k] = p(x, y, k);
}

// Permutohedral::compute_upper_minus_lower_dc on one value per pixel, out (N) added to
void ref_lattice_dc(void* h, float* out, int N, int low, int ml, int mh, int high) {
  MatrixXf o = as_mat(out, 1, N);
  ((RefLattice*)h)->compute_upper_minus_lower_dc(o, low, ml, mh, high);
  std::copy(o.data(), o.data() + N, out);
}
// Permutohedral::gradient: df (d x N) of a^T K b, a and b V x N
void ref_lattice_gradient(void* h, const float* a, const float* b, int V, int N, float* df) {
  ((RefLattice*)h)->gradient(df, a, b, V);
}
// a standalone PairwisePotential(features, Potts(w), ktype, ntype), optional
// setKernelParameters(kp), then kernelGradient(b, Q) into out; returns its size
int ref_pairwise_kernel_gradient(const float* feat, int d, int N, float w, int ktype, int ntype, const float* kp,
                                 int nkp, const float* b, const float* Q, int L, float* out) {
  PairwisePotential p(as_mat(feat, d, N), new PottsCompatibility(w), (KernelType)ktype, (NormalizationType)ntype);
  if (kp) {
    VectorXf v(nkp);
    std::copy(kp, kp + nkp, v.data());
    p.setKernelParameters(v);
  }
  VectorXf g = p.kernelGradient(as_mat(b, L, N), as_mat(Q, L, N));
  std::copy(g.data(), g.data() + g.size(), out);
  return (int)g.size();
}

// LogisticUnaryEnergy(L, f) (src/unary.cpp:53-68): get() into U (L x N) and, when
// b is given, gradient(b) (L*F values) into g
void ref_logistic_unary(const float* Lw, int L, int F, const float* f, int N, const float* b, float* U, float* g) {
  LogisticUnaryEnergy u(as_mat(Lw, L, F), as_mat(f, F, N));
  MatrixXf r = u.get();
  std::copy(r.data(), r.data() + r.size(), U);
  if (b) {
    VectorXf gg = u.gradient(as_mat(b, L, N));
    std::copy(gg.data(), gg.data() + gg.size(), g);
  }
}

// ---- DenseCRF --------------------------------------------------------------
void* ref_crf_new(int W, int H, int L) { return new RefCRF(W, H, L); }
void ref_crf_free(void* h) { delete (RefCRF*)h; }
void ref_crf_set_unary(void* h, const float* U) {
  RefCRF* c = (RefCRF*)h;
  c->setUnaryEnergy(as_mat(U, c->L(), c->N()));
}
void ref_crf_add_pairwise(void* h, const float* feat, int d, float w, int ntype) {
  RefCRF* c = (RefCRF*)h;
  c->addPairwiseEnergy(as_mat(feat, d, c->N()), new PottsCompatibility(w), DIAG_KERNEL, (NormalizationType)ntype);
}
// ctype 0 Potts (params[0] = w), 1 Diagonal (v[L]), 2 Matrix (m[L*L] column-major)
void ref_crf_add_pairwise_compat(void* h, const float* feat, int d, int ctype, const float* params, int ntype) {
  RefCRF* c = (RefCRF*)h;
  const int L = c->L();
  LabelCompatibility* lc;
  if (ctype == 0) {
    lc = new PottsCompatibility(params[0]);
  } else if (ctype == 1) {
    VectorXf v(L);
    for (int i = 0; i < L; ++i) v[i] = params[i];
    lc = new DiagonalCompatibility(v);
  } else {
    lc = new MatrixCompatibility(as_mat(params, L, L));
  }
  c->addPairwiseEnergy(as_mat(feat, d, c->N()), lc, DIAG_KERNEL, (NormalizationType)ntype);
}
// PairwisePotential::gradient(b, Q); returns the number of parameters
int ref_crf_pairwise_gradient(void* h, int k, const float* b, const float* Q, float* g) {
  RefCRF* c = (RefCRF*)h;
  VectorXf r = c->pw(k)->gradient(as_mat(b, c->L(), c->N()), as_mat(Q, c->L(), c->N()));
  if (g) std::copy(r.data(), r.data() + r.size(), g);
  return (int)r.size();
}
void ref_crf_add_gaussian(void* h, float sx, float sy, float w) {
  ((RefCRF*)h)->addPairwiseGaussian(sx, sy, new PottsCompatibility(w));
}
void ref_crf_add_bilateral(void* h, float sx, float sy, float sr, float sg, float sb, const unsigned char* im,
                           float w) {
  ((RefCRF*)h)->addPairwiseBilateral(sx, sy, sr, sg, sb, im, new PottsCompatibility(w));
}
void ref_crf_get_features(void* h, int k, float* out) {
  MatrixXf f = ((RefCRF*)h)->pw(k)->features();
  std::copy(f.data(), f.data() + f.size(), out);
}
void ref_crf_add_layer(void* h, const int* region, int R, const float* e, const float* c) {
  ((RefCRF*)h)->add_layer(region, R, e, c);
}
int ref_crf_R(void* h) { return ((RefCRF*)h)->R(); }
// DenseCRF2D::addSuperPixel (mean-shift segmentation of img), then the layer state
int ref_crf_add_superpixel(void* h, const unsigned char* img, int sr, int rr, int minreg, float cst, float nrm) {
  std::vector<unsigned char> im(img, img + (size_t)((RefCRF*)h)->N() * 3);
  ((RefCRF*)h)->addSuperPixel(im.data(), sr, rr, minreg, cst, nrm);
  return ((RefCRF*)h)->R();
}
void ref_crf_layers(void* h, int* region, float* e, float* c) { ((RefCRF*)h)->layers(region, e, c); }
// which: 0 apply, 1 applyTranspose, 2 no-normalisation copy's apply
void ref_crf_apply(void* h, int k, const float* in, float* out, int which) {
  RefCRF* c = (RefCRF*)h;
  MatrixXf i = as_mat(in, c->L(), c->N()), o;
  if (which == 0) c->pw(k)->apply(o, i);
  else if (which == 1) c->pw(k)->applyTranspose(o, i);
  else c->nn(k)->apply(o, i);
  std::copy(o.data(), o.data() + o.size(), out);
}
// brute-force filters of potential k: 0 apply_bf(Q), 1 apply_upper_minus_lower_bf(ind),
// 2 apply_upper_minus_lower_bf_ord(ind, Q), 3 apply_upper_minus_lower_dc(ind); ind is
// L x N column-major
void ref_crf_apply_bf(void* h, int k, int which, const int* ind, const float* q, float* out) {
  RefCRF* c = (RefCRF*)h;
  const int L = c->L(), N = c->N();
  MatrixXf o(L, N), Q = q ? as_mat(q, L, N) : MatrixXf::Zero(L, N);
  MatrixXi I(L, N);
  if (ind) std::copy(ind, ind + (size_t)L * N, I.data());
  o.fill(0);
  if (which == 0) c->pw(k)->apply_bf(o, Q);
  else if (which == 1) c->pw(k)->apply_upper_minus_lower_bf(o, I);
  else if (which == 2) c->pw(k)->apply_upper_minus_lower_bf_ord(o, I, Q);
  else c->pw(k)->apply_upper_minus_lower_dc(o, I);
  std::copy(o.data(), o.data() + o.size(), out);
}
int ref_crf_apply_ord(void* h, int k, const float* q, float* out) {
  RefCRF* c = (RefCRF*)h;
  MatrixXf i = as_mat(q, c->L(), c->N()), o(c->L(), c->N());
  try {
    c->pw(k)->apply_upper_minus_lower_ord(o, i);
  } catch (std::exception& e) {
    g_err = e.what();
    return -1;
  }
  std::copy(o.data(), o.data() + o.size(), out);
  return 0;
}
void ref_crf_unary_init(void* h, float* Q) {
  MatrixXf q = ((RefCRF*)h)->unary_init();
  std::copy(q.data(), q.data() + q.size(), Q);
}
void ref_crf_qp(void* h, const float* init, float* out) {
  RefCRF* c = (RefCRF*)h;
  MatrixXf q = c->qp_inference_super_pixels_non_convex(as_mat(init, c->L(), c->N()));
  std::copy(q.data(), q.data() + q.size(), out);
}
void ref_crf_zlabels(void* h, float* z) {
  const MatrixXf& m = ((RefCRF*)h)->z();
  std::copy(m.data(), m.data() + m.size(), z);
}
int ref_crf_lp(void* h, const float* init, float* out, const float* fp, const int* ip) {
  RefCRF* c = (RefCRF*)h;
  LP_inf_params P;
  P.prox_reg_const = fp[0];
  P.dual_gap_tol = fp[1];
  P.prox_energy_tol = fp[2];
  P.qp_tol = fp[3];
  P.qp_const = fp[4];
  P.prox_max_iter = ip[0];
  P.fw_max_iter = ip[1];
  P.qp_max_iter = ip[2];
  P.best_int = ip[3] != 0;
  P.less_confident_percent = fp[5];
  P.confidence_tol = fp[6];
  try {
    MatrixXf q = c->lp_inference_prox_super_pixels(as_mat(init, c->L(), c->N()), P);
    std::copy(q.data(), q.data() + q.size(), out);
  } catch (std::exception& e) {
    g_err = e.what();
    return -1;
  }
  return 0;
}
double ref_crf_energy_true(void* h, const int* lab) {
  RefCRF* c = (RefCRF*)h;
  return c->assignment_energy_true(as_lab(lab, c->N()));
}
double ref_crf_energy_ho(void* h, const int* lab) {
  RefCRF* c = (RefCRF*)h;
  return c->assignment_energy_higher_order(as_lab(lab, c->N()));
}
double ref_crf_lp_energy(void* h, const float* Q) {
  RefCRF* c = (RefCRF*)h;
  return c->compute_energy_LP_higher_order(as_mat(Q, c->L(), c->N()));
}
double ref_crf_energy_higher_order(void* h, const float* Q) {
  RefCRF* c = (RefCRF*)h;
  return c->compute_energy_higher_order(as_mat(Q, c->L(), c->N()));
}
double ref_crf_qp_value(void* h, const float* Q) {
  RefCRF* c = (RefCRF*)h;
  return c->compute_LR_QP_value(as_mat(Q, c->L(), c->N()), MatrixXf::Zero(c->L(), c->N()));
}
void ref_crf_ucondgrad(void* h, const float* Q, float* Us) {
  RefCRF* c = (RefCRF*)h;
  MatrixXf u(c->L(), c->N());
  c->ucond(u, as_mat(Q, c->L(), c->N()));
  std::copy(u.data(), u.data() + u.size(), Us);
}
void ref_crf_current_map(void* h, const float* Q, int* lab) {
  RefCRF* c = (RefCRF*)h;
  VectorXs m = c->currentMap(as_mat(Q, c->L(), c->N()));
  for (int i = 0; i < c->N(); ++i) lab[i] = m[i];
}
void ref_feasible(const float* Q, float* out, int N, int L) {
  MatrixXf q = as_mat(Q, L, N), t(L, N);
  MatrixXi ind(L, N);
  VectorXd sum(N);
  VectorXi K(N);
  feasible_Q(t, ind, sum, K, q);
  std::copy(t.data(), t.data() + t.size(), out);
}
void ref_renormalize(float* Q, int N, int L) {
  MatrixXf q = as_mat(Q, L, N);
  renormalize(q);
  std::copy(q.data(), q.data() + q.size(), Q);
}

}  // extern "C"
