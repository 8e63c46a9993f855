// Dense CRF model and the Frank-Wolfe QP on gfx950.
//
// All solver scalars (energies, line-search sums, step sizes, stop flags) live
// in device memory: an iteration is a fixed sequence of launches with no host
// synchronisation, so it can be replayed from a HIP graph.  A kernel whose
// "done" flag is set returns at once, which is how the reference's data
// dependent stopping rules are honoured without reading back every iteration.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "crf.hpp"
#include "crf_kernels.hpp"

namespace dcrf {

// accessors of the superpixel state (defined with it in ho.hip)
float* ho_hb(CRF::HOState* h);
int ho_smax(CRF::HOState* h);

// scalar slots (kS_*, kF_*) are defined in crf_kernels.hpp
constexpr int S_ECUR = kS_ECUR, S_DOTS = kS_DOTS, S_EPART = kS_EPART, S_HO = kS_HO, S_TMP = kS_TMP,
              S_COUNT = kS_COUNT;
constexpr int F_DONE = kF_DONE, F_ITERS = kF_ITERS, F_COUNT = kF_COUNT;

struct CRF::Impl {
  DBuf<float> Q, gy, psx, tmp, ones, rpix, mm0, mm, mmparts;
  DBuf<int> lab, iscratch;
  DBuf<double> parts, sc;
  DBuf<int> fl;
  DBuf<double> st_energy, st_num, st_denom, st_step;
  int stats_cap = 0;
  void ensure(int N, int L) {
    const size_t NL = (size_t)N * L;
    Q.resize(NL);
    gy.resize(NL);
    psx.resize(NL);
    tmp.resize(NL);
    lab.resize(N);
    parts.resize((size_t)kRedBlocks * 8);
    sc.resize(S_COUNT);
    fl.resize(F_COUNT);
    rpix.resize(N);
    mm0.resize(2);
    mm.resize(2);
    mmparts.resize(2 * 256);
    iscratch.resize(4);
  }
  void ensure_stats(int cap) {
    if (cap <= stats_cap) return;
    st_energy.resize(cap + 1);
    st_num.resize(cap);
    st_denom.resize(cap);
    st_step.resize(cap);
    stats_cap = cap;
  }
};

CRF::CRF(int N, int L) : N_(N), L_(L), im_(new Impl) {
  DCRF_REQUIRE(N > 0 && L > 0, "DenseCRF: N and L must be positive");
  DCRF_REQUIRE(L <= kMaxLabels, "DenseCRF: at most 64 labels are supported");
  DCRF_HIP(hipStreamCreateWithFlags(&s_, hipStreamNonBlocking));
  U_.resize((size_t)N * L);
  DCRF_HIP(hipMemsetAsync(U_.get(), 0, U_.bytes(), s_));
  im_->ensure(N, L);
  DCRF_HIP(hipMemsetAsync(im_->mm0.get(), 0, 2 * sizeof(float), s_));
  DCRF_HIP(hipStreamSynchronize(s_));
}

CRF::~CRF() {
  if (s_) {
    (void)hipStreamSynchronize(s_);
    pw_.clear();
    im_.reset();
    (void)hipStreamDestroy(s_);
  }
}

void CRF::set_unary(const float* dU) {
  touch_model();
  DCRF_HIP(hipMemcpyAsync(U_.get(), dU, U_.bytes(), hipMemcpyDeviceToDevice, s_));
}

// DenseKernel::initLattice (src/pairwise.cpp:40-57): norm = 1/sqrt(K*1 + 1e-20)
// (symmetric) or 1/(K*1 + 1e-20), evaluated in double and rounded to float.
void potential_init(Potential& p, const float* d_feat, int N, int d, float w, int ntype, hipStream_t s) {
  DCRF_REQUIRE(ntype >= 0 && ntype <= 3, "addPairwiseEnergy: bad normalisation type");
  auto lat = std::make_unique<Lattice>(s);
  lat->build(d_feat, N, d);
  potential_init_lattice(p, std::move(lat), w, ntype, s);
  p.feat.resize((size_t)N * d);
  DCRF_HIP(hipMemcpyAsync(p.feat.get(), d_feat, (size_t)N * d * sizeof(float), hipMemcpyDeviceToDevice, s));
}

void potential_init_lattice(Potential& p, std::unique_ptr<Lattice> lat, float w, int ntype, hipStream_t s) {
  DCRF_REQUIRE(ntype >= 0 && ntype <= 3, "addPairwiseEnergy: bad normalisation type");
  const int N = lat->N();
  p.d = lat->d();
  p.w = w;
  p.nn_w = w;
  p.ntype = ntype;
  p.lat = std::move(lat);
  p.norm.resize(N);
  if (ntype != kNoNorm) {
    DBuf<float> ones(N);
    k_fill<<<grid_for(N), kBlock, 0, s>>>(ones.get(), (size_t)N, 1.f);
    SplatArgs in;
    in.kind = SplatIn::kPlain;
    in.in = ones.get();
    SliceArgs out;
    out.out = p.norm.get();
    p.lat->filter(in, out, 1, false);
    k_norm_finish<<<grid_for(N), kBlock, 0, s>>>(p.norm.get(), N, ntype == kNormSymmetric ? 1 : 0);
    check_launch("k_norm_finish");
    DCRF_HIP(hipStreamSynchronize(s));
  } else {
    k_fill<<<grid_for(N), kBlock, 0, s>>>(p.norm.get(), (size_t)N, 1.f);
    DCRF_HIP(hipStreamSynchronize(s));
  }
}

void CRF::add_pairwise(const float* d_feat, int d, float w, int ntype) {
  touch_model();
  auto p = std::make_unique<Potential>();
  potential_init(*p, d_feat, N_, d, w, ntype, s_);
  pw_.push_back(std::move(p));
}

// which normalisations DenseKernel::filter applies (src/pairwise.cpp:60,73)
static inline bool pre_norm(int nt, bool tr) {
  return nt == kNormSymmetric || (nt == kNormBefore && !tr) || (nt == kNormAfter && tr);
}
static inline bool post_norm(int nt, bool tr) {
  return nt == kNormSymmetric || (nt == kNormBefore && tr) || (nt == kNormAfter && !tr);
}

// PairwisePotential::apply / applyTranspose (src/pairwise.cpp:203-213):
// kernel filter then Potts (out = -w * out, src/labelcompatibility.cpp:49-52).
// no_norm selects the NO_NORMALIZATION copy that addPairwiseEnergy keeps for
// energies (src/densecrf.cpp:164-168); both share one lattice here.
void potential_apply(Potential& p, const float* in, float* out, int V, bool transpose, bool no_norm) {
  const int nt = no_norm ? kNoNorm : p.ntype;
  SplatArgs a;
  a.kind = pre_norm(nt, transpose) ? SplatIn::kScaled : SplatIn::kPlain;
  a.in = in;
  a.norm = p.norm.get();
  SliceArgs o;
  o.out = out;
  o.norm = post_norm(nt, transpose) ? p.norm.get() : nullptr;
  if (no_norm) {  // the no-norm copy: Potts with the first parameter
    o.scale = -p.nn_w;
    o.apply_scale = true;
    p.lat->filter(a, o, V, transpose);
    return;
  }
  compat_slice(p, o, V, p.lat->stream(), [&](const SliceArgs& oo) { p.lat->filter(a, oo, V, transpose); });
}

void potential_apply_ord(Potential& p, const float* q01, const float* d_zero_minmax, float* out, int V) {
  SliceArgs o;
  o.out = out;
  compat_slice(p, o, V, p.lat->stream(),
               [&](const SliceArgs& oo) { p.lat->ordered(q01, d_zero_minmax, V, oo); });
}

void CRF::apply(int k, const float* in, float* out, bool transpose, bool no_norm) {
  potential_apply(pairwise(k), in, out, L_, transpose, no_norm);
}

void CRF::unary_init(float* Q) {
  k_softmax_neg<<<grid_for(N_), kBlock, 0, s_>>>(U_.get(), Q, N_, L_);
  check_launch("k_softmax_neg");
}

// ---------------------------------------------------------------------------
// QP: qp_inference_super_pixels_non_convex (src/densecrf.cpp:340-442)
// ---------------------------------------------------------------------------
QPResult CRF::qp(const float* init, float* out, int max_iter, bool ref_stop, int stats_cap) {
  Impl& m = *im_;
  const size_t NL = (size_t)N_ * L_;
  const FastDiv divL(L_);
  hipStream_t s = s_;
  DCRF_REQUIRE(max_iter > 0 || ref_stop, "qp: max_iter must be positive without the reference stopping rule");
  const int cap_iters = max_iter > 0 ? max_iter : (1 << 30);
  stats_cap = std::max(0, std::min(stats_cap, 1 << 20));
  m.ensure_stats(std::max(stats_cap, 1));
  HOState* ho = ho_state();
  float* Q = m.Q.get();
  float* gy = m.gy.get();
  float* psx = m.psx.get();
  float* tmp = m.tmp.get();
  int* lab = m.lab.get();
  double* sc = m.sc.get();
  int* fl = m.fl.get();

  DCRF_HIP(hipMemcpyAsync(Q, init, NL * sizeof(float), hipMemcpyDeviceToDevice, s));
  DCRF_HIP(hipMemcpyAsync(gy, U_.get(), NL * sizeof(float), hipMemcpyDeviceToDevice, s));
  k_zero_scalars<<<1, 64, 0, s>>>(sc, S_COUNT, fl, F_COUNT);
  // grad_y = U + sum_k 2 psi_k Q ; E0 = U.Q + sum_k Q.psi_k Q  (:377-386, :940-957)
  launch_dot(s, U_.get(), Q, NL, m.parts.get(), sc + S_TMP + 0);
  for (int k = 0; k < num_pairwise(); ++k) {
    apply(k, Q, tmp, false, false);
    k_gy_init_accum<<<kRedBlocks, kBlock, 0, s>>>(gy, tmp, Q, 1, OneImage{(long)NL}, m.parts.get());
    check_launch("k_gy_init_accum");
    finalize_sums<1><<<1, kBlock, 0, s>>>(m.parts.get(), kRedBlocks, sc + S_TMP + 1 + k);
  }
  if (ho) ho_qp_init(s, Q, gy, sc + S_HO);
  k_qp_e0<<<1, 1, 0, s>>>(sc, num_pairwise(), ho != nullptr, m.st_energy.get(), stats_cap, 1);
  check_launch("k_qp_e0");

  {
    KTimer tm("qp_argmin", s);
    k_argmin_rg<<<rg_grid(N_, L_, kRedBlocks), kBlock, 0, s>>>(gy, lab, N_, L_);
    check_launch("k_argmin_rg");
  }
  auto iteration = [&]() {
    if (ho) ho_qp_direction(s, Q, lab, sc + S_HO, fl);
    for (int k = 0; k < num_pairwise(); ++k) {
      Potential& p = pairwise(k);
      SplatArgs a;
      a.kind = pre_norm(p.ntype, false) ? SplatIn::kSxScaled : SplatIn::kSx;
      a.in = Q;
      a.lab = lab;
      a.norm = p.norm.get();
      SliceArgs o;
      o.out = psx;
      o.norm = post_norm(p.ntype, false) ? p.norm.get() : nullptr;
      o.accumulate = k > 0;
      compat_slice(p, o, L_, s, [&](const SliceArgs& oo) { p.lat->filter(a, oo, L_, false); });
    }
    {
      KTimer tm("qp_dots", s);
      k_qp_dots<<<kRedBlocks, kBlock, 0, s>>>(U_.get(), Q, lab, psx, OneImage{N_}, divL, m.parts.get(), fl);
      check_launch("k_qp_dots");
    }
    finalize_sums<3><<<1, kBlock, 0, s>>>(m.parts.get(), kRedBlocks, sc + S_DOTS);
    k_qp_step<<<1, 1, 0, s>>>(sc, fl, ho != nullptr, m.st_num.get(), m.st_denom.get(), m.st_step.get(),
                              m.st_energy.get(), stats_cap, 1);
    check_launch("k_qp_step");
    if (ho) ho_qp_zstep(s, sc, fl);
    {
      KTimer tm("qp_update", s);
      k_qp_update_rg<<<kRedBlocks, kBlock, 0, s>>>(U_.get(), Q, gy, psx, lab, ho ? ho_hb(ho) : nullptr,
                                                    ho ? ho_smax(ho) : 0, OneImage{N_}, L_, sc, m.parts.get(), fl);
      check_launch("k_qp_update_rg");
    }
    finalize_sums<1><<<1, kBlock, 0, s>>>(m.parts.get(), kRedBlocks, sc + S_EPART);
    if (ho) ho_qp_after(s, Q, sc + S_HO, fl);
    k_qp_check<<<1, 1, 0, s>>>(sc, fl, ref_stop ? 1 : 0, ho != nullptr, m.st_energy.get(), stats_cap, 1);
    check_launch("k_qp_check");
  };

  // launches run ahead; the done flag is read back once per group
  const int group = 8;
  int issued = 0, host_flags[F_COUNT] = {0};
  while (issued < cap_iters) {
    const int n = std::min(group, cap_iters - issued);
    for (int i = 0; i < n; ++i) iteration();
    issued += n;
    DCRF_HIP(hipMemcpyAsync(host_flags, fl, sizeof(host_flags), hipMemcpyDeviceToHost, s));
    DCRF_HIP(hipStreamSynchronize(s));
    if (host_flags[F_DONE]) break;
  }
  DCRF_HIP(hipMemcpyAsync(out, Q, NL * sizeof(float), hipMemcpyDeviceToDevice, s));
  if (ho) ho_qp_finish(s);

  QPResult r;
  r.iterations = host_flags[F_ITERS];
  r.updates = r.iterations - (host_flags[F_DONE] == 2 ? 1 : 0);
  if (stats_cap > 0) {
    // energy: before the first and after each entered iteration (unchanged on
    // the one that stops on num==0 || denom==0); step: per update
    const int ne = std::min(r.iterations, stats_cap);
    const int ns = std::min(r.iterations - (host_flags[F_DONE] == 2 ? 1 : 0), stats_cap);
    r.energy.resize(ne + 1);
    r.num.resize(ne);
    r.denom.resize(ne);
    r.step.resize(ns);
    DCRF_HIP(hipMemcpyAsync(r.energy.data(), m.st_energy.get(), (ne + 1) * sizeof(double), hipMemcpyDeviceToHost, s));
    if (ne) {
      DCRF_HIP(hipMemcpyAsync(r.num.data(), m.st_num.get(), ne * sizeof(double), hipMemcpyDeviceToHost, s));
      DCRF_HIP(hipMemcpyAsync(r.denom.data(), m.st_denom.get(), ne * sizeof(double), hipMemcpyDeviceToHost, s));
    }
    if (ns) DCRF_HIP(hipMemcpyAsync(r.step.data(), m.st_step.get(), ns * sizeof(double), hipMemcpyDeviceToHost, s));
  }
  DCRF_HIP(hipStreamSynchronize(s));
  return r;
}

double CRF::qp_value(const float* Q) {
  Impl& m = *im_;
  const size_t NL = (size_t)N_ * L_;
  hipStream_t s = s_;
  k_zero_scalars<<<1, 64, 0, s>>>(m.sc.get(), S_COUNT, m.fl.get(), F_COUNT);
  launch_dot(s, U_.get(), Q, NL, m.parts.get(), m.sc.get() + S_TMP + 0);
  for (int k = 0; k < num_pairwise(); ++k) {
    apply(k, Q, m.tmp.get(), false, false);
    launch_dot(s, Q, m.tmp.get(), NL, m.parts.get(), m.sc.get() + S_TMP + 1 + k);
  }
  std::vector<double> h(S_COUNT);
  DCRF_HIP(hipMemcpyAsync(h.data(), m.sc.get(), S_COUNT * sizeof(double), hipMemcpyDeviceToHost, s));
  DCRF_HIP(hipStreamSynchronize(s));
  double e = 0;  // compute_LR_QP_value(Q, 0) (:940-957)
  e += h[S_TMP];
  e -= 0.0;
  for (int k = 0; k < num_pairwise(); ++k) e += h[S_TMP + 1 + k];
  e += 0.0;
  return e;
}

}  // namespace dcrf

namespace dcrf {

// PairwisePotential::apply_upper_minus_lower_ord (src/pairwise.cpp:214-228)
void CRF::apply_ord(int k, const float* q01, float* out, bool check) {
  if (check) check_unit_interval(q01, (size_t)N_ * L_, s_);
  potential_apply_ord(pairwise(k), q01, im_->mm0.get(), out, L_);
}

// {min, max(Q - min)} of a device array into mm (rescale, src/eigen_utils.cpp:120-140);
// fl[2] = 1 when the maximum is negative or NaN is met
__global__ void __launch_bounds__(kBlock) k_mm_parts(const float* __restrict__ a, size_t n, float* __restrict__ parts) {
  __shared__ float smn[kBlock], smx[kBlock];
  float mn = INFINITY, mx = -INFINITY;
  bool nan = false;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    const float v = a[i];
    nan = nan || v != v;
    mn = fminf(mn, v);
    mx = fmaxf(mx, v);
  }
  smn[threadIdx.x] = nan ? NAN : mn;
  smx[threadIdx.x] = nan ? NAN : mx;
  __syncthreads();
  for (int o = kBlock / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) {
      const float a2 = smn[threadIdx.x + o], b2 = smx[threadIdx.x + o];
      smn[threadIdx.x] = (a2 != a2 || smn[threadIdx.x] != smn[threadIdx.x]) ? NAN : fminf(smn[threadIdx.x], a2);
      smx[threadIdx.x] = (b2 != b2 || smx[threadIdx.x] != smx[threadIdx.x]) ? NAN : fmaxf(smx[threadIdx.x], b2);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    parts[2 * blockIdx.x] = smn[0];
    parts[2 * blockIdx.x + 1] = smx[0];
  }
}
__global__ void k_mm_final(const float* __restrict__ parts, int nb, float* __restrict__ mm, int* __restrict__ fl) {
  float mn = parts[0], mx = parts[1];
  bool nan = mn != mn || mx != mx;
  for (int b = 1; b < nb; ++b) {
    nan = nan || parts[2 * b] != parts[2 * b] || parts[2 * b + 1] != parts[2 * b + 1];
    mn = fminf(mn, parts[2 * b]);
    mx = fmaxf(mx, parts[2 * b + 1]);
  }
  mm[0] = mn;
  mm[1] = mx - mn;
  if (nan || mx - mn < 0) fl[2] = 1;
}

void CRF::minmax_dev(const float* Q, size_t n, float* mm, int* fl, hipStream_t s) {
  k_mm_parts<<<256, kBlock, 0, s>>>(Q, n, im_->mmparts.get());
  k_mm_final<<<1, 1, 0, s>>>(im_->mmparts.get(), 256, mm, fl);
  check_launch("minmax");
}

void CRF::current_map(const float* Q, int* lab) {
  k_argmax<<<grid_for(N_), kBlock, 0, s_>>>(Q, lab, N_, L_);
  check_launch("k_argmax");
}

void CRF::unary_energy(const int* lab, float* out) {
  k_unary_energy<<<grid_for(N_), kBlock, 0, s_>>>(U_.get(), lab, N_, L_, out);
  check_launch("k_unary_energy");
}

// pairwise_energy_true (src/densecrf.cpp:902-927): Q = 1 - onehot(l) through
// the no-norm potential, r_i = -Q(l_i, i); term -1 adds the terms in order
void CRF::pairwise_energy_true(const int* lab, int term, float* out) {
  Impl& m = *im_;
  DCRF_REQUIRE(term >= -1 && term < num_pairwise(), "pairwise_energy_true: no such term");
  k_fill<<<grid_for(N_), kBlock, 0, s_>>>(out, (size_t)N_, 0.f);
  for (int k = term < 0 ? 0 : term; k < (term < 0 ? num_pairwise() : term + 1); ++k) {
    Potential& p = pairwise(k);
    SplatArgs a;
    a.kind = SplatIn::kNotOnehot;
    a.lab = lab;
    SliceArgs o;
    o.out = m.tmp.get();
    o.scale = -p.nn_w;
    o.apply_scale = true;
    p.lat->filter(a, o, L_, false);
    k_pairwise_true_accum<<<grid_for(N_), kBlock, 0, s_>>>(m.tmp.get(), lab, out, N_, L_);
    check_launch("k_pairwise_true_accum");
  }
}

// per-pixel unary + true pairwise energy (unaryEnergy + pairwise_energy_true,
// src/densecrf.cpp:887-927); returns the FP64 total
double CRF::assignment_energy_true(const int* lab) {
  Impl& m = *im_;
  k_fill<<<grid_for(N_), kBlock, 0, s_>>>(m.rpix.get(), (size_t)N_, 0.f);
  for (int k = 0; k < num_pairwise(); ++k) {
    Potential& p = pairwise(k);
    SplatArgs a;
    a.kind = SplatIn::kNotOnehot;
    a.lab = lab;
    SliceArgs o;
    o.out = m.tmp.get();
    o.scale = -p.nn_w;  // the no-norm copy (Potts)
    o.apply_scale = true;
    p.lat->filter(a, o, L_, false);
    k_pairwise_true_accum<<<grid_for(N_), kBlock, 0, s_>>>(m.tmp.get(), lab, m.rpix.get(), N_, L_);
    check_launch("k_pairwise_true_accum");
  }
  k_assign_energy<<<kRedBlocks, kBlock, 0, s_>>>(U_.get(), lab, m.rpix.get(), OneImage{N_}, L_, m.parts.get());
  check_launch("k_assign_energy");
  finalize_sums<1><<<1, kBlock, 0, s_>>>(m.parts.get(), kRedBlocks, m.sc.get() + S_TMP);
  double e = 0;
  DCRF_HIP(hipMemcpyAsync(&e, m.sc.get() + S_TMP, sizeof(double), hipMemcpyDeviceToHost, s_));
  DCRF_HIP(hipStreamSynchronize(s_));
  return e;
}

}  // namespace dcrf
