// Label compatibilities beyond Potts (Diagonal, Matrix), the reference's
// PairwisePotential::gradient, and an unrolled mean-field forward pass with
// its hand-written adjoint.
//
// Reference pieces: LabelCompatibility and its gradients
// (src/labelcompatibility.cpp:47-129), PairwisePotential::apply /
// applyTranspose / gradient (src/pairwise.cpp:203-213, 390-395),
// expAndNormalize (src/densecrf.cpp:328-336).  The fork has no inference loop
// of its own; the mean-field update here is the classic one built from those
// pieces: Q0 = softmax(-U), then Q <- softmax(-U - sum_k apply_k(Q)).
//
// Per-pixel kernels use the row-group mapping (lanes = labels of one pixel,
// ⌊64/L⌋ pixels per wave) with the pixel's row staged in a wave-private LDS
// segment, so every global access is a coalesced label-contiguous row.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "crf.hpp"
#include "crf_kernels.hpp"

namespace dcrf {
namespace {

constexpr int kMaxPw = 8;          // potentials fused into one mean-field update launch
constexpr int kOuterBlocks = 2048;  // blocks of the L x L gradient reduction

// out (+)= C x per pixel; C Diagonal (cm = v[L]) or Matrix (cm = W[L*L]
// column-major).  The sum over j runs in index order like the oracle's.
__global__ void __launch_bounds__(kBlock) k_compat_rg(int N, int L, int ctype, const float* __restrict__ cm,
                                                      const float* __restrict__ x, float* __restrict__ out,
                                                      int accumulate) {
  __shared__ float seg[kBlock];
  __shared__ float sW[kMaxLabels * kMaxLabels];
  const int nW = ctype == kMatrix ? L * L : L;
  for (int i = threadIdx.x; i < nW; i += blockDim.x) sW[i] = cm[i];
  __syncthreads();
  const long ntiles = rg_tiles(N, L);
  for (long tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    const RG r = rg_at(L, N, tile);
    float* row = seg + (threadIdx.x & ~63) + r.g * L;
    float xv = 0.f;
    if (r.valid) {
      xv = x[(size_t)r.row * L + r.k];
      row[r.k] = xv;
    }
    __syncthreads();
    if (r.valid) {
      float y;
      if (ctype == kDiagonal) {
        y = sW[r.k] * xv;
      } else {
        y = 0.f;
        for (int j = 0; j < L; ++j) y = y + sW[j * L + r.k] * row[j];
      }
      const size_t idx = (size_t)r.row * L + r.k;
      out[idx] = accumulate ? out[idx] + y : y;
    }
    __syncthreads();
  }
}

// Per-block partials of G = sum_i b_i m_i^T (G[a*L + c] = sum_i b(a,i) m(c,i)).
// One wave per block, ⌊64/L⌋ pixels per step; lane (pixel group, label a)
// keeps its row G[a][0..L) in registers (MAXL >= L) and adds b(a,i) m(c,i)
// for the pixel's m row staged in LDS.  At the end the wave's rows meet in
// LDS and the block writes their sum (FP64, fixed order).
template <int MAXL>
__global__ void __launch_bounds__(64) k_outer_w(int N, int L, const float* __restrict__ b,
                                                const float* __restrict__ m, double* __restrict__ parts) {
  __shared__ float red[64 * kMaxLabels];  // gpw * L * L <= 64 * L
  __shared__ float seg[64];
  const int gpw = 64 / L, LL = L * L, lane = threadIdx.x;
  const int g = lane / L, k = lane - g * L;
  float acc[MAXL];
#pragma unroll
  for (int c = 0; c < MAXL; ++c) acc[c] = 0.f;
  const long nt = ((long)N + gpw - 1) / gpw;
  for (long tile = blockIdx.x; tile < nt; tile += gridDim.x) {
    const long row = tile * gpw + g;
    const bool valid = g < gpw && row < N;
    float bv = 0.f;
    if (valid) {
      const size_t idx = (size_t)row * L + k;
      bv = b[idx];
      seg[g * L + k] = m[idx];
    }
    __syncthreads();
    if (valid) {
      const float* mr = seg + g * L;
#pragma unroll
      for (int c = 0; c < MAXL; ++c)
        if (c < L) acc[c] = acc[c] + bv * mr[c];
    }
    __syncthreads();
  }
  if (g < gpw) {
    float* dst = red + (size_t)(g * L + k) * L;
#pragma unroll
    for (int c = 0; c < MAXL; ++c)
      if (c < L) dst[c] = acc[c];
  }
  __syncthreads();
  for (int e = lane; e < LL; e += 64) {
    double s = 0.0;
    for (int q = 0; q < gpw; ++q) s += (double)red[(size_t)q * LL + e];
    parts[(size_t)blockIdx.x * LL + e] = s;
  }
}

// G[e] = sum over blocks of the partials (one block per entry, fixed order)
__global__ void __launch_bounds__(kBlock) k_outer_sum(const double* __restrict__ parts, int nblocks, int LL,
                                                      double* __restrict__ G) {
  double v[1] = {0.0};
  for (int q = threadIdx.x; q < nblocks; q += blockDim.x) v[0] += parts[(size_t)q * LL + blockIdx.x];
  block_sum<1>(v);
  if (threadIdx.x == 0) G[blockIdx.x] = v[0];
}

// the compatibility's gradient packing of G, added into acc:
//   Potts     -sum(b o Q)            = -trace G          (src/labelcompatibility.cpp:72-76)
//   Diagonal  row sums of b o Q      = diag G            (:91-93)
//   Matrix    upper triangle of G + G^T off the diagonal  (:122-129)
__global__ void k_outer_pack(const double* __restrict__ G, int L, int ctype, double* __restrict__ acc) {
  const int LL = L * L;
  if (ctype == kPotts) {
    if (threadIdx.x == 0) {
      double t = 0.0;
      for (int a = 0; a < L; ++a) t += G[a * L + a];
      acc[0] = acc[0] - t;
    }
  } else if (ctype == kDiagonal) {
    for (int a = threadIdx.x; a < L; a += blockDim.x) acc[a] = acc[a] + G[a * L + a];
  } else {
    for (int e = threadIdx.x; e < LL; e += blockDim.x) {
      const int i = e / L, j = e - i * L;
      if (j < i) continue;
      const int k = i * L - i * (i - 1) / 2 + (j - i);  // packed index of (i, j), j >= i
      acc[k] = acc[k] + G[i * L + j] + (i != j ? G[j * L + i] : 0.0);
    }
  }
}

struct MFTerm {
  const float* m;   // K_k Q of this potential
  const float* cm;  // Diagonal v / Matrix W (device), unused for Potts
  int ctype;
  float w;          // Potts weight
};
struct MFArgs {
  MFTerm t[kMaxPw];
  int K;
};

// One mean-field update: a = -U - sum_k C_k(m_k) (in potential order), then
// Q = softmax(a) per pixel (max, exp, sum in label order).
__global__ void __launch_bounds__(kBlock) k_mf_update_rg(int N, int L, const float* __restrict__ U, MFArgs A,
                                                         float* __restrict__ Q) {
  __shared__ float seg[kBlock];
  const long ntiles = rg_tiles(N, L);
  for (long tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    const RG r = rg_at(L, N, tile);
    float* row = seg + (threadIdx.x & ~63) + r.g * L;
    const size_t idx = (size_t)r.row * L + r.k;
    float a = r.valid ? -U[idx] : 0.f;
#pragma unroll
    for (int k = 0; k < kMaxPw; ++k) {
      if (k >= A.K) break;  // uniform: A.K is a kernel argument
      const MFTerm& t = A.t[k];
      float c = 0.f;
      if (t.ctype == kMatrix) {
        if (r.valid) row[r.k] = t.m[idx];
        __syncthreads();
        if (r.valid)
          for (int j = 0; j < L; ++j) c = c + t.cm[j * L + r.k] * row[j];
        __syncthreads();
      } else if (r.valid) {
        const float mv = t.m[idx];
        c = t.ctype == kPotts ? -t.w * mv : t.cm[r.k] * mv;
      }
      a = a - c;
    }
    if (r.valid) row[r.k] = a;
    __syncthreads();
    float e = 0.f;
    if (r.valid) {
      float mx = row[0];
      for (int j = 1; j < L; ++j) mx = fmaxf(mx, row[j]);
      e = expf(a - mx);
    }
    __syncthreads();
    if (r.valid) row[r.k] = e;
    __syncthreads();
    if (r.valid) {
      float s = 0.f;
      for (int j = 0; j < L; ++j) s = s + row[j];
      Q[idx] = e / s;
    }
    __syncthreads();
  }
}

// Softmax adjoint of one update: ga = Q o (gQ - <Q, gQ>) per pixel;
// dU = -ga (first) or dU - ga; b = -ga (the gradient with respect to the
// pairwise output, which enters a with a minus sign) when b is non-null.
__global__ void __launch_bounds__(kBlock) k_mf_back_rg(int N, int L, const float* __restrict__ Q,
                                                       const float* __restrict__ gQ, float* __restrict__ dU,
                                                       int first, float* __restrict__ b) {
  __shared__ float seg[kBlock];
  const long ntiles = rg_tiles(N, L);
  for (long tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    const RG r = rg_at(L, N, tile);
    float* row = seg + (threadIdx.x & ~63) + r.g * L;
    const size_t idx = (size_t)r.row * L + r.k;
    float q = 0.f, g = 0.f;
    if (r.valid) {
      q = Q[idx];
      g = gQ[idx];
      row[r.k] = q * g;
    }
    __syncthreads();
    if (r.valid) {
      float s = 0.f;
      for (int j = 0; j < L; ++j) s = s + row[j];
      const float ga = q * (g - s);
      dU[idx] = first ? -ga : dU[idx] - ga;
      if (b) b[idx] = -ga;
    }
    __syncthreads();
  }
}

// -sum log Q(y_i, i) over pixels with a label in [0, L), and
// dQ(l, i) = -[l == y_i] / Q(y_i, i)
__global__ void __launch_bounds__(kBlock) k_xent_rg(int N, int L, const float* __restrict__ Q,
                                                    const int* __restrict__ lab, float* __restrict__ dQ,
                                                    double* __restrict__ parts) {
  double e = 0.0;
  const long ntiles = rg_tiles(N, L);
  for (long tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    const RG r = rg_at(L, N, tile);
    if (!r.valid) continue;
    const size_t idx = (size_t)r.row * L + r.k;
    const int y = lab[r.row];
    float g = 0.f;
    if (y == r.k) {
      const float q = Q[idx];
      e -= log((double)q);
      g = -1.f / q;
    }
    dQ[idx] = g;
  }
  double v[1] = {e};
  block_sum<1>(v);
  if (threadIdx.x == 0) parts[blockIdx.x] = v[0];
}

}  // namespace

// ---------------------------------------------------------------------------
// compatibilities
// ---------------------------------------------------------------------------
void potential_set_compat(Potential& p, int ctype, const float* params, int L, hipStream_t s) {
  DCRF_REQUIRE(ctype >= kPotts && ctype <= kMatrix, "label compatibility: unknown type");
  DCRF_REQUIRE(params != nullptr, "label compatibility: null parameters");
  p.ctype = ctype;
  if (ctype == kPotts) {
    p.w = params[0];
    p.nn_w = p.w;
    p.cL = 0;
    p.cparam.clear();
    return;
  }
  DCRF_REQUIRE(L > 0 && L <= kMaxLabels, "label compatibility: 1..64 labels");
  p.cL = L;
  if (ctype == kDiagonal) {
    p.cparam.assign(params, params + L);
  } else {
    // MatrixCompatibility stores w_ = 0.5 * (m + m^T) (src/labelcompatibility.cpp:94-97)
    p.cparam.assign((size_t)L * L, 0.f);
    for (int j = 0; j < L; ++j)
      for (int i = 0; i < L; ++i)
        p.cparam[(size_t)j * L + i] = 0.5f * (params[(size_t)j * L + i] + params[(size_t)i * L + j]);
  }
  p.nn_w = p.cparam[0];  // parameters()(0): v[0] or W(0,0)
  p.cdev.resize(p.cparam.size());
  DCRF_HIP(hipMemcpyAsync(p.cdev.get(), p.cparam.data(), p.cparam.size() * sizeof(float), hipMemcpyHostToDevice, s));
  DCRF_HIP(hipStreamSynchronize(s));
}

int potential_nparams(const Potential& p, int L) {
  return p.ctype == kPotts ? 1 : p.ctype == kDiagonal ? L : L * (L + 1) / 2;
}

void potential_parameters(const Potential& p, int L, float* out) {
  if (p.ctype == kPotts) {
    out[0] = p.w;
  } else if (p.ctype == kDiagonal) {
    std::copy(p.cparam.begin(), p.cparam.end(), out);
  } else {
    for (int i = 0, k = 0; i < L; ++i)
      for (int j = i; j < L; ++j, ++k) out[k] = p.cparam[(size_t)j * L + i];
  }
}

void potential_set_parameters(Potential& p, int L, const float* v, hipStream_t s) {
  if (p.ctype == kPotts) {
    p.w = v[0];
    return;
  }
  DCRF_REQUIRE(L == p.cL, "setParameters: label count differs from the compatibility's");
  if (p.ctype == kDiagonal) {
    std::copy(v, v + L, p.cparam.begin());
  } else {
    for (int i = 0, k = 0; i < L; ++i)
      for (int j = i; j < L; ++j, ++k) p.cparam[(size_t)j * L + i] = p.cparam[(size_t)i * L + j] = v[k];
  }
  DCRF_HIP(hipMemcpyAsync(p.cdev.get(), p.cparam.data(), p.cparam.size() * sizeof(float), hipMemcpyHostToDevice, s));
  DCRF_HIP(hipStreamSynchronize(s));
}

void compat_slice(Potential& p, const SliceArgs& o, int V, hipStream_t s,
                  const std::function<void(const SliceArgs&)>& filt) {
  if (p.ctype == kPotts) {
    SliceArgs q = o;
    q.scale = -p.w;
    q.apply_scale = true;
    filt(q);
    return;
  }
  DCRF_REQUIRE(V == p.cL, "label compatibility: field has a different label count");
  const int N = p.lat->N();
  p.ctmp.resize((size_t)N * V);
  SliceArgs q = o;
  q.out = p.ctmp.get();
  q.accumulate = false;
  q.apply_scale = false;
  filt(q);
  KTimer tm("compat", s);
  k_compat_rg<<<rg_grid(N, V, 4096), kBlock, 0, s>>>(N, V, p.ctype, p.cdev.get(), p.ctmp.get(), o.out,
                                                     o.accumulate ? 1 : 0);
  check_launch("k_compat_rg");
}

void compat_apply_rows(Potential& p, const float* in, float* out, long nrows, int V, hipStream_t s) {
  DCRF_REQUIRE(p.ctype != kPotts, "compat_apply_rows: Potts is a scale");
  DCRF_REQUIRE(V == p.cL, "label compatibility: field has a different label count");
  if (nrows <= 0) return;
  k_compat_rg<<<rg_grid(nrows, V, 4096), kBlock, 0, s>>>((int)nrows, V, p.ctype, p.cdev.get(), in, out, 0);
  check_launch("k_compat_rg");
}

// compatibility gradient of b against an already filtered field m, into acc
static void compat_grad_accum_dev(int ctype, int N, int L, const float* b, const float* m, double* acc,
                                  double* parts, hipStream_t s) {
  const long tiles = ((long)N + 64 / L - 1) / (64 / L);
  const unsigned nb = (unsigned)std::max(1L, std::min<long>(tiles, kOuterBlocks));
  if (L == 21)
    k_outer_w<21><<<nb, 64, 0, s>>>(N, L, b, m, parts);
  else if (L <= 8)
    k_outer_w<8><<<nb, 64, 0, s>>>(N, L, b, m, parts);
  else if (L <= 16)
    k_outer_w<16><<<nb, 64, 0, s>>>(N, L, b, m, parts);
  else if (L <= 32)
    k_outer_w<32><<<nb, 64, 0, s>>>(N, L, b, m, parts);
  else
    k_outer_w<64><<<nb, 64, 0, s>>>(N, L, b, m, parts);
  check_launch("k_outer_w");
  double* G = parts + (size_t)kOuterBlocks * L * L;
  k_outer_sum<<<L * L, kBlock, 0, s>>>(parts, (int)nb, L * L, G);
  k_outer_pack<<<1, kBlock, 0, s>>>(G, L, ctype, acc);
  check_launch("k_outer_pack");
}

// scratch doubles needed by compat_grad_accum_dev
static size_t outer_parts_size(int L) { return (size_t)(kOuterBlocks + 1) * L * L; }

void potential_gradient_dev(Potential& p, int L, const float* b, const float* Q, double* acc, double* parts,
                            hipStream_t s) {
  const int N = p.lat->N();
  DBuf<float> fq((size_t)N * L);
  // DenseKernel::apply: the normalised filter without the compatibility
  SplatArgs a;
  const int nt = p.ntype;
  const bool pre = nt == kNormSymmetric || nt == kNormBefore;
  const bool post = nt == kNormSymmetric || nt == kNormAfter;
  a.kind = pre ? SplatIn::kScaled : SplatIn::kPlain;
  a.in = Q;
  a.norm = p.norm.get();
  SliceArgs o;
  o.out = fq.get();
  o.norm = post ? p.norm.get() : nullptr;
  p.lat->filter(a, o, L, false);
  compat_grad_accum_dev(p.ctype, N, L, b, fq.get(), acc, parts, s);
  DCRF_HIP(hipStreamSynchronize(s));
}

// ---------------------------------------------------------------------------
// mean field
// ---------------------------------------------------------------------------
struct CRF::MFState {
  int T = -1;
  unsigned long long model_gen = 0;  // model generation the checkpoints belong to
  long long id = 0;                  // number of this forward run
  DBuf<float> Qs, Ms, g1, g2, b, dU;
  DBuf<double> parts, acc;
};

int CRF::meanfield_iterations() const { return mf_ ? mf_->T : -1; }
long long CRF::meanfield_id() const { return mf_ ? mf_->id : 0; }

// DenseKernel::apply / applyTranspose of potential p (no compatibility), out (+)= K in
static void kernel_filter(Potential& p, const float* in, float* out, int V, bool transpose) {
  const int nt = p.ntype;
  SplatArgs a;
  a.kind = (nt == kNormSymmetric || (nt == kNormBefore && !transpose) || (nt == kNormAfter && transpose))
               ? SplatIn::kScaled
               : SplatIn::kPlain;
  a.in = in;
  a.norm = p.norm.get();
  SliceArgs o;
  o.out = out;
  o.norm = (nt == kNormSymmetric || (nt == kNormBefore && transpose) || (nt == kNormAfter && !transpose))
               ? p.norm.get()
               : nullptr;
  p.lat->filter(a, o, V, transpose);
}

void CRF::meanfield(int T, float* Qout) {
  DCRF_REQUIRE(T >= 0 && T <= 1000, "meanfield: 0..1000 iterations");
  const int K = num_pairwise();
  DCRF_REQUIRE(K <= kMaxPw, "meanfield: at most 8 pairwise potentials");
  for (int k = 0; k < K; ++k)
    DCRF_REQUIRE(pairwise(k).ctype == kPotts || pairwise(k).cL == L_,
                 "meanfield: compatibility built for another label count");
  if (!mf_) mf_ = std::make_unique<MFState>();
  MFState& st = *mf_;
  const size_t NL = (size_t)N_ * L_;
  hipStream_t s = s_;
  st.Qs.resize(NL * (T + 1));
  st.Ms.resize(std::max<size_t>(NL * T * K, 1));
  st.T = T;
  st.model_gen = model_gen_;
  st.id = ++mf_count_;
  k_softmax_neg<<<grid_for(N_), kBlock, 0, s>>>(U_.get(), st.Qs.get(), N_, L_);
  check_launch("k_softmax_neg");
  const unsigned grid = rg_grid(N_, L_, 8192);
  for (int t = 0; t < T; ++t) {
    const float* Q = st.Qs.get() + NL * t;
    MFArgs A{};
    A.K = K;
    for (int k = 0; k < K; ++k) {
      Potential& p = pairwise(k);
      float* m = st.Ms.get() + NL * ((size_t)t * K + k);
      kernel_filter(p, Q, m, L_, false);
      A.t[k] = MFTerm{m, p.ctype == kPotts ? nullptr : p.cdev.get(), p.ctype, p.w};
    }
    KTimer tm("mf_update", s);
    k_mf_update_rg<<<grid, kBlock, 0, s>>>(N_, L_, U_.get(), A, st.Qs.get() + NL * (t + 1));
    check_launch("k_mf_update_rg");
  }
  if (Qout)
    DCRF_HIP(hipMemcpyAsync(Qout, st.Qs.get() + NL * T, NL * sizeof(float), hipMemcpyDeviceToDevice, s));
}

void CRF::meanfield_backward(const float* dQ, float* dU, double* dparams) {
  DCRF_REQUIRE(mf_ && mf_->T >= 0, "meanfield_backward: run meanfield first");
  MFState& st = *mf_;
  DCRF_REQUIRE(st.model_gen == model_gen_,
               "meanfield_backward: the unary or a pairwise term changed after the last meanfield");
  const int T = st.T, K = num_pairwise();
  const size_t NL = (size_t)N_ * L_;
  hipStream_t s = s_;
  std::vector<int> goff(K + 1, 0);
  for (int k = 0; k < K; ++k) goff[k + 1] = goff[k] + potential_nparams(pairwise(k), L_);
  st.g1.resize(NL);
  st.g2.resize(NL);
  st.b.resize(NL);
  st.parts.resize(outer_parts_size(L_));
  st.acc.resize(std::max(goff[K], 1));
  DCRF_HIP(hipMemcpyAsync(st.g1.get(), dQ, NL * sizeof(float), hipMemcpyDeviceToDevice, s));
  k_zero_scalars<<<1, 256, 0, s>>>(st.acc.get(), std::max(goff[K], 1), nullptr, 0);
  const unsigned grid = rg_grid(N_, L_, 8192);
  float* gQ = st.g1.get();
  float* gprev = st.g2.get();
  for (int t = T; t >= 1; --t) {
    {
      KTimer tm("mf_back", s);
      k_mf_back_rg<<<grid, kBlock, 0, s>>>(N_, L_, st.Qs.get() + NL * t, gQ, dU, t == T ? 1 : 0, st.b.get());
      check_launch("k_mf_back_rg");
    }
    for (int k = 0; k < K; ++k) {
      Potential& p = pairwise(k);
      {
        KTimer tm("mf_grad", s);
        compat_grad_accum_dev(p.ctype, N_, L_, st.b.get(), st.Ms.get() + NL * ((size_t)(t - 1) * K + k),
                              st.acc.get() + goff[k], st.parts.get(), s);
      }
      // gQ_{t-1} (+)= applyTranspose_k(b) (src/pairwise.cpp:209-213)
      SliceArgs o;
      o.out = gprev;
      o.accumulate = k > 0;
      const int nt = p.ntype;
      o.norm = (nt == kNormSymmetric || nt == kNormBefore) ? p.norm.get() : nullptr;
      SplatArgs a;
      a.kind = (nt == kNormSymmetric || nt == kNormAfter) ? SplatIn::kScaled : SplatIn::kPlain;
      a.in = st.b.get();
      a.norm = p.norm.get();
      compat_slice(p, o, L_, s, [&](const SliceArgs& oo) { p.lat->filter(a, oo, L_, true); });
    }
    if (K == 0) k_fill<<<grid_for(NL), kBlock, 0, s>>>(gprev, NL, 0.f);
    std::swap(gQ, gprev);
  }
  {
    KTimer tm("mf_back", s);
    k_mf_back_rg<<<grid, kBlock, 0, s>>>(N_, L_, st.Qs.get(), gQ, dU, T == 0 ? 1 : 0, nullptr);
    check_launch("k_mf_back_rg");
  }
  if (dparams && goff[K] > 0)
    DCRF_HIP(hipMemcpyAsync(dparams, st.acc.get(), goff[K] * sizeof(double), hipMemcpyDeviceToHost, s));
  DCRF_HIP(hipStreamSynchronize(s));
}

double CRF::cross_entropy(const float* Q, const int* lab, float* dQ) {
  Impl& m = *im_;
  k_xent_rg<<<kRedBlocks, kBlock, 0, s_>>>(N_, L_, Q, lab, dQ, m.parts.get());
  check_launch("k_xent_rg");
  finalize_sums<1><<<1, kBlock, 0, s_>>>(m.parts.get(), kRedBlocks, m.sc.get() + kS_TMP);
  double e = 0;
  DCRF_HIP(hipMemcpyAsync(&e, m.sc.get() + kS_TMP, sizeof(double), hipMemcpyDeviceToHost, s_));
  DCRF_HIP(hipStreamSynchronize(s_));
  return e;
}

void CRF::pairwise_gradient(int k, const float* b, const float* Q, float* host_out) {
  Potential& p = pairwise(k);
  const int np = potential_nparams(p, L_);
  DBuf<double> parts(outer_parts_size(L_)), acc(np);
  k_zero_scalars<<<1, 256, 0, s_>>>(acc.get(), np, nullptr, 0);
  potential_gradient_dev(p, L_, b, Q, acc.get(), parts.get(), s_);
  std::vector<double> h(np);
  DCRF_HIP(hipMemcpy(h.data(), acc.get(), np * sizeof(double), hipMemcpyDeviceToHost));
  for (int i = 0; i < np; ++i) host_out[i] = (float)h[i];
}

}  // namespace dcrf
