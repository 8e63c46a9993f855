// Brute-force (dense Gaussian) forms of the pairwise filters, the reference's
// test and comparison paths (no lattice):
//   * PairwisePotential::apply_bf (src/pairwise.cpp:368-383):
//       out(l, c) = sum_b exp(-|f_c - f_b|^2) Q(l, b);
//   * apply_upper_minus_lower_bf (src/pairwise.cpp:294-320): per label, with
//     the pixels in the order ind(l, .),
//       out(l, c) = sum_{b > c} K(c, b) - sum_{b < c} K(c, b);
//   * apply_upper_minus_lower_bf_ord (src/pairwise.cpp:321-353): the same
//     with the split by value instead of position,
//       out(l, c) = sum_{Q_b <= Q_c} K(c, b) - sum_{Q_b >= Q_c} K(c, b),
//     Q_x = Q(l, ind(l, x));
// each followed by the compatibility.  The two sorted forms leave row l in
// the sorted order (position c is pixel ind(l, c)), as the reference does.
//
// One thread per output pixel c; the b loop runs in the reference's order
// (ascending, the upper sum before the lower one) over tiles of pixels staged
// in LDS, so every sum keeps its order; each term is the double exp added to
// the float sum before rounding, as the reference's float += double does.
// O(N^2 L) work: for the small inputs these paths are meant for.
#include "crf.hpp"

namespace dcrf {
namespace {

constexpr int kBfTile = 64;
constexpr int kBfMaxD = 16;

// exp(-|a - b|^2) as the reference evaluates it: the float squared norm in
// Eigen 3.2.7's SSE2 reduction order for a fresh VectorXf (one or two 4-wide
// partial sums, (p0 + p2) + (p1 + p3), then the tail in order;
// include/Eigen/src/Core/Redux.h:193-245), exp of the promoted value in double
__device__ inline double bf_kernel(const float* a, const float* b, int d) {
  float t[kBfMaxD];
  for (int k = 0; k < d; ++k) {
    const float x = a[k] - b[k];
    t[k] = x * x;
  }
  float r;
  if (d < 4) {
    r = t[0];
    for (int i = 1; i < d; ++i) r = r + t[i];
  } else {
    const int a1 = d / 4 * 4, a2 = d / 8 * 8;
    float p0[4] = {t[0], t[1], t[2], t[3]};
    if (a1 > 4) {
      float p1[4] = {t[4], t[5], t[6], t[7]};
      for (int i = 8; i < a2; i += 8)
        for (int q = 0; q < 4; ++q) {
          p0[q] = p0[q] + t[i + q];
          p1[q] = p1[q] + t[i + 4 + q];
        }
      for (int q = 0; q < 4; ++q) p0[q] = p0[q] + p1[q];
      if (a1 > a2)
        for (int q = 0; q < 4; ++q) p0[q] = p0[q] + t[a2 + q];
    }
    r = (p0[0] + p0[2]) + (p0[1] + p0[3]);
    for (int i = a1; i < d; ++i) r = r + t[i];
  }
  return exp((double)-r);
}

template <int MAXL>
__global__ void __launch_bounds__(kBfTile) k_bf_apply(int N, int d, int L, const float* __restrict__ f,
                                                      const float* __restrict__ Q, float* __restrict__ out) {
  __shared__ float sf[kBfTile * kBfMaxD];
  __shared__ float sq[kBfTile * MAXL];
  const int c = blockIdx.x * kBfTile + threadIdx.x;
  float fc[kBfMaxD];
  for (int k = 0; k < d; ++k) fc[k] = c < N ? f[(size_t)c * d + k] : 0.f;
  float acc[MAXL];
#pragma unroll
  for (int l = 0; l < MAXL; ++l) acc[l] = 0.f;
  for (int b0 = 0; b0 < N; b0 += kBfTile) {
    const int nb = min(kBfTile, N - b0);
    __syncthreads();
    for (int i = threadIdx.x; i < nb * d; i += kBfTile) sf[i] = f[(size_t)b0 * d + i];
    for (int i = threadIdx.x; i < nb * L; i += kBfTile) sq[(i / L) * MAXL + i % L] = Q[(size_t)b0 * L + i];
    __syncthreads();
    for (int b = 0; b < nb; ++b) {
      const double K = bf_kernel(fc, &sf[b * d], d);
#pragma unroll
      for (int l = 0; l < MAXL; ++l)
        if (l < L) acc[l] = (float)((double)acc[l] + K * (double)sq[b * MAXL + l]);
    }
  }
  if (c < N)
    for (int l = 0; l < L; ++l) out[(size_t)c * L + l] = acc[l];
}

// grid (tiles of c, labels)
template <bool ORD>
__global__ void __launch_bounds__(kBfTile) k_bf_uml(int N, int d, int L, const float* __restrict__ f,
                                                    const int* __restrict__ ind, const float* __restrict__ Q,
                                                    float* __restrict__ out) {
  __shared__ float sf[kBfTile * kBfMaxD];
  __shared__ float sq[kBfTile];
  const int l = blockIdx.y;
  const int c = blockIdx.x * kBfTile + threadIdx.x;
  float fc[kBfMaxD];
  float qc = 0.f;
  if (c < N) {
    const size_t pc = (size_t)ind[(size_t)c * L + l];
    for (int k = 0; k < d; ++k) fc[k] = f[pc * d + k];
    if (ORD) qc = Q[pc * L + l];
  } else {
    for (int k = 0; k < d; ++k) fc[k] = 0.f;
  }
  float acc = 0.f;
  for (int pass = 0; pass < 2; ++pass) {  // upper (:300-305 / :329-335), then lower (:307-312 / :337-343)
    for (int b0 = 0; b0 < N; b0 += kBfTile) {
      const int nb = min(kBfTile, N - b0);
      __syncthreads();
      for (int i = threadIdx.x; i < nb * d; i += kBfTile) {
        const int b = i / d;
        sf[i] = f[(size_t)ind[(size_t)(b0 + b) * L + l] * d + i % d];
      }
      if (ORD)
        for (int i = threadIdx.x; i < nb; i += kBfTile) sq[i] = Q[(size_t)ind[(size_t)(b0 + i) * L + l] * L + l];
      __syncthreads();
      for (int b = 0; b < nb; ++b) {
        const int bb = b0 + b;
        bool take;
        if (ORD)
          take = pass == 0 ? !(sq[b] > qc) : !(sq[b] < qc);
        else
          take = pass == 0 ? bb > c : bb < c;
        if (!take) continue;
        const double K = bf_kernel(fc, &sf[b * d], d);
        acc = (float)(pass == 0 ? (double)acc + K : (double)acc - K);
      }
    }
  }
  if (c < N) out[(size_t)c * L + l] = acc;
}

}  // namespace

// compatibility_->apply(out, out) after a brute-force filter into `tgt`
static void bf_compat(Potential& p, float* tgt, float* out, int N, int L, hipStream_t s) {
  const size_t n = (size_t)N * L;
  if (p.ctype == kPotts) {
    k_rscale<<<grid_for(n), kBlock, 0, s>>>(out, n, -p.w);
    check_launch("k_rscale");
  } else {
    compat_apply_rows(p, tgt, out, N, L, s);
  }
}

void potential_apply_bf(Potential& p, const float* Q, float* out, int L) {
  DCRF_REQUIRE(p.feat.get() != nullptr, "brute-force filter: the potential keeps no features");
  DCRF_REQUIRE(p.d <= kBfMaxD && L <= 64, "brute-force filter: at most 16 features and 64 labels");
  hipStream_t s = p.lat->stream();
  const int N = p.lat->N();
  float* tgt = out;
  if (p.ctype != kPotts) {
    p.ctmp.resize((size_t)N * L);
    tgt = p.ctmp.get();
  }
  const unsigned g = (unsigned)((N + kBfTile - 1) / kBfTile);
  DCRF_MAXL_DISPATCH(L, (k_bf_apply<MAXL><<<g, kBfTile, 0, s>>>(N, p.d, L, p.feat.get(), Q, tgt)));
  check_launch("k_bf_apply");
  bf_compat(p, tgt, out, N, L, s);
}

void potential_apply_uml_bf(Potential& p, const int* ind, const float* Q, float* out, int L, bool ord) {
  DCRF_REQUIRE(p.feat.get() != nullptr, "brute-force filter: the potential keeps no features");
  DCRF_REQUIRE(p.d <= kBfMaxD, "brute-force filter: at most 16 features");
  hipStream_t s = p.lat->stream();
  const int N = p.lat->N();
  float* tgt = out;
  if (p.ctype != kPotts) {
    p.ctmp.resize((size_t)N * L);
    tgt = p.ctmp.get();
  }
  const dim3 g((unsigned)((N + kBfTile - 1) / kBfTile), (unsigned)L);
  if (ord)
    k_bf_uml<true><<<g, kBfTile, 0, s>>>(N, p.d, L, p.feat.get(), ind, Q, tgt);
  else
    k_bf_uml<false><<<g, kBfTile, 0, s>>>(N, p.d, L, p.feat.get(), ind, nullptr, tgt);
  check_launch("k_bf_uml");
  bf_compat(p, tgt, out, N, L, s);
}

}  // namespace dcrf
