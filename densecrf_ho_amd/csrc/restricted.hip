// Restricted ordered filter: Permutohedral::compute_upper_minus_lower_ord_restricted
// (src/permutohedral.cpp:702-880, 1074-1078).
//
// The ordered filter evaluated at a subset pI of the pixels (ascending).  The
// pixels of pI are splatted with their own values rq (|pI| x V); every other
// pixel with its value in a full N x V field.  The lattice rows are the
// reference's split arrays: 10 bins per (point, value), point p at row p + 1,
// row 0 the zero row of absent neighbours.
//
// State kept on the lattice between calls, as the reference keeps it in its
// Permutohedral object:
//   * the splat of the pixels outside pI, one copy per pass, and
//   * the active points (those a pixel of pI touches),
// both written by a call with `store`; a call without `store` adds its pI
// pixels to the stored splat and blurs the stored active points.  The blur
// updates active points only and swaps two row buffers per axis, so an
// inactive point reads as its splat after an even number of axes and as 0
// after an odd number (the buffer that was cleared); the kernels keep exactly
// that buffer discipline.  The output is added to, as in the reference.
#include "lattice.hpp"

namespace dcrf {

struct RestrictedState {
  int V = 0;
  bool have = false;
  int nact = 0;
  DBuf<float> stored[2];  // splat of the pixels outside pI, lower / upper pass
  DBuf<float> A, B;       // row buffers of the blur
  DBuf<int> active;       // active point ids
  DBuf<int> rmap;         // pixel -> index in pI, or -1
  DBuf<int> flag, pos;    // active flags and their exclusive scan (M + 1)
};

namespace {

constexpr int kRBins = 10;  // RESOLUTION, include/permutohedral.h:41

__global__ void __launch_bounds__(kBlock) k_rmap(const int* __restrict__ pI, int nI, int* __restrict__ rmap) {
  const int ii = blockIdx.x * blockDim.x + threadIdx.x;
  if (ii < nI) rmap[pI[ii]] = ii;
}

__global__ void __launch_bounds__(kBlock) k_rflag(const int* __restrict__ pI, int nI, const int* __restrict__ offset,
                                                  int d1, int* __restrict__ flag) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)nI * d1) return;
  const int ii = (int)(t / d1), j = (int)(t % d1);
  flag[offset[(size_t)pI[ii] * d1 + j]] = 1;
}

__global__ void __launch_bounds__(kBlock) k_rcompact(const int* __restrict__ flag, const int* __restrict__ pos, int M,
                                                     int* __restrict__ act) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < M && flag[i]) act[pos[i]] = i;
}

// addSplitArray (src/permutohedral.cpp:548-569): w into bins 0..bin_lo (lower)
// or bin_up..9 (upper)
template <bool UP>
__device__ inline void split_add(float (&acc)[kRBins], float w, float q) {
  if (UP) {
    const int c0 = bin_up(q);
#pragma unroll
    for (int c = 0; c < kRBins; ++c) acc[c] = c >= c0 ? acc[c] + w : acc[c];
  } else {
    const int c1 = bin_lo(q);
#pragma unroll
    for (int c = 0; c < kRBins; ++c) acc[c] = c <= c1 ? acc[c] + w : acc[c];
  }
}

// One lane per (point, value): the point's entries in pixel order, first the
// pixels outside pI (from ext, stored when `store`), then those of pI (from
// rq).  Every point's row is written, active or not.
template <bool UP>
__global__ void __launch_bounds__(kBlock) k_rsplat(int M, int V, const int* __restrict__ seg,
                                                   const int* __restrict__ sp_pix, const float* __restrict__ sp_w,
                                                   const int* __restrict__ rmap, const float* __restrict__ ext,
                                                   const float* __restrict__ rq, int store, float* __restrict__ stored,
                                                   float* __restrict__ A) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)M * V) return;
  const int i = (int)(t / V), k = (int)(t % V);
  const size_t row = ((size_t)(i + 1) * V + k) * kRBins;
  const int e0 = seg[i], e1 = seg[i + 1];
  float acc[kRBins];
  if (store) {
#pragma unroll
    for (int c = 0; c < kRBins; ++c) acc[c] = 0.f;
    for (int e = e0; e < e1; ++e) {
      const int pix = sp_pix[e];
      if (rmap[pix] >= 0) continue;
      split_add<UP>(acc, sp_w[e], ext[(size_t)pix * V + k]);
    }
#pragma unroll
    for (int c = 0; c < kRBins; ++c) stored[row + c] = acc[c];
  } else {
#pragma unroll
    for (int c = 0; c < kRBins; ++c) acc[c] = stored[row + c];
  }
  for (int e = e0; e < e1; ++e) {
    const int ii = rmap[sp_pix[e]];
    if (ii < 0) continue;
    split_add<UP>(acc, sp_w[e], rq[(size_t)ii * V + k]);
  }
#pragma unroll
  for (int c = 0; c < kRBins; ++c) A[row + c] = acc[c];
}

// weightedAddSplitArray (:570-578) on the active points of one axis; one lane
// per (active point, value, bin)
__global__ void __launch_bounds__(kBlock) k_rblur(const int* __restrict__ act, int nact, const int* __restrict__ nbr_j,
                                                  int VR, const float* __restrict__ src, float* __restrict__ dst) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)nact * VR) return;
  const int a = (int)(t / VR), r = (int)(t % VR);
  const int i = act[a];
  const int n1 = nbr_j[2 * i] + 1, n2 = nbr_j[2 * i + 1] + 1;
  dst[(size_t)(i + 1) * VR + r] = src[(size_t)(i + 1) * VR + r] + 0.5f * (src[(size_t)n1 * VR + r] + src[(size_t)n2 * VR + r]);
}

// sliceSplitArray (:579-589) for the pixels of pI: out += bin * (-w alpha)
// (lower) or bin * (w alpha) (upper), vertices in order
template <bool UP>
__global__ void __launch_bounds__(kBlock) k_rslice(int nI, int V, int d1, const int* __restrict__ pI,
                                                   const int* __restrict__ offset, const float* __restrict__ bary,
                                                   const float* __restrict__ rq, const float* __restrict__ vals,
                                                   float alpha, float* __restrict__ out) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)nI * V) return;
  const int ii = (int)(t / V), k = (int)(t % V);
  const size_t pix = (size_t)pI[ii];
  const float q = rq[t];
  const int c = UP ? bin_up(q) : bin_lo(q);
  float o = out[t];
  for (int j = 0; j < d1; ++j) {
    const int p = offset[pix * d1 + j] + 1;
    const float w = bary[pix * d1 + j];
    const float coef = UP ? w * alpha : -w * alpha;
    o = o + vals[((size_t)p * V + k) * kRBins + c] * coef;
  }
  out[t] = o;
}

__global__ void __launch_bounds__(kBlock) k_rscale(float* __restrict__ x, size_t n, float a) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) x[t] = a * x[t];
}

// less_confident_pixels (src/densecrf.cpp:457-462): flag of every pixel whose
// largest label value is <= tol
__global__ void __launch_bounds__(kBlock) k_lc_flag(const float* __restrict__ Q, int N, int L, float tol,
                                                    int* __restrict__ flag) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i > N) return;
  if (i == N) {
    flag[N] = 0;
    return;
  }
  const float* col = Q + (size_t)i * L;
  float mx = col[0];
  for (int l = 1; l < L; ++l) mx = col[l] > mx ? col[l] : mx;
  flag[i] = mx <= tol ? 1 : 0;
}

}  // namespace

// the pixel indices (ascending) of less_confident_pixels into idx (device,
// N entries); returns their count
int less_confident_pixels_dev(const float* Q, int N, int L, float tol, int* idx, hipStream_t s) {
  DBuf<int> flag(N + 1), pos(N + 1);
  k_lc_flag<<<grid_for((size_t)N + 1), kBlock, 0, s>>>(Q, N, L, tol, flag.get());
  scan_excl(flag.get(), pos.get(), N + 1, s);
  k_rcompact<<<grid_for(N), kBlock, 0, s>>>(flag.get(), pos.get(), N, idx);
  check_launch("less_confident_pixels");
  int n = 0;
  DCRF_HIP(hipMemcpyAsync(&n, pos.get() + N, sizeof(int), hipMemcpyDeviceToHost, s));
  DCRF_HIP(hipStreamSynchronize(s));
  return n;
}

void Lattice::ordered_restricted(const float* rq, const int* pI, int nI, const float* ext, int V, bool store,
                                 float* out) {
  if (!rs_) rs_ = std::make_shared<RestrictedState>();
  RestrictedState& r = *rs_;
  DCRF_REQUIRE(store || (r.have && r.V == V),
               "restricted ordered filter: no stored splat for this value count (call with store first)");
  DCRF_REQUIRE(nchunks_ >= 0 && N_ > 0, "restricted ordered filter: lattice not built");
  const int M = M_, d1 = d_ + 1;
  const int VR = V * kRBins;
  const size_t rows = (size_t)(M + 1) * VR;
  hipStream_t s = stream_;
  r.rmap.resize(N_);
  k_fill_int<<<grid_for(N_, kBlock, 4096), kBlock, 0, s>>>(r.rmap.get(), N_, -1);
  if (nI > 0) k_rmap<<<grid_for(nI), kBlock, 0, s>>>(pI, nI, r.rmap.get());
  if (store) {
    r.have = false;
    r.V = V;
    r.stored[0].resize(rows);
    r.stored[1].resize(rows);
    r.flag.resize(M + 1);
    r.pos.resize(M + 1);
    k_fill_int<<<grid_for(M + 1, kBlock, 4096), kBlock, 0, s>>>(r.flag.get(), M + 1, 0);
    if (nI > 0) k_rflag<<<grid_for((size_t)nI * d1), kBlock, 0, s>>>(pI, nI, offset_.get(), d1, r.flag.get());
    scan_excl(r.flag.get(), r.pos.get(), M + 1, s);
    DCRF_HIP(hipMemcpyAsync(&r.nact, r.pos.get() + M, sizeof(int), hipMemcpyDeviceToHost, s));
    DCRF_HIP(hipStreamSynchronize(s));
    r.active.resize(std::max(r.nact, 1));
    k_rcompact<<<grid_for(M), kBlock, 0, s>>>(r.flag.get(), r.pos.get(), M, r.active.get());
  }
  check_launch("restricted setup");
  r.A.resize(rows);
  r.B.resize(rows);
  const float alpha = 1.0f / (1 + powf(2.f, (float)-d_));
  for (int pass = 0; pass < 2; ++pass) {
    const bool up = pass == 1;
    DCRF_HIP(hipMemsetAsync(r.A.get(), 0, (size_t)VR * sizeof(float), s));  // row 0
    DCRF_HIP(hipMemsetAsync(r.B.get(), 0, rows * sizeof(float), s));
    const unsigned gs = grid_for((size_t)M * V);
    if (up)
      k_rsplat<true><<<gs, kBlock, 0, s>>>(M, V, seg_.get(), sp_pix_.get(), sp_w_.get(), r.rmap.get(), ext, rq,
                                           store ? 1 : 0, r.stored[1].get(), r.A.get());
    else
      k_rsplat<false><<<gs, kBlock, 0, s>>>(M, V, seg_.get(), sp_pix_.get(), sp_w_.get(), r.rmap.get(), ext, rq,
                                            store ? 1 : 0, r.stored[0].get(), r.A.get());
    float* src = r.A.get();
    float* dst = r.B.get();
    if (r.nact > 0)
      for (int j = 0; j < d1; ++j) {
        k_rblur<<<grid_for((size_t)r.nact * VR), kBlock, 0, s>>>(r.active.get(), r.nact, nbr_.get() + (size_t)j * M * 2,
                                                                 VR, src, dst);
        std::swap(src, dst);
      }
    else if (d1 % 2)  // no active point: the buffers still swap once per axis
      src = r.B.get();
    if (nI > 0) {
      const unsigned gl = grid_for((size_t)nI * V);
      if (up)
        k_rslice<true><<<gl, kBlock, 0, s>>>(nI, V, d1, pI, offset_.get(), bary_.get(), rq, src, alpha, out);
      else
        k_rslice<false><<<gl, kBlock, 0, s>>>(nI, V, d1, pI, offset_.get(), bary_.get(), rq, src, alpha, out);
    }
    check_launch("restricted ordered filter");
  }
  if (store) r.have = true;
}

// rout.fill(0), the kernel's restricted filter, compatibility_->apply(rout, rout)
void potential_apply_ord_restricted(Potential& p, float* rout, const float* rq, const int* pI, int nI,
                                    const float* ext, int V, bool store) {
  hipStream_t s = p.lat->stream();
  const size_t n = (size_t)nI * V;
  float* tgt = rout;
  if (p.ctype != kPotts) {
    p.ctmp.resize(std::max<size_t>(n, 1));
    tgt = p.ctmp.get();
  }
  if (n) k_fill<<<grid_for(n), kBlock, 0, s>>>(tgt, n, 0.f);
  p.lat->ordered_restricted(rq, pI, nI, ext, V, store, tgt);
  if (p.ctype == kPotts) {
    if (n) k_rscale<<<grid_for(n), kBlock, 0, s>>>(rout, n, -p.w);  // -w * Q (src/labelcompatibility.cpp:49-52)
    check_launch("k_rscale");
  } else {
    compat_apply_rows(p, tgt, rout, nI, V, s);
  }
}

}  // namespace dcrf
