// Per-pixel and reduction kernels shared by the solvers.  Included once, by
// the engine's single translation unit.
#pragma once

#include <algorithm>

#include "common.hpp"

namespace dcrf {

constexpr int kMaxLabels = 64;

namespace {

__global__ void __launch_bounds__(kBlock) k_fill(float* __restrict__ p, size_t n, float v) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) p[t] = v;
}

__global__ void k_fill_int(int* __restrict__ p, int n, int v) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = v;
}
__global__ void k_add_scalar(double* dst, const double* src) { dst[0] = dst[0] + src[0]; }
__global__ void k_sub_scalar(double* dst, const double* src) { dst[0] = dst[0] - src[0]; }

__global__ void k_zero_scalars(double* __restrict__ sc, int ns, int* __restrict__ fl, int nf) {
  for (int i = threadIdx.x; i < ns; i += blockDim.x) sc[i] = 0.0;
  for (int i = threadIdx.x; i < nf; i += blockDim.x) fl[i] = 0;
}

// norm = 1 / sqrt(K1 + 1e-20) or 1 / (K1 + 1e-20), double then float
__global__ void __launch_bounds__(kBlock) k_norm_finish(float* __restrict__ n, int N, int symmetric) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const double v = (double)n[i] + 1e-20;
  n[i] = symmetric ? (float)(1.0 / sqrt(v)) : (float)(1.0 / v);
}

// Reduction order of Eigen 3.2.7 on SSE2 for a short dense vector whose first
// 16-byte-aligned element is `a` (include/Eigen/src/Core/Redux.h:193-245):
// two 4-wide accumulators over the aligned span, (p0+p2)+(p1+p3), then the
// unaligned head and tail in order.  `val(i)` supplies element i.
template <class F>
__device__ inline float eigen_sum(int n, int a, F val) {
  if (a > n) a = n;
  const int as2 = ((n - a) / 8) * 8, as1 = ((n - a) / 4) * 4;
  const int ae2 = a + as2, ae = a + as1;
  float res;
  if (as1) {
    float p0 = val(a), p1 = val(a + 1), p2 = val(a + 2), p3 = val(a + 3);
    if (as1 > 4) {
      float q0 = val(a + 4), q1 = val(a + 5), q2 = val(a + 6), q3 = val(a + 7);
      for (int i = a + 8; i < ae2; i += 8) {
        p0 = p0 + val(i);
        p1 = p1 + val(i + 1);
        p2 = p2 + val(i + 2);
        p3 = p3 + val(i + 3);
        q0 = q0 + val(i + 4);
        q1 = q1 + val(i + 5);
        q2 = q2 + val(i + 6);
        q3 = q3 + val(i + 7);
      }
      p0 = p0 + q0;
      p1 = p1 + q1;
      p2 = p2 + q2;
      p3 = p3 + q3;
      if (ae > ae2) {
        p0 = p0 + val(ae2);
        p1 = p1 + val(ae2 + 1);
        p2 = p2 + val(ae2 + 2);
        p3 = p3 + val(ae2 + 3);
      }
    }
    res = (p0 + p2) + (p1 + p3);
    for (int i = 0; i < a; ++i) res = res + val(i);
    for (int i = ae; i < n; ++i) res = res + val(i);
  } else {
    res = val(0);
    for (int i = 1; i < n; ++i) res = res + val(i);
  }
  return res;
}
// first aligned element of column `col` of an L-row matrix with aligned storage
__device__ inline int col_astart(long col, int L) { return (int)((4 - ((col * L) & 3)) & 3); }

// ---------------------------------------------------------------------------
// Row groups: fields are rows of L <= 64 values per pixel.  A wave covers
// gpw = floor(64/L) rows, lane -> (group g, value k), so global loads and
// stores coalesce across lanes.  Per-row reductions go through a wave-private
// 64-float LDS segment, read in the order the reference sums.  Kernels walk
// tiles of kBlock/64 waves with a block-uniform grid-stride loop so that
// __syncthreads() is reached by every thread.
// ---------------------------------------------------------------------------
struct RG {
  int g, k;
  long row;
  bool valid;
};
__device__ inline RG rg_at(int L, long nrows, long tile) {
  const int gpw = 64 / L;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  RG r;
  r.g = lane / L;
  r.k = lane - r.g * L;
  r.row = (tile * (kBlock / 64) + wave) * gpw + r.g;
  r.valid = r.g < gpw && r.row < nrows;
  return r;
}
__host__ __device__ inline long rg_tiles(long nrows, int L) {
  const long rows_per_tile = (long)(64 / L) * (kBlock / 64);
  return (nrows + rows_per_tile - 1) / rows_per_tile;
}
inline unsigned rg_grid(long nrows, int L, long cap = 1 << 30) {
  const long t = rg_tiles(nrows, L);
  return (unsigned)std::max(1L, std::min(t, cap));
}

// Row quads: the same tiling with one lane per (row, quad of values): a wave
// covers floor(64 / ceil(L/4)) rows (10 at L = 21) and each lane moves four
// values with one 16-byte access (rows of L floats are 4-byte aligned; the last
// quad of a row, when L % 4 != 0, moves its values one by one).  Per-row
// reductions go through a wave-private LDS segment of at most 256 floats.
struct RQ {
  int g, q;
  long row;
  bool valid;
};
__device__ inline RQ rq_at(int L, long nrows, long tile) {
  const int nq = (L + 3) >> 2, gpw = 64 / nq;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  RQ r;
  r.g = lane / nq;
  r.q = lane - r.g * nq;
  r.row = (tile * (kBlock / 64) + wave) * gpw + r.g;
  r.valid = r.g < gpw && r.row < nrows;
  return r;
}
__host__ __device__ inline long rq_tiles(long nrows, int L) {
  const long rows_per_tile = (long)(64 / ((L + 3) / 4)) * (kBlock / 64);
  return (nrows + rows_per_tile - 1) / rows_per_tile;
}
inline unsigned rq_grid(long nrows, int L, long cap = 1 << 30) {
  const long t = rq_tiles(nrows, L);
  return (unsigned)std::max(1L, std::min(t, cap));
}
// four consecutive values of a row starting at value k0 (n of them valid)
__device__ inline void ld_q4(const float* p, int n, float (&v)[4]) {
  if (n == 4) {
    const Float4u u = *reinterpret_cast<const Float4u*>(p);
    v[0] = u.x, v[1] = u.y, v[2] = u.z, v[3] = u.w;
  } else {
    for (int i = 0; i < 4; ++i) v[i] = i < n ? p[i] : 0.f;
  }
}
__device__ inline void st_q4(float* p, int n, const float (&v)[4]) {
  if (n == 4) {
    *reinterpret_cast<Float4u*>(p) = Float4u{v[0], v[1], v[2], v[3]};
  } else {
    for (int i = 0; i < n; ++i) p[i] = v[i];
  }
}

// ---------------------------------------------------------------------------
// Work split of a launch over images.  OneImage spreads every block over all
// N pixels.  ImageBlocks (a batch) gives each image its own contiguous range
// of blocks, so a block only touches one image's pixels, reads that image's
// scalars, and per-block partial sums combine per image.
// ---------------------------------------------------------------------------
struct SegPos {
  int img;
  long p0, n;   // first pixel and pixel count of the block's image
  long lb, nb;  // the block's index among its image's blocks, and their number
};
struct OneImage {
  long N;
  __device__ inline SegPos pos() const { return SegPos{0, 0, N, (long)blockIdx.x, (long)gridDim.x}; }
};
struct ImageBlocks {
  const int* blk_img;    // image of each block
  const int* blk_first;  // first block of each image (images + 1 entries)
  const long* pix;       // first pixel of each image (images + 1 entries)
  __device__ inline SegPos pos() const {
    const int i = blk_img[blockIdx.x];
    const int f = blk_first[i];
    return SegPos{i, pix[i], pix[i + 1] - pix[i], (long)blockIdx.x - f, (long)(blk_first[i + 1] - f)};
  }
};

// per-image sums of per-block partials: block i of the launch sums the NV
// partials of blocks [first[i], first[i+1]) (or [0, nparts) when first is
// null) in block order and writes them to out[i * stride ..]
template <int NV>
__global__ void finalize_seg(const double* __restrict__ parts, const int* __restrict__ first, int nparts,
                             double* __restrict__ out, int stride) {
  const int b0 = first ? first[blockIdx.x] : 0, b1 = first ? first[blockIdx.x + 1] : nparts;
  double v[NV];
#pragma unroll
  for (int i = 0; i < NV; ++i) v[i] = 0;
  for (int b = b0 + threadIdx.x; b < b1; b += blockDim.x)
#pragma unroll
    for (int i = 0; i < NV; ++i) v[i] += parts[(size_t)b * NV + i];
  block_sum<NV>(v);
  if (threadIdx.x == 0)
#pragma unroll
    for (int i = 0; i < NV; ++i) out[(size_t)blockIdx.x * stride + i] = v[i];
}

// less_confident_pixels (src/densecrf.cpp:457-462) as a count: per image, the
// number of pixels whose largest label value is <= tol, added to count[img]
// (integer sums, so the result does not depend on the grid)
template <class SEG>
__global__ void __launch_bounds__(kBlock) k_less_confident(SEG sg, int L, const float* __restrict__ Q, float tol,
                                                           int* __restrict__ count, int stride) {
  __shared__ int sc;
  const SegPos q = sg.pos();
  if (threadIdx.x == 0) sc = 0;
  __syncthreads();
  int c = 0;
  for (long i = q.lb * blockDim.x + threadIdx.x; i < q.n; i += q.nb * blockDim.x) {
    const float* col = Q + (size_t)(q.p0 + i) * L;
    float mx = col[0];
    for (int l = 1; l < L; ++l) mx = col[l] > mx ? col[l] : mx;
    c += mx <= tol ? 1 : 0;
  }
  if (c) atomicAdd(&sc, c);
  __syncthreads();
  if (threadIdx.x == 0 && sc) atomicAdd(count + (size_t)q.img * stride, sc);
}

// expAndNormalize(-U) (src/densecrf.cpp:328-336), one thread per pixel
__global__ void __launch_bounds__(kBlock) k_softmax_neg(const float* __restrict__ U, float* __restrict__ Q, int N, int L) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const float* u = U + (size_t)i * L;
  float* q = Q + (size_t)i * L;
  float mx = -u[0];
  for (int l = 1; l < L; ++l) mx = fmaxf(mx, -u[l]);
  for (int l = 0; l < L; ++l) q[l] = expf(-u[l] - mx);
  const float s = eigen_sum(L, 0, [&](int l) { return q[l]; });
  for (int l = 0; l < L; ++l) q[l] = q[l] / s;
}

// Σ a[i]*b[i] in FP64 over n elements: per-block partials, then one block
template <class SEG>
__global__ void __launch_bounds__(kBlock) k_dot(const float* __restrict__ a, const float* __restrict__ b, int L, SEG seg,
                                                double* __restrict__ parts) {
  const SegPos q = seg.pos();
  double v[1] = {0.0};
  const size_t e = (size_t)(q.p0 + q.n) * L;
  for (size_t i = (size_t)q.p0 * L + (size_t)q.lb * blockDim.x + threadIdx.x; i < e; i += (size_t)q.nb * blockDim.x)
    v[0] += (double)a[i] * (double)b[i];
  block_sum<1>(v);
  if (threadIdx.x == 0) parts[blockIdx.x] = v[0];
}
inline void launch_dot(hipStream_t s, const float* a, const float* b, size_t n, double* parts, double* out) {
  k_dot<<<kRedBlocks, kBlock, 0, s>>>(a, b, 1, OneImage{(long)n}, parts);
  check_launch("k_dot");
  finalize_sums<1><<<1, kBlock, 0, s>>>(parts, kRedBlocks, out);
  check_launch("finalize_sums");
}

// grad_y += 2 * tmp (src/densecrf.cpp:380) and partial Σ Q*tmp (:953)
template <class SEG>
__global__ void __launch_bounds__(kBlock) k_gy_init_accum(float* __restrict__ gy, const float* __restrict__ tmp, const float* __restrict__ Q,
                                int L, SEG seg, double* __restrict__ parts) {
  const SegPos q = seg.pos();
  double v[1] = {0.0};
  const size_t e = (size_t)(q.p0 + q.n) * L;
  for (size_t i = (size_t)q.p0 * L + (size_t)q.lb * blockDim.x + threadIdx.x; i < e; i += (size_t)q.nb * blockDim.x) {
    const float t = tmp[i];
    gy[i] = gy[i] + 2.f * t;
    v[0] += (double)Q[i] * (double)t;
  }
  block_sum<1>(v);
  if (threadIdx.x == 0) parts[blockIdx.x] = v[0];
}

// slot layout shared with crf.hip
constexpr int kS_ECUR = 0, kS_EPREV = 1, kS_NUM = 2, kS_DENOM = 3, kS_STEP = 4, kS_DOTS = 5, kS_EPART = 8,
              kS_HO = 9, kS_TMP = 17;
constexpr int kF_DONE = 0, kF_ITERS = 1;
// per-image strides of the QP scalar slots and flags
constexpr int kS_COUNT = 32, kF_COUNT = 8;
// higher-order scalar slots (relative to kS_HO)
constexpr int kH_A = 0, kH_B = 1, kH_SXZ = 2, kH_DEN = 3, kH_E = 4;

// E0 = U.Q - 0 + Σ_k Q.psi_k Q + 0 (+ higher order)  (:386-387, :940-957)
// (one thread per image; scalars of image i at sc + i * kS_COUNT)
__global__ void k_qp_e0(double* sc_all, int nk, int has_ho, double* st_energy, int cap, int nimg) {
  for (int img = threadIdx.x; img < nimg; img += blockDim.x) {
    double* sc = sc_all + (size_t)img * kS_COUNT;
    double e = 0;
    e += sc[kS_TMP + 0];
    e -= 0.0;
    for (int k = 0; k < nk; ++k) e += sc[kS_TMP + 1 + k];
    e += 0.0;
    if (has_ho) e += sc[kS_HO + kH_E];
    sc[kS_ECUR] = e;
    sc[kS_EPREV] = 1.7976931348623157e308;
    if (cap > 0) st_energy[0] = e;
  }
}

// descent_direction (src/qp.cpp:5-15): first minimum of each column
__global__ void __launch_bounds__(kBlock) k_argmin(const float* __restrict__ g, int* __restrict__ lab, int N, int L, const int* __restrict__ fl) {
  if (fl[kF_DONE]) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const float* r = g + (size_t)i * L;
  int m = 0;
  float v = r[0];
  for (int l = 1; l < L; ++l) {
    const float x = r[l];
    if (x < v) {
      v = x;
      m = l;
    }
  }
  lab[i] = m;
}

// first minimum of each row (descent_direction, src/qp.cpp:5-15), row groups
__global__ void __launch_bounds__(kBlock) k_argmin_rg(const float* __restrict__ g, int* __restrict__ lab, int N, int L) {
  __shared__ float seg[kBlock];
  const long ntiles = rg_tiles(N, L);
  for (long tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    const RG r = rg_at(L, N, tile);
    float* row = seg + (threadIdx.x & ~63) + r.g * L;
    if (r.valid) row[r.k] = g[(size_t)r.row * L + r.k];
    wave_lds_sync();  // a row group stays within its wave
    if (r.valid && r.k == 0) {
      int m = 0;
      float v = row[0];
      for (int l = 1; l < L; ++l)
        if (row[l] < v) {
          v = row[l];
          m = l;
        }
      lab[r.row] = m;
    }
    wave_lds_sync();
  }
}

// line-search sums U.sx, Q.psisx, sx.psisx with sx = onehot(lab) - Q (:408-414)
template <class SEG>
__global__ void __launch_bounds__(kBlock) k_qp_dots(const float* __restrict__ U, const float* __restrict__ Q, const int* __restrict__ lab,
                          const float* __restrict__ psx, SEG seg, FastDiv divL, double* __restrict__ parts,
                          const int* __restrict__ fl) {
  const SegPos qs = seg.pos();
  if (fl[qs.img * kF_COUNT + kF_DONE]) return;
  const int L = (int)divL.d;
  double v[3] = {0.0, 0.0, 0.0};
  const size_t e = (size_t)(qs.p0 + qs.n) * L;
  for (size_t i = (size_t)qs.p0 * L + (size_t)qs.lb * blockDim.x + threadIdx.x; i < e; i += (size_t)qs.nb * blockDim.x) {
    const uint32_t pix = divL.div((uint32_t)i);
    const int l = (int)(i - (size_t)pix * L);
    const float q = Q[i], p = psx[i];
    const float sx = (lab[pix] == l ? 1.f : 0.f) - q;
    v[0] += (double)U[i] * (double)sx;
    v[1] += (double)q * (double)p;
    v[2] += (double)sx * (double)p;
  }
  block_sum<3>(v);
  if (threadIdx.x == 0) {
    parts[blockIdx.x * 3 + 0] = v[0];
    parts[blockIdx.x * 3 + 1] = v[1];
    parts[blockIdx.x * 3 + 2] = v[2];
  }
}

// step size (:412-421); done = 2 when the reference breaks on denom==0||num==0
__global__ void k_qp_step(double* sc_all, int* fl_all, int has_ho, double* st_num, double* st_denom, double* st_step,
                          double* st_energy, int cap, int nimg) {
  for (int img = threadIdx.x; img < nimg; img += blockDim.x) {
  double* sc = sc_all + (size_t)img * kS_COUNT;
  int* fl = fl_all + (size_t)img * kF_COUNT;
  if (fl[kF_DONE]) continue;
  sc[kS_EPREV] = sc[kS_ECUR];
  double num = sc[kS_DOTS + 0] + 2 * sc[kS_DOTS + 1];
  double denom = sc[kS_DOTS + 2];
  if (has_ho) {
    num = num + sc[kS_HO + kH_A] + sc[kS_HO + kH_B] + sc[kS_HO + kH_SXZ];
    denom = denom + sc[kS_HO + kH_DEN];
  }
  double step = -num / (2 * denom);
  const int it = fl[kF_ITERS]++;  // iterations entered, as the reference's counter
  if (it < cap) {
    st_num[it] = num;
    st_denom[it] = denom;
  }
  sc[kS_NUM] = num;
  sc[kS_DENOM] = denom;
  if (denom == 0 || num == 0) {
    if (it + 1 < cap + 1) st_energy[it + 1] = sc[kS_ECUR];  // no update: energy unchanged
    fl[kF_DONE] = 2;
    continue;
  }
  if (denom < 0) step = 1;
  sc[kS_STEP] = step;
  if (it < cap) st_step[it] = step;
  }
}

// Q += step*sx; grad_y += 2*step*psisx + step*hb; partial Q.(grad_y+U)
// (:424-434), then the first minimum of the new grad_y row: the next
// iteration's conditional gradient (src/qp.cpp:5-15)
template <class SEG>
__global__ void __launch_bounds__(kBlock) k_qp_update_rg(const float* __restrict__ U, float* __restrict__ Q, float* __restrict__ gy,
                               const float* __restrict__ psx, int* __restrict__ lab, const float* __restrict__ hb,
                               int hb_npix, SEG sg, int L, const double* __restrict__ sc, double* __restrict__ parts,
                               const int* __restrict__ fl) {
  const SegPos qs = sg.pos();
  if (fl[qs.img * kF_COUNT + kF_DONE]) return;
  __shared__ float seg[kBlock];
  const double step = sc[(size_t)qs.img * kS_COUNT + kS_STEP];
  const float st = (float)step, st2 = (float)(2 * step);
  double e = 0.0;
  const long ntiles = rg_tiles(qs.n, L);
  for (long tile = qs.lb; tile < ntiles; tile += qs.nb) {
    RG r = rg_at(L, qs.n, tile);
    r.row += qs.p0;
    float* row = seg + (threadIdx.x & ~63) + r.g * L;
    if (r.valid) {
      const size_t idx = (size_t)r.row * L + r.k;
      const float q0 = Q[idx];
      const float sx = (lab[r.row] == r.k ? 1.f : 0.f) - q0;
      const float q = q0 + st * sx;
      const float h = (hb != nullptr && r.row < hb_npix) ? hb[idx] : 0.f;
      const float g = gy[idx] + (st2 * psx[idx] + st * h);
      Q[idx] = q;
      gy[idx] = g;
      e += (double)q * (double)(g + U[idx]);
      row[r.k] = g;
    }
    wave_lds_sync();  // a row group stays within its wave
    if (r.valid && r.k == 0) {
      int m = 0;
      float v = row[0];
      for (int l = 1; l < L; ++l)
        if (row[l] < v) {
          v = row[l];
          m = l;
        }
      lab[r.row] = m;
    }
    wave_lds_sync();
  }
  double v[1] = {e};
  block_sum<1>(v);
  if (threadIdx.x == 0) parts[blockIdx.x] = v[0];
}

// new energy and the reference's stopping rule old - energy > DIFF_ENG=100
// (:390, :434-435)
__global__ void k_qp_check(double* sc_all, int* fl_all, int ref_stop, int has_ho, double* st_energy, int cap,
                           int nimg) {
  for (int img = threadIdx.x; img < nimg; img += blockDim.x) {
    double* sc = sc_all + (size_t)img * kS_COUNT;
    int* fl = fl_all + (size_t)img * kF_COUNT;
    if (fl[kF_DONE]) continue;
    double e = sc[kS_EPART];
    if (has_ho) e += sc[kS_HO + kH_E];
    sc[kS_ECUR] = e;
    const int it = fl[kF_ITERS];
    if (it < cap + 1) st_energy[it] = e;
    if (ref_stop && !((sc[kS_EPREV] - e) > 100.0)) fl[kF_DONE] = 1;
  }
}

// currentMap (src/densecrf.cpp:929-938): first maximum of each column
__global__ void __launch_bounds__(kBlock) k_argmax(const float* __restrict__ Q, int* __restrict__ lab, int N, int L) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const float* r = Q + (size_t)i * L;
  int m = 0;
  float v = r[0];
  for (int l = 1; l < L; ++l) {
    const float x = r[l];
    if (x > v) {
      v = x;
      m = l;
    }
  }
  lab[i] = m;
}

// r[i] += -out(l_i, i) of one no-normalisation potential (:915-925)
__global__ void __launch_bounds__(kBlock) k_pairwise_true_accum(const float* __restrict__ o, const int* __restrict__ lab,
                                      float* __restrict__ r, int N, int L) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const int l = lab[i];
  const float v = (l >= 0 && l < L) ? -o[(size_t)i * L + l] : 0.f;
  r[i] = r[i] + v;
}

// unaryEnergy (:887-899): U(l_i, i), 0 for a label out of range
__global__ void __launch_bounds__(kBlock) k_unary_energy(const float* __restrict__ U, const int* __restrict__ lab,
                                                         int N, int L, float* __restrict__ r) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const int l = lab[i];
  r[i] = (l >= 0 && l < L) ? U[(size_t)i * L + l] : 0.f;
}

// Σ_i (unary(l_i, i) + pairwise_i) in FP64 (:873-885)
template <class SEG>
__global__ void __launch_bounds__(kBlock) k_assign_energy(const float* __restrict__ U, const int* __restrict__ lab,
                                const float* __restrict__ r, SEG seg, int L, double* __restrict__ parts) {
  const SegPos q = seg.pos();
  double v[1] = {0.0};
  for (long i = q.p0 + q.lb * blockDim.x + threadIdx.x; i < q.p0 + q.n; i += q.nb * blockDim.x) {
    const int l = lab[i];
    const float u = (l >= 0 && l < L) ? U[(size_t)i * L + l] : 0.f;
    v[0] += (double)(u + r[i]);
  }
  block_sum<1>(v);
  if (threadIdx.x == 0) parts[blockIdx.x] = v[0];
}

}  // namespace
}  // namespace dcrf
