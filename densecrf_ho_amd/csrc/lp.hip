// Proximal LP (DenseCRF::lp_inference_prox_super_pixels, src/densecrf.cpp:526-799)
// and the LP-side energies on gfx950.
//
// Per inner Frank-Wolfe step:
//   k_lp_Y        alpha,u <- previous step's update (fused), Y = lambda*C(alpha+u-U) + cur_Q
//   k_lp_sweeps   qp-gamma: B sweeps per pixel in registers, per-sweep partial
//                 sums of qp_values; k_lp_sweep_decide finds the stopping sweep
//                 and the replay pass writes gamma after exactly that many sweeps
//   k_lp_betaQ    beta, prox Q and its global min/max (rescale)
//   ordered filters -> sa ; computeUCondGrad -> su (superpixels)
//   k_lp_gap      dual gap and |tmp|^2 in FP64 -> delta
// Per outer step: k_lp_project (feasible_Q + renormalize + argmax) and the
// integer energy of the labelling.
#include <algorithm>
#include <cmath>
#include <limits>

#include "crf.hpp"
#include "crf_kernels.hpp"
#include "tuner.hpp"

namespace dcrf {
namespace {

// LP scalar slots (doubles) and flags (ints) in LPState::sc / fl
// kL_H0..kL_H0+2: the last three sweep totals of the inner step (device plan)
constexpr int kL_GAP = 0, kL_TT = 1, kL_DELTA = 2, kL_QPV = 3, kL_EINT = 4, kL_EHO = 5, kL_EPAIR = 6, kL_TMP = 8,
              kL_H0 = 12, kL_NSC = 16;
// kLF_TOTAL: sweeps done so far in the current inner step.  Batches planned
// on the device: kLF_NEXT the next batch's length, kLF_CUR the length of the
// batch being decided, kLF_NB the batches that ran, kLF_NH the totals seen.
constexpr int kLF_SWEEP_DONE = 0, kLF_SWEEPS = 1, kLF_ERR = 2, kLF_TOTAL = 3, kLF_NEXT = 4, kLF_CUR = 5,
              kLF_NB = 6, kLF_NH = 7, kLF_NFL = 8;
constexpr int kSweepBatch = 16;

// sum of the first L entries of a register array in Eigen's order for an
// aligned vector (first aligned element 0): blocks of 8 into two 4-wide
// accumulators, one more block of 4 if present, (p0+p2)+(p1+p3), then the tail
template <int MAXL, class F>
__device__ inline float eigen_sum_fn(int L, F v);
template <int MAXL>
__device__ inline float eigen_sum_reg(const float (&a)[MAXL], int L) {
  return eigen_sum_fn<MAXL>(L, [&](int i) { return a[i]; });
}
// the same sum over v(0..L-1), the values produced on the fly
template <int MAXL, class F>
__device__ inline float eigen_sum_fn(int L, F v) {
  const int as2 = (L / 8) * 8, as1 = (L / 4) * 4;
  if (as1 == 0) {
    float r = v(0);
#pragma unroll
    for (int i = 1; i < MAXL; ++i)
      if (i < L) r = r + v(i);
    return r;
  }
  float p0 = v(0), p1 = v(1), p2 = v(2), p3 = v(3);
  if (as1 > 4) {
    float q0 = v(4), q1 = v(5), q2 = v(6), q3 = v(7);
#pragma unroll
    for (int b = 8; b + 7 < MAXL; b += 8)
      if (b < as2) {
        p0 = p0 + v(b);
        p1 = p1 + v(b + 1);
        p2 = p2 + v(b + 2);
        p3 = p3 + v(b + 3);
        q0 = q0 + v(b + 4);
        q1 = q1 + v(b + 5);
        q2 = q2 + v(b + 6);
        q3 = q3 + v(b + 7);
      }
    p0 = p0 + q0;
    p1 = p1 + q1;
    p2 = p2 + q2;
    p3 = p3 + q3;
    if (as1 > as2) {
#pragma unroll
      for (int b = 8; b + 3 < MAXL; b += 8)
        if (b == as2) {
          p0 = p0 + v(b);
          p1 = p1 + v(b + 1);
          p2 = p2 + v(b + 2);
          p3 = p3 + v(b + 3);
        }
    }
  }
  float r = (p0 + p2) + (p1 + p3);
#pragma unroll
  for (int i = 4; i < MAXL; ++i)
    if (i >= as1 && i < L) r = r + v(i);
  return r;
}
template <int MAXL>
__device__ inline float eigen_dot_reg(const float (&a)[MAXL], const float (&b)[MAXL], int L) {
  float p[MAXL];
#pragma unroll
  for (int i = 0; i < MAXL; ++i) p[i] = a[i] * b[i];
  return eigen_sum_reg<MAXL>(p, L);
}

struct LPConst {
  float lambda, c_m1_over_M, c_2l_over_M, c_1_over_M, c_1m2_over_M, qp_delta;
};

// alpha/u update of the previous step (:737-738, fused here), then
// Y = lambda * (tmp - sum(tmp)/M) + cur_Q with tmp = alpha + u - U (:627-635).
// Row groups; sum(tmp) is the sum of an aligned copy (qp_gamma_multiplyC).
template <class SEG>
__global__ void __launch_bounds__(kBlock) k_lp_Y(SEG sg, int L, LPConst c, float* __restrict__ alpha, float* __restrict__ u,
                       const float* __restrict__ sa, const float* __restrict__ su, const double* __restrict__ sc,
                       int apply_update, const float* __restrict__ U, const float* __restrict__ curQ,
                       float* __restrict__ Y) {
  __shared__ float seg[kBlock / 64][256];
  const SegPos qs = sg.pos();
  const float dl = apply_update ? (float)sc[(size_t)qs.img * kL_NSC + kL_DELTA] : 0.f;
  const long ntiles = rq_tiles(qs.n, L);
  for (long tile = qs.lb; tile < ntiles; tile += qs.nb) {
    RQ r = rq_at(L, qs.n, tile);
    r.row += qs.p0;
    float* row = seg[threadIdx.x >> 6] + r.g * L;
    const int k0 = 4 * r.q, nk = min(4, L - k0);
    const size_t idx = (size_t)r.row * L + k0;
    float t[4] = {0.f, 0.f, 0.f, 0.f}, cq[4];
    if (r.valid) {
      float a[4], uu[4] = {0.f, 0.f, 0.f, 0.f}, uv[4];
      ld_q4(curQ + idx, nk, cq);  // loaded with the other inputs, used after the row sum
      ld_q4(alpha + idx, nk, a);
      if (u) ld_q4(u + idx, nk, uu);
      ld_q4(U + idx, nk, uv);
      if (apply_update) {
        float sv[4];
        ld_q4(sa + idx, nk, sv);
#pragma unroll
        for (int i = 0; i < 4; ++i) a[i] = a[i] + dl * (sv[i] - a[i]);
        st_q4(alpha + idx, nk, a);
        if (u) {
          float sw[4];
          ld_q4(su + idx, nk, sw);
#pragma unroll
          for (int i = 0; i < 4; ++i) uu[i] = uu[i] + dl * (sw[i] - uu[i]);
          st_q4(u + idx, nk, uu);
        }
      }
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        t[i] = (a[i] + uu[i]) - uv[i];
        if (i < nk) row[k0 + i] = t[i];
      }
    }
    wave_lds_sync();  // the row sums stay within the wave
    if (r.valid) {
      const float S = eigen_sum(L, 0, [&](int i) { return row[i]; });
      float y[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) y[i] = ((c.c_m1_over_M * S + t[i]) * c.lambda) + cq[i];
      st_q4(Y + idx, nk, y);
    }
    wave_lds_sync();
  }
}

// one qp-gamma sweep of a pixel (qp_gamma_step :495-514, qp_gamma_multiplyC
// :517-523, qp_values :667); returns the pixel's qp value
// The row of Y is held as its positive and negative parts nh = max(y, 0),
// ph = max(-y, 0) (the terms of the update); y = nh - ph where the qp value
// needs it (exact: one of the two is zero).
// sum: in, the Eigen-order sum of g; out, that of the updated g (the next
// sweep's sum: the same operations on the same values)
template <int MAXL>
__device__ inline __attribute__((always_inline)) float qp_sweep(float (&g)[MAXL], const float (&ph)[MAXL],
                                                                const float (&nh)[MAXL], int L, const LPConst& c,
                                                                float& sum_io) {
  const float sum = sum_io;
#pragma unroll
  for (int l = 0; l < MAXL; ++l)
    if (l < L) {
      const float t = (sum - g[l]) * c.c_2l_over_M;
      const float stp = (t + ph[l]) + c.qp_delta;
      float t2 = (c.c_1_over_M * sum + c.c_1m2_over_M * g[l]) * c.lambda;
      t2 = (t2 + nh[l]) + c.qp_delta;
      g[l] = g[l] * (stp / t2);
    }
  const float s2 = eigen_sum_reg<MAXL>(g, L);
  sum_io = s2;
  const float cs2 = c.c_m1_over_M * s2;
  const float d1 = eigen_sum_fn<MAXL>(L, [&](int l) { return g[l] * ((cs2 + g[l]) * c.lambda); });
  const float d2 = eigen_sum_fn<MAXL>(L, [&](int l) { return g[l] * (nh[l] - ph[l]); });
  return (float)(0.5 * (double)d1 + (double)d2);
}

// qp-gamma sweeps (:650-675), one thread per pixel holding its row in
// registers; rows move through LDS so the block's kBlock x L tiles load and
// store coalesced.
//   pass 1 (replay == 0): `nsw` sweeps from gin, gamma after nsw sweeps to
//     gout, per-sweep block partial sums of qp_values to parts
//   replay (replay == 1): if the loop stopped after s < nsw sweeps
//     (fl[kLF_SWEEPS]), redo s sweeps from gin into gout; otherwise no-op
//   An image whose loop ended in an earlier batch of this inner step only
//   carries gamma over (gin -> gout) in pass 1.  Block b works on pixels
//   [lb * kBlock, lb * kBlock + kBlock) of its image.
//   nsw < 0: the lengths come from the device plan (pass 1 fl[kLF_NEXT], the
//   replay fl[kLF_CUR]).  Without `carry`, an image whose loop has ended does
//   nothing in pass 1 (one image: the host takes gamma from the buffer of the
//   last batch that ran).
template <int MAXL, int LT = 0, class SEG>  // LT > 0: the label count is known at compile time
__global__ void __launch_bounds__(kBlock, 6) k_lp_sweeps(SEG sg, int Lr, LPConst c, const float* __restrict__ gin, float* __restrict__ gout,
                            const float* __restrict__ Y, int nsw, int replay, int carry,
                            double* __restrict__ parts, const int* __restrict__ fl_all) {
  __shared__ float tile[kBlock * MAXL];
  const int L = LT > 0 ? LT : Lr;
  const SegPos qs = sg.pos();
  const int* fl = fl_all + (size_t)qs.img * kLF_NFL;
  const bool dev = nsw < 0;
  if (dev) nsw = replay ? fl[kLF_CUR] : fl[kLF_NEXT];
  if (replay && fl[kLF_SWEEPS] >= nsw) return;
  const long p0 = qs.p0 + qs.lb * kBlock;
  const int np = (int)min((long)kBlock, qs.n - qs.lb * kBlock);
  const size_t base = (size_t)p0 * L;
  const int t = threadIdx.x;
  if (!replay && fl[kLF_SWEEP_DONE]) {
    if (carry)
      for (int i = t; i < np * L; i += kBlock) gout[base + i] = gin[base + i];
    return;
  }
  const int n = replay ? fl[kLF_SWEEPS] : nsw;
  const bool act = t < np;
  __shared__ double red[kSweepBatch][kBlock / 64];  // per sweep and wave: wave sum of the qp values
  float g[MAXL], ph[MAXL], nh[MAXL];
  if constexpr (LT > 0 && LT == MAXL) {
    // the label count fixed at compile time: each thread loads its own two
    // rows with 16-byte loads, all issued at once (no LDS round trips)
    const size_t r0 = base + (size_t)t * LT;
    float y[MAXL];
#pragma unroll
    for (int l = 0; l < MAXL; ++l) g[l] = y[l] = 0.f;
    if (act) {
#pragma unroll
      for (int l = 0; l + 3 < LT; l += 4) {
        const Float4u a = *reinterpret_cast<const Float4u*>(gin + r0 + l);
        const Float4u b = *reinterpret_cast<const Float4u*>(Y + r0 + l);
        g[l] = a.x, g[l + 1] = a.y, g[l + 2] = a.z, g[l + 3] = a.w;
        y[l] = b.x, y[l + 1] = b.y, y[l + 2] = b.z, y[l + 3] = b.w;
      }
#pragma unroll
      for (int l = LT & ~3; l < LT; ++l) {
        g[l] = gin[r0 + l];
        y[l] = Y[r0 + l];
      }
    }
#pragma unroll
    for (int l = 0; l < MAXL; ++l) {
      ph[l] = fmaxf(-y[l], 0.f);
      nh[l] = fmaxf(y[l], 0.f);
    }
  } else {
    for (int i = t; i < np * L; i += kBlock) tile[i] = gin[base + i];
    __syncthreads();
#pragma unroll
    for (int l = 0; l < MAXL; ++l) g[l] = (act && l < L) ? tile[t * L + l] : 0.f;
    __syncthreads();
    for (int i = t; i < np * L; i += kBlock) tile[i] = Y[base + i];
    __syncthreads();
#pragma unroll
    for (int l = 0; l < MAXL; ++l) {
      const float y = (act && l < L) ? tile[t * L + l] : 0.f;
      ph[l] = fmaxf(-y, 0.f);
      nh[l] = fmaxf(y, 0.f);
    }
  }
  float gsum = eigen_sum_reg<MAXL>(g, L);
  for (int sw = 0; sw < n; ++sw) {
    const float v = act ? qp_sweep<MAXL>(g, ph, nh, L, c, gsum) : 0.f;
    if (!replay) {
      const double ws = wave_sum((double)v);
      if ((t & 63) == 0) red[sw][t >> 6] = ws;
    }
  }
  __syncthreads();
  if (!replay && t == 0)  // block sum per sweep: the waves' sums in wave order
    for (int sw = 0; sw < n; ++sw) {
      double bs = 0;
      for (int w = 0; w < kBlock / 64; ++w) bs += red[sw][w];
      parts[(size_t)sw * gridDim.x + blockIdx.x] = bs;
    }
  if constexpr (LT > 0 && LT == MAXL) {
    if (act) {
      const size_t r0 = base + (size_t)t * LT;
#pragma unroll
      for (int l = 0; l + 3 < LT; l += 4)
        *reinterpret_cast<Float4u*>(gout + r0 + l) = Float4u{g[l], g[l + 1], g[l + 2], g[l + 3]};
#pragma unroll
      for (int l = LT & ~3; l < LT; ++l) gout[r0 + l] = g[l];
    }
  } else {
#pragma unroll
    for (int l = 0; l < MAXL; ++l)
      if (act && l < L) tile[t * L + l] = g[l];
    __syncthreads();
    for (int i = t; i < np * L; i += kBlock) gout[base + i] = tile[i];
  }
}

// per-sweep totals of the batch's block partials: block (image, sweep) sums
// the partials of the image's blocks [first[i], first[i+1]) of the sweep
// launch (all nblocks when first is null) into tot[image * kSweepBatch +
// sweep]: each thread's blocks in order, the wave sums, the waves in order.
// Images whose loop already ended in this inner step are skipped.
// (dev: the batch's length is fl[kLF_NEXT], the grid covers kSweepBatch)
__global__ void __launch_bounds__(1024) k_lp_sweep_sums(const double* __restrict__ parts, int nblocks,
                                                        const int* __restrict__ first, double* __restrict__ tot,
                                                        const int* __restrict__ fl_all, int dev) {
  __shared__ double red[1024 / 64];
  const int img = blockIdx.x, sw = blockIdx.y;
  if (fl_all[(size_t)img * kLF_NFL + kLF_SWEEP_DONE]) return;
  if (dev && sw >= fl_all[(size_t)img * kLF_NFL + kLF_NEXT]) return;
  const int b0 = first ? first[img] : 0, b1 = first ? first[img + 1] : nblocks;
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  double acc = 0.0;
  for (int b = b0 + threadIdx.x; b < b1; b += blockDim.x) acc += parts[(size_t)sw * nblocks + b];
  const double w = wave_sum(acc);
  if (lane == 0) red[wid] = w;
  __syncthreads();
  if (threadIdx.x == 0) {
    double v = 0;
    for (int i = 0; i < (int)blockDim.x / 64; ++i) v += red[i];
    tot[(size_t)img * kSweepBatch + sw] = v;
  }
}

// stopping sweep of the batch (:670-675): the first sweep whose float total
// differs from the previous one by less than qp_tol; fl[kLF_SWEEPS] = sweeps
// to replay, fl[kLF_SWEEP_DONE] = 1 if the loop ends in this batch.
// sc[kL_QPV] carries the previous total across batches.  One thread per
// image; an image whose loop already ended in this inner step is left as it
// is (its replay is a no-op).
// Length of the next batch of qp-gamma sweeps of an inner step.  The loop
// stops at the first sweep whose total moves by less than qp_tol
// (src/densecrf.cpp:670-675), and the totals' steps shrink about
// geometrically: after a first batch of kSweepFirst sweeps the remaining
// count is predicted from the ratio of the last two steps.  A batch that runs
// past the stop costs a replay and one that ends short another launch, each
// about four sweeps' time (1080p); measured on the 1080p LP this plan
// computes about 530 sweeps for 503 used, where following the previous inner
// step's count computed about 960.  It runs on the device, in
// k_lp_sweep_decide, so the host can enqueue batches without reading the
// totals.  Only launch shapes depend on it, never a result.
constexpr int kSweepFirst = 5;
// the rule, from the last three totals h[0..2] (oldest first) of the nh seen
__device__ inline int sweep_plan_rule(const double* h, int nh, float qp_tol) {
  if (nh < 3) return kSweepFirst;
  const double d1 = fabs(h[2] - h[1]), d0 = fabs(h[1] - h[0]);
  if (d1 < (double)qp_tol) return 1;
  double r = d0 > 0 ? d1 / d0 : 0.5;
  if (r >= 0.999) r = 0.95;
  if (r <= 0.0) r = 0.5;
  const double m = ceil(log((double)qp_tol / d1) / log(r));
  return (int)fmax(1.0, fmin((double)kSweepBatch, m));
}
// nsw < 0: the batch's length is fl[kLF_NEXT] (device plan); the totals then
// feed the plan of the next batch.  host_status (pinned host memory, may be
// null) receives {every image's loop ended, image 0's sweeps so far, image
// 0's batches run, then each image's sweeps so far}.  One block.
__global__ void k_lp_sweep_decide(const double* __restrict__ tot, int nimg, int nsw, float qp_tol, int max_total,
                                  double* __restrict__ sc_all, int* __restrict__ fl_all,
                                  int* __restrict__ host_status) {
  const bool dev = nsw < 0;
  for (int img = threadIdx.x; img < nimg; img += blockDim.x) {
    double* sc = sc_all + (size_t)img * kL_NSC;
    int* fl = fl_all + (size_t)img * kLF_NFL;
    if (dev) {
      nsw = fl[kLF_NEXT];
      fl[kLF_CUR] = nsw;
    }
    if (fl[kLF_SWEEP_DONE]) {
      fl[kLF_SWEEPS] = nsw;
      continue;
    }
    const double* tt = tot + (size_t)img * kSweepBatch;
    const int done_before = fl[kLF_TOTAL];
    float prev = done_before == 0 ? 3.402823466e38f : (float)sc[kL_QPV];
    int stop = nsw, finished = 0;
    for (int sw = 0; sw < nsw; ++sw) {
      const float cur = (float)tt[sw];
      const int total = done_before + sw + 1;
      if (fabsf(prev - cur) < qp_tol || total >= max_total) {
        stop = sw + 1;
        finished = 1;
        break;
      }
      prev = cur;
    }
    sc[kL_QPV] = (double)prev;
    fl[kLF_SWEEPS] = stop;
    fl[kLF_SWEEP_DONE] = finished;
    fl[kLF_TOTAL] = done_before + stop;
    fl[kLF_NB] += 1;
    if (dev && !finished) {  // every total of the batch was used: the plan's history
      double* h = sc + kL_H0;
      for (int sw = 0; sw < nsw; ++sw) {
        h[0] = h[1];
        h[1] = h[2];
        h[2] = tt[sw];
      }
      fl[kLF_NH] += nsw;
      fl[kLF_NEXT] = max(1, min(sweep_plan_rule(h, fl[kLF_NH], qp_tol), max_total - fl[kLF_TOTAL]));
    }
  }
  if (host_status) {
    __syncthreads();
    for (int img = threadIdx.x; img < nimg; img += blockDim.x)
      host_status[3 + img] = fl_all[(size_t)img * kLF_NFL + kLF_TOTAL];
    if (threadIdx.x == 0) {
      int all = 1;
      for (int img = 0; img < nimg; ++img) all = all && fl_all[(size_t)img * kLF_NFL + kLF_SWEEP_DONE];
      host_status[0] = all;
      host_status[1] = fl_all[kLF_TOTAL];
      host_status[2] = fl_all[kLF_NB];
    }
  }
}


// start of an inner step's sweep loop, one thread per image
// (first: the first batch's length under the device plan)
__global__ void k_lp_sweep_reset(int* __restrict__ fl_all, int nimg, int first) {
  for (int i = threadIdx.x; i < nimg; i += blockDim.x) {
    int* fl = fl_all + (size_t)i * kLF_NFL;
    fl[kLF_SWEEP_DONE] = 0;
    fl[kLF_TOTAL] = 0;
    fl[kLF_NEXT] = first;
    fl[kLF_NB] = 0;
    fl[kLF_NH] = 0;
  }
}

// NaN-propagating min / max (Eigen's minCoeff/maxCoeff see NaN as unordered;
// any NaN makes the rescale fail here)
__device__ inline float nmin(float a, float b) { return (a != a || b != b) ? NAN : fminf(a, b); }
__device__ inline float nmax(float a, float b) { return (a != a || b != b) ? NAN : fmaxf(a, b); }

// beta = -colsum(alpha+u+gamma-U)/M, Q = lambda*(alpha+beta+u+gamma-U) + cur_Q
// (:684-697), row groups; the column sum follows Eigen's order for the
// column's alignment.  Block min/max of Q for rescale (:699).
template <class SEG>
__global__ void __launch_bounds__(kBlock) k_lp_betaQ(SEG sg, int L, float lambda, const float* __restrict__ alpha, const float* __restrict__ u,
                           const float* __restrict__ gamma, const float* __restrict__ U,
                           const float* __restrict__ curQ, float* __restrict__ Q, float* __restrict__ mmparts) {
  __shared__ float seg[kBlock / 64][256];
  __shared__ float smn[kBlock], smx[kBlock];
  float mn = INFINITY, mx = -INFINITY;
  const SegPos qs = sg.pos();
  const long ntiles = rq_tiles(qs.n, L);
  for (long tile = qs.lb; tile < ntiles; tile += qs.nb) {
    RQ r = rq_at(L, qs.n, tile);
    r.row += qs.p0;
    float* row = seg[threadIdx.x >> 6] + r.g * L;
    const int k0 = 4 * r.q, nk = min(4, L - k0);
    const size_t idx = (size_t)r.row * L + k0;
    float a[4], uu[4] = {0.f, 0.f, 0.f, 0.f}, gm[4], uv[4], cq[4];
    if (r.valid) {
      ld_q4(curQ + idx, nk, cq);  // loaded with the other inputs, used after the row sum
      ld_q4(alpha + idx, nk, a);
      if (u) ld_q4(u + idx, nk, uu);
      ld_q4(gamma + idx, nk, gm);
      ld_q4(U + idx, nk, uv);
#pragma unroll
      for (int i = 0; i < 4; ++i)
        if (i < nk) row[k0 + i] = ((a[i] + uu[i]) + gm[i]) - uv[i];
    }
    wave_lds_sync();  // the row sums stay within the wave
    if (r.valid) {
      // alignment of the column within its own image's matrix
      float b = -eigen_sum(L, col_astart(r.row - qs.p0, L), [&](int i) { return row[i]; });
      b = b / (float)L;
      float qv[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        qv[i] = lambda * ((((a[i] + b) + uu[i]) + gm[i]) - uv[i]) + cq[i];
        if (i < nk) {
          mn = nmin(mn, qv[i]);
          mx = nmax(mx, qv[i]);
        }
      }
      st_q4(Q + idx, nk, qv);
    }
    wave_lds_sync();
  }
  smn[threadIdx.x] = mn;
  smx[threadIdx.x] = mx;
  __syncthreads();
  for (int o = kBlock / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) {
      smn[threadIdx.x] = nmin(smn[threadIdx.x], smn[threadIdx.x + o]);
      smx[threadIdx.x] = nmax(smx[threadIdx.x], smx[threadIdx.x + o]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    mmparts[2 * blockIdx.x] = smn[0];
    mmparts[2 * blockIdx.x + 1] = smx[0];
  }
}

// rescale (src/eigen_utils.cpp:120-140): mm = {min, max(Q - min)}; error when
// max < 0 (only possible with NaN here).  One block.
// (one block per image over its blocks [first[i], first[i+1]), or all nb
// blocks when first is null; mm and fl of image i at mm + 2i, fl + i*kLF_NFL)
__global__ void k_lp_minmax_final(const float* __restrict__ parts, int nb, const int* __restrict__ first,
                                  float* __restrict__ mm_all, int* __restrict__ fl_all) {
  __shared__ float smn[1024], smx[1024];
  float* mm = mm_all + 2 * (size_t)blockIdx.x;
  int* fl = fl_all + (size_t)blockIdx.x * kLF_NFL;
  const int b0 = first ? first[blockIdx.x] : 0, b1 = first ? first[blockIdx.x + 1] : nb;
  float mn = INFINITY, mx = -INFINITY;
  for (int b = b0 + threadIdx.x; b < b1; b += blockDim.x) {
    mn = nmin(mn, parts[2 * b]);
    mx = nmax(mx, parts[2 * b + 1]);
  }
  smn[threadIdx.x] = mn;
  smx[threadIdx.x] = mx;
  __syncthreads();
  for (int o = blockDim.x / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) {
      smn[threadIdx.x] = nmin(smn[threadIdx.x], smn[threadIdx.x + o]);
      smx[threadIdx.x] = nmax(smx[threadIdx.x], smx[threadIdx.x + o]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const float maxval = smx[0] - smn[0];
    mm[0] = smn[0];
    mm[1] = maxval;
    if (smn[0] != smn[0] || smx[0] != smx[0] || maxval < 0) fl[kLF_ERR] = 1;
  }
}

// tmp = alpha + u - (sa + su); gap = tmp.Q; tt = tmp.tmp (:725-729)
template <class SEG>
__global__ void __launch_bounds__(kBlock) k_lp_gap(SEG sg, int L, const float* __restrict__ alpha, const float* __restrict__ u,
                         const float* __restrict__ sa, const float* __restrict__ su, const float* __restrict__ Q,
                         double* __restrict__ parts) {
  const SegPos qs = sg.pos();
  double v[2] = {0.0, 0.0};
  const size_t e = (size_t)(qs.p0 + qs.n) * L, step = (size_t)qs.nb * blockDim.x;
  size_t i = (size_t)qs.p0 * L + (size_t)qs.lb * blockDim.x + threadIdx.x;
  // four elements' loads in flight together, then their terms in order
  for (; i + 3 * step < e; i += 4 * step) {
    float t[4], q[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const size_t ij = i + j * step;
      const float uu = u ? u[ij] : 0.f, ss = su ? su[ij] : 0.f;
      t[j] = (alpha[ij] + uu) - (sa[ij] + ss);
      q[j] = Q[ij];
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      v[0] += (double)t[j] * (double)q[j];
      v[1] += (double)t[j] * (double)t[j];
    }
  }
  for (; i < e; i += step) {
    const float uu = u ? u[i] : 0.f, ss = su ? su[i] : 0.f;
    const float t = (alpha[i] + uu) - (sa[i] + ss);
    v[0] += (double)t * (double)Q[i];
    v[1] += (double)t * (double)t;
  }
  block_sum<2>(v);
  if (threadIdx.x == 0) {
    parts[2 * blockIdx.x] = v[0];
    parts[2 * blockIdx.x + 1] = v[1];
  }
}
// delta = clamp((float)(gap / (lambda * tt)), 0, 1) (:729-730); stats
// (one thread per image; the record st_gap / st_delta is kept for image 0)
__global__ void k_lp_delta(double* __restrict__ sc_all, float lambda, double* __restrict__ st_gap,
                           double* __restrict__ st_delta, int idx, int cap, int nimg) {
  for (int img = threadIdx.x; img < nimg; img += blockDim.x) {
    double* sc = sc_all + (size_t)img * kL_NSC;
    const double gap = sc[kL_GAP], tt = sc[kL_TT];
    float d = (float)(gap / ((double)lambda * tt));
    d = (d < 0.0f) ? 0.0f : d;  // std::max(d, 0.0f)
    d = (1.0f < d) ? 1.0f : d;  // std::min(d, 1.0f)
    sc[kL_DELTA] = d;
    if (img == 0 && idx < cap) {
      st_gap[idx] = gap;
      st_delta[idx] = d;
    }
  }
}

// feasible_Q (:470-492), then renormalize (:444-455), then currentMap
// (:929-938); row groups.  The descending sort of sortCols becomes a rank per
// value (ties broken by index), the threshold loop runs on the sorted row in
// LDS, and the column sums follow Eigen's order for the column's alignment.
template <class SEG>
__global__ void __launch_bounds__(kBlock) k_lp_project(SEG sg, int L, const float* __restrict__ Qin, float* __restrict__ Qout,
                             int* __restrict__ lab, int do_feasible) {
  __shared__ float sq[kBlock], ss[kBlock];
  const SegPos qs = sg.pos();
  const long ntiles = rg_tiles(qs.n, L);
  for (long tile = qs.lb; tile < ntiles; tile += qs.nb) {
    RG r = rg_at(L, qs.n, tile);
    const long lcol = r.row;  // column within its own image (sets the Eigen sum alignment)
    r.row += qs.p0;
    const int wb = (threadIdx.x & ~63) + r.g * L;
    float* rowq = sq + wb;
    float* rows = ss + wb;
    const size_t idx = (size_t)r.row * L + r.k;
    const float q = r.valid ? Qin[idx] : 0.f;
    if (r.valid) rowq[r.k] = q;
    __syncthreads();
    float v = q;
    if (do_feasible) {
      float cs = 0.f;
      if (r.valid) {
        cs = eigen_sum(L, col_astart(lcol, L), [&](int i) { return rowq[i]; });
        int rank = 0;
        for (int j = 0; j < L; ++j) {
          const float qj = rowq[j];
          rank += (qj > q) || (qj == q && j < r.k);
        }
        rows[rank] = q;
      }
      __syncthreads();
      if (r.valid) {
        double sum = (double)(float)(cs - 1.f);
        int K = -1;
        for (int k = L; k > 0; --k) {
          const double uk = (double)rows[k - 1];
          if (sum / k < uk) {
            K = k;
            break;
          }
          sum -= uk;
        }
        double t = (double)q - sum / K;
        t = (t < 0.0) ? 0.0 : t;  // std::max(t, 0.0)
        t = (1.0 < t) ? 1.0 : t;  // std::min(t, 1.0)
        v = (float)t;
      }
      __syncthreads();
      if (r.valid) rowq[r.k] = v;
      __syncthreads();
    }
    if (r.valid) {
      const double sum = (double)eigen_sum(L, col_astart(lcol, L), [&](int i) { return rowq[i]; });
      const float out = (sum == 0 || sum != sum) ? (float)(1.0 / L) : v / (float)sum;
      Qout[idx] = out;
      rows[r.k] = out;
    }
    __syncthreads();
    if (r.valid && r.k == 0) {
      int m = 0;
      float best = rows[0];
      for (int l = 1; l < L; ++l)
        if (rows[l] > best) {
          best = rows[l];
          m = l;
        }
      lab[r.row] = m;
    }
    __syncthreads();
  }
}

// ---- the same per pixel, one thread per pixel with its row in registers ----
// (LT labels known at compile time).  The row moves through LDS so the
// block's loads and stores coalesce.

// Eigen's sum (the eigen_sum order) of a register row whose first 16-byte
// aligned element is A
template <int LT, int A>
__device__ inline float eigen_sum_a(const float (&x)[LT]) {
  constexpr int a = A > LT ? LT : A;
  constexpr int as2 = ((LT - a) / 8) * 8, as1 = ((LT - a) / 4) * 4;
  constexpr int ae2 = a + as2, ae = a + as1;
  if constexpr (as1 > 0) {
    float p0 = x[a], p1 = x[a + 1], p2 = x[a + 2], p3 = x[a + 3];
    if constexpr (as1 > 4) {
      float q0 = x[a + 4], q1 = x[a + 5], q2 = x[a + 6], q3 = x[a + 7];
#pragma unroll
      for (int i = a + 8; i < ae2; i += 8) {
        p0 = p0 + x[i];
        p1 = p1 + x[i + 1];
        p2 = p2 + x[i + 2];
        p3 = p3 + x[i + 3];
        q0 = q0 + x[i + 4];
        q1 = q1 + x[i + 5];
        q2 = q2 + x[i + 6];
        q3 = q3 + x[i + 7];
      }
      p0 = p0 + q0;
      p1 = p1 + q1;
      p2 = p2 + q2;
      p3 = p3 + q3;
      if constexpr (ae > ae2) {
        p0 = p0 + x[ae2];
        p1 = p1 + x[ae2 + 1];
        p2 = p2 + x[ae2 + 2];
        p3 = p3 + x[ae2 + 3];
      }
    }
    float res = (p0 + p2) + (p1 + p3);
#pragma unroll
    for (int i = 0; i < a; ++i) res = res + x[i];
#pragma unroll
    for (int i = ae; i < LT; ++i) res = res + x[i];
    return res;
  } else {
    float res = x[0];
#pragma unroll
    for (int i = 1; i < LT; ++i) res = res + x[i];
    return res;
  }
}
template <int LT>
__device__ inline float eigen_sum_col(const float (&x)[LT], int a) {
  switch (a) {
    case 0:
      return eigen_sum_a<LT, 0>(x);
    case 1:
      return eigen_sum_a<LT, 1>(x);
    case 2:
      return eigen_sum_a<LT, 2>(x);
    default:
      return eigen_sum_a<LT, 3>(x);
  }
}
// descending sort of the first LT of a row padded (in thought) to N = 2^m
// entries of -inf: Batcher's odd-even merge network, comparators that touch
// a padding entry dropped (a padding entry never moves below LT).  Only the
// sorted values are used, so the order among equal values does not matter.
template <int N, int LT>
__device__ inline void sort_desc(float (&u)[LT]) {
#pragma unroll
  for (int p = 1; p < N; p <<= 1)
#pragma unroll
    for (int k = p; k >= 1; k >>= 1)
#pragma unroll
      for (int j = k % p; j + k < N; j += 2 * k)
#pragma unroll
        for (int i = 0; i < k; ++i) {
          const int x = i + j, y = i + j + k;
          if (y < LT && x / (2 * p) == y / (2 * p)) {
            const float hi = fmaxf(u[x], u[y]), lo = fminf(u[x], u[y]);
            u[x] = hi;
            u[y] = lo;
          }
        }
}
template <int LT>
constexpr int pow2_at_least() {
  int n = 1;
  while (n < LT) n <<= 1;
  return n;
}
template <int LT, class SEG>
__global__ void __launch_bounds__(kBlock) k_lp_project_px(SEG sg, const float* Qin, float* Qout,
                                                          int* __restrict__ lab,
                                                          int do_feasible) {
  const SegPos qs = sg.pos();
  const int t = threadIdx.x;
  const long ntiles = (qs.n + kBlock - 1) / kBlock;
  for (long tl = qs.lb; tl < ntiles; tl += qs.nb) {
    const long c0 = tl * kBlock;  // the tile's first column within its image
    const int np = (int)min((long)kBlock, qs.n - c0);
    const bool act = t < np;
    // the thread's own row, 16-byte loads (in place when Qin == Qout: a row
    // is read before it is written, by the same thread)
    const size_t r0 = (size_t)(qs.p0 + c0 + t) * LT;
    float q[LT], v[LT];
#pragma unroll
    for (int l = 0; l < LT; ++l) q[l] = 0.f;
    if (act) {
#pragma unroll
      for (int l = 0; l + 3 < LT; l += 4) {
        const Float4u a4 = *reinterpret_cast<const Float4u*>(Qin + r0 + l);
        q[l] = a4.x, q[l + 1] = a4.y, q[l + 2] = a4.z, q[l + 3] = a4.w;
      }
#pragma unroll
      for (int l = LT & ~3; l < LT; ++l) q[l] = Qin[r0 + l];
    }
    const int a = col_astart(c0 + t, LT);
    if (do_feasible) {
      const float cs = eigen_sum_col<LT>(q, a);
#pragma unroll
      for (int l = 0; l < LT; ++l) v[l] = q[l];
      sort_desc<pow2_at_least<LT>(), LT>(v);
      // largest k with sum / k < u_k, the sum falling from the top (:480-486)
      double sum = (double)(float)(cs - 1.f);
      int K = -1;
      for (int k = LT; k > 0; --k) {
        const double uk = (double)v[k - 1];
        if (sum / k < uk) {
          K = k;
          break;
        }
        sum -= uk;
      }
      const double sk = sum / K;
#pragma unroll
      for (int l = 0; l < LT; ++l) {
        double tt = (double)q[l] - sk;
        tt = (tt < 0.0) ? 0.0 : tt;  // std::max(t, 0.0)
        tt = (1.0 < tt) ? 1.0 : tt;  // std::min(t, 1.0)
        v[l] = (float)tt;
      }
    } else {
#pragma unroll
      for (int l = 0; l < LT; ++l) v[l] = q[l];
    }
    const double sum = (double)eigen_sum_col<LT>(v, a);
    const bool uni = sum == 0 || sum != sum;
#pragma unroll
    for (int l = 0; l < LT; ++l) v[l] = uni ? (float)(1.0 / LT) : v[l] / (float)sum;
    int m = 0;
    float best = v[0];
#pragma unroll
    for (int l = 1; l < LT; ++l)
      if (v[l] > best) {
        best = v[l];
        m = l;
      }
    if (act) {
      lab[qs.p0 + c0 + t] = m;
#pragma unroll
      for (int l = 0; l + 3 < LT; l += 4)
        *reinterpret_cast<Float4u*>(Qout + r0 + l) = Float4u{v[l], v[l + 1], v[l + 2], v[l + 3]};
#pragma unroll
      for (int l = LT & ~3; l < LT; ++l) Qout[r0 + l] = v[l];
    }
  }
}

// max_rounding (:832-840)
__global__ void __launch_bounds__(kBlock) k_onehot(const int* __restrict__ lab, int N, int L, float* __restrict__ Q) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (size_t)N * L) return;
  const int i = (int)(t / L), l = (int)(t - (size_t)i * L);
  Q[t] = lab[i] == l ? 1.f : 0.f;
}

}  // namespace

// dispatch over the register width of per-pixel kernels
#define DCRF_MAXL_DISPATCH(L, BODY)          \
  do {                                       \
    if ((L) <= 8) {                          \
      constexpr int MAXL = 8;                \
      BODY;                                  \
    } else if ((L) <= 16) {                  \
      constexpr int MAXL = 16;               \
      BODY;                                  \
    } else if ((L) <= 24) {                  \
      constexpr int MAXL = 24;               \
      BODY;                                  \
    } else if ((L) <= 32) {                  \
      constexpr int MAXL = 32;               \
      BODY;                                  \
    } else {                                 \
      constexpr int MAXL = 64;               \
      BODY;                                  \
    }                                        \
  } while (0)

struct CRF::LPState {
  DBuf<float> alpha, u, sa, su, gamma, gamma2, Y, Q, curQ, best, tmp;
  DBuf<int> lab, iscratch, lc;
  DBuf<double> parts, sc, swtot;
  DBuf<float> mmparts, mm;
  DBuf<int> fl;
  DBuf<double> st_gap, st_delta;
  int cap = 0;
  HBuf<int> hstatus;  // the sweep decision's status (k_lp_sweep_decide)
};

// labelling energy: unary + true pairwise (+ superpixel term), into sc[dst]
void CRF::int_energy_dev(const int* lab, double* dst, double* scratch, hipStream_t s) {
  Impl& m = *im_;
  k_fill<<<grid_for(N_), kBlock, 0, s>>>(m.rpix.get(), (size_t)N_, 0.f);
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
    k_pairwise_true_accum<<<grid_for(N_), kBlock, 0, s>>>(m.tmp.get(), lab, m.rpix.get(), N_, L_);
  }
  k_assign_energy<<<kRedBlocks, kBlock, 0, s>>>(U_.get(), lab, m.rpix.get(), OneImage{N_}, L_, m.parts.get());
  finalize_sums<1><<<1, kBlock, 0, s>>>(m.parts.get(), kRedBlocks, dst);
  if (R_ > 0) {
    ho_assign_dev(lab, scratch, m.iscratch.get(), s);
    k_add_scalar<<<1, 1, 0, s>>>(dst, scratch);
  }
  check_launch("int_energy");
}

// LP energy (:982-1037) of a device Q into *dst (device double)
void CRF::lp_energy_dev(const float* Q, double* dst, double* scratch, hipStream_t s) {
  Impl& m = *im_;
  const size_t NL = (size_t)N_ * L_;
  launch_dot(s, U_.get(), Q, NL, m.parts.get(), dst);
  minmax_dev(Q, NL, m.mm.get(), m.fl.get(), s);
  for (int k = 0; k < num_pairwise(); ++k) {
    Potential& p = pairwise(k);
    SliceArgs o;
    o.out = m.tmp.get();
    compat_slice(p, o, L_, s,
                 [&](const SliceArgs& oo) { p.lat->ordered(Q, m.mm.get(), L_, oo, &obins_, k == 0); });
    launch_dot(s, Q, m.tmp.get(), NL, m.parts.get(), scratch);
    k_sub_scalar<<<1, 1, 0, s>>>(dst, scratch);
  }
  if (R_ > 0) {
    ho_maxmin_dev(Q, scratch, s);
    k_add_scalar<<<1, 1, 0, s>>>(dst, scratch);
  }
  check_launch("lp_energy");
}

double CRF::lp_energy(const float* Q) {
  Impl& m = *im_;
  k_zero_scalars<<<1, 64, 0, s_>>>(m.sc.get(), 32, m.fl.get(), 8);
  lp_energy_dev(Q, m.sc.get() + kS_TMP, m.sc.get() + kS_TMP + 1, s_);
  double e = 0;
  int err = 0;
  DCRF_HIP(hipMemcpyAsync(&e, m.sc.get() + kS_TMP, sizeof(double), hipMemcpyDeviceToHost, s_));
  DCRF_HIP(hipMemcpyAsync(&err, m.fl.get() + 2, sizeof(int), hipMemcpyDeviceToHost, s_));
  DCRF_HIP(hipStreamSynchronize(s_));
  if (err) throw Error(kMaxvalNegative, "Maxval < 0, skipping");
  return e;
}

double CRF::assignment_energy_ho(const int* lab) {
  Impl& m = *im_;
  int_energy_dev(lab, m.sc.get() + kS_TMP, m.sc.get() + kS_TMP + 1, s_);
  double e = 0;
  DCRF_HIP(hipMemcpyAsync(&e, m.sc.get() + kS_TMP, sizeof(double), hipMemcpyDeviceToHost, s_));
  DCRF_HIP(hipStreamSynchronize(s_));
  return e;
}

LPResult CRF::lp(const float* init, float* out, const LPParams& P, int stats_cap) {
  if (!lp_) lp_ = std::make_unique<LPState>();
  LPState& st = *lp_;
  Impl& m = *im_;
  const size_t NL = (size_t)N_ * L_;
  const int N = N_, L = L_;
  hipStream_t s = s_;
  const bool ho = R_ > 0;
  if (ho) ho_state();
  for (DBuf<float>* b : {&st.alpha, &st.sa, &st.gamma, &st.gamma2, &st.Y, &st.Q, &st.curQ, &st.best, &st.tmp})
    b->resize(NL);
  if (ho) {
    st.u.resize(NL);
    st.su.resize(NL);
  }
  st.lab.resize(N);
  st.iscratch.resize(4);
  st.lc.resize(1);
  const int nbp = (int)grid_for(N);
  st.parts.resize(std::max<size_t>((size_t)kSweepBatch * nbp, (size_t)2 * kRedBlocks));
  st.swtot.resize(kSweepBatch);
  st.sc.resize(kL_NSC);
  st.mmparts.resize(2 * (size_t)std::max<long>(nbp, 16384));  // up to the largest betaQ grid tried
  st.mm.resize(2);
  st.fl.resize(kLF_NFL);
  st.hstatus.resize(4);
  const int cap = std::max(1, std::min(stats_cap, 1 << 16));
  if (cap > st.cap) {
    st.st_gap.resize(cap);
    st.st_delta.resize(cap);
    st.cap = cap;
  }
  float* u = ho ? st.u.get() : nullptr;
  float* su = ho ? st.su.get() : nullptr;
  double* sc = st.sc.get();
  int* fl = st.fl.get();
  k_zero_scalars<<<1, 64, 0, s>>>(sc, kL_NSC, fl, kLF_NFL);

  const float lambda = P.prox_reg_const;
  LPConst c;
  c.lambda = lambda;
  c.c_m1_over_M = (float)(-1.0 / (double)(float)L);
  c.c_2l_over_M = (float)(2 * lambda) / (float)L;
  c.c_1_over_M = (float)(1.0 / (double)(float)L);
  c.c_1m2_over_M = (float)(1 - 2.0 / (double)(float)L);
  c.qp_delta = P.qp_const;

  auto read_d = [&](const double* d) {
    double v = 0;
    DCRF_HIP(hipMemcpyAsync(&v, d, sizeof(double), hipMemcpyDeviceToHost, s));
    DCRF_HIP(hipStreamSynchronize(s));
    return v;
  };
  auto check_err = [&]() {
    int e = 0;
    DCRF_HIP(hipMemcpyAsync(&e, fl + kLF_ERR, sizeof(int), hipMemcpyDeviceToHost, s));
    DCRF_HIP(hipStreamSynchronize(s));
    if (e) throw Error(kMaxvalNegative, "Maxval < 0, skipping");
  };

  // ---- initialisation (:534-550) ------------------------------------------------
  k_lp_project<<<rg_grid(N, L, 8192), kBlock, 0, s>>>(OneImage{N}, L, init, st.Q.get(), st.lab.get(), 0);
  DCRF_HIP(hipMemcpyAsync(st.best.get(), st.Q.get(), NL * sizeof(float), hipMemcpyDeviceToDevice, s));
  int_energy_dev(st.lab.get(), sc + kL_EINT, sc + kL_TMP, s);
  double int_energy = read_d(sc + kL_EINT);
  lp_energy_dev(st.Q.get(), sc + kL_TMP + 1, sc + kL_TMP + 2, s);
  check_err();
  double energy = read_d(sc + kL_TMP + 1);
  LPResult res;
  res.init_lp_energy = energy;
  res.init_int_energy = int_energy;
  if (energy > int_energy) {
    k_onehot<<<grid_for(NL), kBlock, 0, s>>>(st.lab.get(), N, L, st.Q.get());
    lp_energy_dev(st.Q.get(), sc + kL_TMP + 1, sc + kL_TMP + 2, s);
    energy = read_d(sc + kL_TMP + 1);
  }
  double best_energy = energy, best_int_energy = int_energy;
  int it = 0, count = 0, inner_idx = 0;
  do {
    ++it;
    k_fill<<<grid_for(NL), kBlock, 0, s>>>(st.gamma.get(), NL, 1.f);
    DCRF_HIP(hipMemcpyAsync(st.curQ.get(), st.Q.get(), NL * sizeof(float), hipMemcpyDeviceToDevice, s));
    k_fill<<<grid_for(NL), kBlock, 0, s>>>(st.alpha.get(), NL, 0.f);
    if (ho) k_fill<<<grid_for(NL), kBlock, 0, s>>>(u, NL, 0.f);
    int pit = 0;
    do {
      ++pit;
      {
        KTimer tm("lp_Y", s);
        auto ylaunch = [&](unsigned g, int upd) {
          k_lp_Y<<<g, kBlock, 0, s>>>(OneImage{N}, L, c, st.alpha.get(), u, st.sa.get(), su, sc, upd, U_.get(),
                                      st.curQ.get(), st.Y.get());
        };
        // tuned on the first step (which only writes Y); later steps also
        // update alpha in place and reuse the choice
        const unsigned gy = pit == 1 ? Tuner::get().pick("lp_Y", N, grid_candidates(rq_grid(N, L)), s,
                                                          [&](unsigned g) { ylaunch(g, 0); })
                                     : Tuner::get().cached("lp_Y", N, rq_grid(N, L, 8192));
        ylaunch(gy, pit > 1 ? 1 : 0);
      }
      // qp-gamma sweeps in batches: pass 1 runs `nsw` sweeps gamma -> gamma2,
      // the decision finds where the reference's loop stops, the replay redoes
      // that many sweeps only when the stop comes before the end of the batch.
      // Batches are planned on the device (k_lp_sweep_decide writes the next
      // one's length), so the host enqueues kSweepAhead batches before it
      // waits for the status: one wait covers most inner steps.  A batch
      // after the loop's end does nothing; gamma is then in the buffer of
      // the last batch that ran.
      constexpr int kSweepAhead = 2;
      int done_sweeps = 0;
      k_lp_sweep_reset<<<1, 64, 0, s>>>(fl, 1, std::min(kSweepFirst, P.qp_max_iter));
      float* gbuf[2] = {st.gamma.get(), st.gamma2.get()};
      int enq = 0, nb = 0;
      while (true) {
        for (int b = 0; b < kSweepAhead; ++b, ++enq) {
          float* gin = gbuf[enq & 1];
          float* gout = gbuf[(enq + 1) & 1];
          {
            KTimer tm("lp_sweeps", s);
            if (L == 21)
              k_lp_sweeps<21, 21><<<nbp, kBlock, 0, s>>>(OneImage{N}, L, c, gin, gout, st.Y.get(), -1, 0, 0,
                                                         st.parts.get(), fl);
            else
              DCRF_MAXL_DISPATCH(L, (k_lp_sweeps<MAXL><<<nbp, kBlock, 0, s>>>(OneImage{N}, L, c, gin, gout,
                                                                             st.Y.get(), -1, 0, 0, st.parts.get(),
                                                                             fl)));
          }
          k_lp_sweep_sums<<<dim3(1, kSweepBatch), 1024, 0, s>>>(st.parts.get(), nbp, nullptr, st.swtot.get(), fl,
                                                                 1);
          k_lp_sweep_decide<<<1, 64, 0, s>>>(st.swtot.get(), 1, -1, P.qp_tol, P.qp_max_iter, sc, fl,
                                             st.hstatus.get());
          {
            KTimer tm("lp_sweeps_replay", s);
            if (L == 21)
              k_lp_sweeps<21, 21><<<nbp, kBlock, 0, s>>>(OneImage{N}, L, c, gin, gout, st.Y.get(), -1, 1, 0,
                                                         st.parts.get(), fl);
            else
              DCRF_MAXL_DISPATCH(L, (k_lp_sweeps<MAXL><<<nbp, kBlock, 0, s>>>(OneImage{N}, L, c, gin, gout,
                                                                             st.Y.get(), -1, 1, 0, st.parts.get(),
                                                                             fl)));
          }
        }
        DCRF_HIP(hipStreamSynchronize(s));
        const volatile int* hs = st.hstatus.get();
        done_sweeps = hs[1];
        nb = hs[2];
        if (hs[0] || done_sweeps >= P.qp_max_iter) break;
      }
      if (nb & 1) std::swap(st.gamma, st.gamma2);  // gamma after the last batch that ran
      res.qp_sweeps.push_back(done_sweeps);
      {
        KTimer tm("lp_betaQ", s);
        auto bq = [&](unsigned g) {
          k_lp_betaQ<<<g, kBlock, 0, s>>>(OneImage{N}, L, lambda, st.alpha.get(), u, st.gamma.get(), U_.get(),
                                          st.curQ.get(), st.Q.get(), st.mmparts.get());
        };
        const unsigned nbq = Tuner::get().pick("lp_betaQ", N, grid_candidates(rq_grid(N, L, 16384)), s, bq);
        bq(nbq);
        k_lp_minmax_final<<<1, 1024, 0, s>>>(st.mmparts.get(), nbq, nullptr, st.mm.get(), fl);
      }
      for (int k = 0; k < num_pairwise(); ++k) {
        Potential& p = pairwise(k);
        SliceArgs o;
        o.out = st.sa.get();
        o.accumulate = k > 0;
        compat_slice(p, o, L, s, [&](const SliceArgs& oo) {
          p.lat->ordered(st.Q.get(), st.mm.get(), L, oo, &obins_, k == 0);
        });
      }
      if (num_pairwise() == 0) k_fill<<<grid_for(NL), kBlock, 0, s>>>(st.sa.get(), NL, 0.f);
      if (ho) ucondgrad_dev(st.Q.get(), su, s);
      {
        KTimer tm("lp_gap", s);
        k_lp_gap<<<kRedBlocks, kBlock, 0, s>>>(OneImage{N}, L, st.alpha.get(), u, st.sa.get(), su, st.Q.get(),
                                               st.parts.get());
        finalize_sums<2><<<1, kBlock, 0, s>>>(st.parts.get(), kRedBlocks, sc + kL_GAP);
        k_lp_delta<<<1, 1, 0, s>>>(sc, lambda, st.st_gap.get(), st.st_delta.get(), inner_idx, cap, 1);
      }
      ++inner_idx;
    } while (pit < P.fw_max_iter);
    check_err();
    // project, renormalize, labelling (:746-752)
    {
      KTimer tm("lp_project", s);
      if (L == 21)
        k_lp_project_px<21><<<(unsigned)std::max<long>(1, std::min<long>((N + kBlock - 1) / kBlock, 16384)), kBlock, 0, s>>>(
            OneImage{N}, st.Q.get(), st.Q.get(), st.lab.get(), 1);
      else
        k_lp_project<<<rg_grid(N, L, 8192), kBlock, 0, s>>>(OneImage{N}, L, st.Q.get(), st.Q.get(), st.lab.get(), 1);
    }
    const double prev_int = int_energy;
    int_energy_dev(st.lab.get(), sc + kL_EINT, sc + kL_TMP, s);
    int_energy = read_d(sc + kL_EINT);
    res.int_energy.push_back(int_energy);
    if (P.best_int) {
      if (std::abs(int_energy - prev_int) < P.prox_energy_tol)
        ++count;
      else
        count = 0;
      if (count >= 5) break;
      if (int_energy < best_int_energy) {
        DCRF_HIP(hipMemcpyAsync(st.best.get(), st.Q.get(), NL * sizeof(float), hipMemcpyDeviceToDevice, s));
        best_int_energy = int_energy;
      }
    } else {
      const double prev = energy;
      lp_energy_dev(st.Q.get(), sc + kL_TMP + 1, sc + kL_TMP + 2, s);
      check_err();
      energy = read_d(sc + kL_TMP + 1);
      if (std::abs(energy - prev) < P.prox_energy_tol)
        ++count;
      else
        count = 0;
      if (count >= 5) break;
      if (energy < best_energy) {
        DCRF_HIP(hipMemcpyAsync(st.best.get(), st.Q.get(), NL * sizeof(float), hipMemcpyDeviceToDevice, s));
        best_energy = energy;
      }
    }
    if (P.less_confident_percent > 0) {  // :784-794
      k_fill_int<<<1, 64, 0, s>>>(st.lc.get(), 1, 0);
      k_less_confident<<<grid_for(N, kBlock, 4096), kBlock, 0, s>>>(OneImage{N}, L, st.best.get(), P.confidence_tol,
                                                          st.lc.get(), 1);
      int n_lc = 0;
      DCRF_HIP(hipMemcpyAsync(&n_lc, st.lc.get(), sizeof(int), hipMemcpyDeviceToHost, s));
      DCRF_HIP(hipStreamSynchronize(s));
      if (double(n_lc) / double(N) * 100 > (double)P.less_confident_percent) break;
    }
  } while (it < P.prox_max_iter);
  res.outer = it;
  DCRF_HIP(hipMemcpyAsync(out, st.best.get(), NL * sizeof(float), hipMemcpyDeviceToDevice, s));
  const int n = std::min(inner_idx, cap);
  res.gap.resize(n);
  res.delta.resize(n);
  if (n) {
    DCRF_HIP(hipMemcpyAsync(res.gap.data(), st.st_gap.get(), n * sizeof(double), hipMemcpyDeviceToHost, s));
    DCRF_HIP(hipMemcpyAsync(res.delta.data(), st.st_delta.get(), n * sizeof(double), hipMemcpyDeviceToHost, s));
  }
  DCRF_HIP(hipStreamSynchronize(s));
  return res;
}

}  // namespace dcrf
