// Pn superpixel (higher-order) terms on the device.
//
// QP side, bug-compatible with DenseCRF::multiplySuperPixels
// (src/densecrf.cpp:58-102): region r of size s_r sums p(l, pixel) over the
// LOOP INDEX pixel = 0..s_r-1, not over its members, and broadcasts back to
// columns 0..s_r-1.  With the distinct region sizes sorted ascending
// (u_0 < u_1 < ...), both become scans over the size groups:
//   pixels -> regions: NtoR(l, r) = sum_{g <= g(r)} sum_{pix in [u_{g-1}, u_g)} p(l, pix)
//   regions -> pixels: RtoN(pix, l) = sum_{g : u_g > pix} sum_{r in group g} p(l, r)
// evaluated in FP64 in a fixed order (the reference accumulates in float).
//
// LP side (correct member ids): computeUCondGrad (:803-830),
// the max-min term of compute_energy_LP_higher_order (:1020-1034) and the
// region term of assignment_energy_higher_order (:856-867).
#include <algorithm>
#include <numeric>

#include "crf.hpp"
#include "crf_kernels.hpp"

namespace dcrf {

struct CRF::HOState {
  int R = 0, G = 0, smax = 0, nmem = 0;
  std::vector<int> h_size;            // per region
  std::vector<int> h_layer_start;     // first region of each layer (+ end)
  std::vector<float> h_e, h_c;
  std::vector<int> h_mstart, h_mpix;  // member CSR (ascending pixel ids)
  DBuf<float> e, c;
  DBuf<int> usz, grp_start, grp_reg, reg_grp, mstart, mpix, layer_of;
  // QP state (L x R) and scratch
  DBuf<float> z, gz, sxz, cgz;
  DBuf<double> nrQ, nrS, segB, prefB;
  DBuf<float> hb;  // smax x L
  // LP scratch: per region min/max positions
  DBuf<int> mpos;  // 4 per region: min_pix, min_lab, max_pix, max_lab
  DBuf<double> parts;
  DBuf<double> eterm;  // compute_energy_higher_order's region term
  bool dirty = true;
  bool have_z = false;  // z holds the last QP's z_labels_
};

namespace {

// prefix over pixels, read at the size-group ends: out[g][l] = sum of
// P(pix, l) over pix < usz[g] (the bug-compatible region sums of
// multiplySuperPixels run over pixel indices 0..s_r-1); MODE 0: P = Q - 1,
// MODE 1: P = onehot(lab) - Q; each term in float, the sums in double.  One
// block per label: each thread adds a contiguous run of pixels, an exclusive
// scan of the runs across the block, then each thread walks its run again
// and writes the group ends that fall in it.
constexpr int kPixPrefixBlock = 1024;
template <int MODE>
__global__ void __launch_bounds__(kPixPrefixBlock) k_ho_pixprefix(const float* __restrict__ Q, const int* __restrict__ lab,
                                                                  int L, const int* __restrict__ usz, int G,
                                                                  double* __restrict__ out, const int* __restrict__ fl) {
  if (fl && fl[kF_DONE]) return;
  __shared__ double wtot[kPixPrefixBlock / 64];
  const int l = blockIdx.x, t = threadIdx.x, lane = t & 63, wid = t >> 6;
  const int smax = usz[G - 1];
  const int per = (smax + (int)blockDim.x - 1) / (int)blockDim.x;
  const int p0 = min(smax, t * per), p1 = min(smax, p0 + per);
  auto term = [&](int p) {
    const float q = Q[(size_t)p * L + l];
    const float v = MODE == 0 ? q - 1.f : (lab[p] == l ? 1.f : 0.f) - q;
    return (double)v;
  };
  double run = 0;
  for (int p = p0; p < p1; ++p) run += term(p);
  double inc = run;
  for (int o = 1; o < 64; o <<= 1) {
    const double a = __shfl_up(inc, o, 64);
    if (lane >= o) inc += a;
  }
  if (lane == 63) wtot[wid] = inc;
  __syncthreads();
  double acc = inc - run;
  for (int w = 0; w < wid; ++w) acc += wtot[w];
  int lo = 0, hi = G;  // first group ending after p0
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (usz[mid] > p0)
      hi = mid;
    else
      lo = mid + 1;
  }
  int g = lo;
  for (int p = p0; p < p1 && g < G; ++p) {
    acc += term(p);
    if (usz[g] == p + 1) out[(size_t)g++ * L + l] = acc;
  }
}

// inclusive prefix over groups (PREFIX) or suffix (!PREFIX) in double, one
// block per label: each thread adds a run of groups, an exclusive scan of the
// runs' totals across the block, then each thread writes its run
template <bool PREFIX>
__global__ void __launch_bounds__(kBlock) k_ho_scan(const double* __restrict__ B, double* __restrict__ out, int G,
                                                    int L, const int* __restrict__ fl) {
  if (fl && fl[kF_DONE]) return;
  __shared__ double wtot[kBlock / 64];
  const int l = blockIdx.x, t = threadIdx.x, lane = t & 63, wid = t >> 6;
  const int per = (G + kBlock - 1) / kBlock, i0 = min(G, t * per), i1 = min(G, i0 + per);
  auto at = [&](int i) { return (size_t)(PREFIX ? i : G - 1 - i) * L + l; };
  double run = 0;
  for (int i = i0; i < i1; ++i) run += B[at(i)];
  // exclusive scan of the runs: within the wave, then over the waves
  double inc = run;
  for (int o = 1; o < 64; o <<= 1) {
    const double a = __shfl_up(inc, o, 64);
    if (lane >= o) inc += a;
  }
  if (lane == 63) wtot[wid] = inc;
  __syncthreads();
  double base = inc - run;
  for (int w = 0; w < wid; ++w) base += wtot[w];
  for (int i = i0; i < i1; ++i) out[at(i)] = (base += B[at(i)]);
}

// per (group, label): sum over the group's regions of p(l, r) where
// p = (z - 1) * e (the argument of the broadcast in :382 and :431)
__global__ void __launch_bounds__(kBlock) k_ho_groupsum_zm1(const float* __restrict__ z, const float* __restrict__ e,
                                  const int* __restrict__ gstart, const int* __restrict__ greg, int G, int L,
                                  double* __restrict__ A, const int* __restrict__ fl) {
  if (fl && fl[kF_DONE]) return;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= G * L) return;
  const int g = t / L, l = t - g * L;
  double s = 0;
  for (int i = gstart[g]; i < gstart[g + 1]; ++i) {
    const int r = greg[i];
    s += (double)((z[(size_t)r * L + l] - 1.f) * e[r]);
  }
  A[(size_t)g * L + l] = s;
}

// hb(pix, l) = suffix(g*) with g* the first group whose size exceeds pix
__global__ void __launch_bounds__(kBlock) k_ho_bcast(const double* __restrict__ suf, const int* __restrict__ usz, int G, int smax, int L,
                           float* __restrict__ hb, const int* __restrict__ fl) {
  if (fl && fl[kF_DONE]) return;
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (size_t)smax * L) return;
  const int pix = (int)(t / L), l = (int)(t - (size_t)pix * L);
  int lo = 0, hi = G - 1;  // smallest g with usz[g] > pix (exists since pix < smax = usz[G-1])
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (usz[mid] > pix)
      hi = mid;
    else
      lo = mid + 1;
  }
  hb[t] = (float)suf[(size_t)lo * L + l];
}

// grad_y += hb on the first smax pixels (:382)
__global__ void __launch_bounds__(kBlock) k_ho_add_hb(float* __restrict__ gy, const float* __restrict__ hb, size_t n) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) gy[t] = gy[t] + hb[t];
}

// after NtoR(Q-1): grad_z and the higher-order energy
// init (:373-375):      grad_z = (1 + NtoR(Q-1)) .* c
// iteration (:432):     grad_z = c + NtoR(Q-1) .* c
// energy (:387, :435):  sum(z .* c) + sum((1 - z) .* c .* NtoR(Q-1))
// Grid-stride over (region, label); one partial per block (see k_ho_finalize).
__global__ void __launch_bounds__(kBlock) k_ho_gradz_energy(const float* __restrict__ z, const float* __restrict__ e,
                                  const int* __restrict__ reg_grp, const double* __restrict__ prefQ, int R, int L,
                                  int init, float* __restrict__ gz, double* __restrict__ nrQ,
                                  double* __restrict__ parts, const int* __restrict__ fl) {
  if (fl && fl[kF_DONE]) return;
  double v[1] = {0.0};
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < R * L; t += gridDim.x * blockDim.x) {
    const int r = t / L, l = t - r * L;
    const double nr = prefQ[(size_t)reg_grp[r] * L + l];
    nrQ[t] = nr;
    const float nrf = (float)nr, c = e[r];
    gz[t] = init ? (1.f + nrf) * c : c + nrf * c;
    v[0] += (double)(z[t] * c) + (double)((1.f - z[t]) * c) * nr;
  }
  block_sum<1>(v);
  if (threadIdx.x == 0) parts[blockIdx.x] = v[0];
}

// direction terms (:396-414): cond_grad_z, sx_z and the four scalars
// Grid-stride over (region, label); four partials per block (see k_ho_finalize).
__global__ void __launch_bounds__(kBlock) k_ho_direction(const float* __restrict__ z, const float* __restrict__ gz, const float* __restrict__ e,
                               const int* __restrict__ reg_grp, const double* __restrict__ prefS,
                               const double* __restrict__ nrQ, int R, int L, float* __restrict__ cgz,
                               float* __restrict__ sxz, double* __restrict__ parts, const int* __restrict__ fl) {
  if (fl[kF_DONE]) return;
  double v[4] = {0.0, 0.0, 0.0, 0.0};
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < R * L; t += gridDim.x * blockDim.x) {
    const int r = t / L, l = t - r * L;
    const float cg = gz[t] >= 0 ? 0.f : 1.f;  // descent_direction_z (src/qp.cpp:17-25)
    const float sz = cg - z[t];
    cgz[t] = cg;
    sxz[t] = sz;
    const float c = e[r];
    const double nrS = prefS[(size_t)reg_grp[r] * L + l];
    const float szc = sz * c;
    v[0] += (double)szc * nrQ[t];              // a
    v[1] += (double)((cg - 1.f) * c) * nrS;    // b
    v[2] += (double)sz;                        // sum(sx_z)
    v[3] += (double)szc * nrS;                 // denominator term
  }
  block_sum<4>(v);
  if (threadIdx.x == 0)
    for (int i = 0; i < 4; ++i) parts[4 * blockIdx.x + i] = v[i];
}

// the block partials of k_ho_direction (NV = 4: a, b, sum sx_z, denominator
// into sho[kH_A..kH_DEN]) or k_ho_gradz_energy (NV = 1: energy into
// sho[kH_E]) summed in block order; nothing once the QP has stopped
template <int NV>
__global__ void __launch_bounds__(kBlock) k_ho_finalize(const double* __restrict__ parts, int nparts,
                                                        double* __restrict__ out, const int* __restrict__ fl) {
  if (fl && fl[kF_DONE]) return;
  double v[NV];
#pragma unroll
  for (int i = 0; i < NV; ++i) v[i] = 0;
  for (int b = threadIdx.x; b < nparts; b += blockDim.x)
#pragma unroll
    for (int i = 0; i < NV; ++i) v[i] += parts[(size_t)b * NV + i];
  block_sum<NV>(v);
  if (threadIdx.x == 0)
#pragma unroll
    for (int i = 0; i < NV; ++i) out[i] = v[i];
}
static_assert(kH_B == kH_A + 1 && kH_SXZ == kH_A + 2 && kH_DEN == kH_A + 3, "direction scalars are consecutive");

// z += step * sx_z (:425)
__global__ void __launch_bounds__(kBlock) k_ho_zstep(float* __restrict__ z, const float* __restrict__ sxz, int n, const double* __restrict__ sc,
                           const int* __restrict__ fl) {
  if (fl[kF_DONE]) return;
  const float st = (float)sc[kS_STEP];
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < n; t += gridDim.x * blockDim.x) z[t] = z[t] + st * sxz[t];
}

// ---- LP side -------------------------------------------------------------------
// computeUCondGrad: a scan of each region (label-major, members in order) with
// int `min`/`max`.  The running int min after element k is trunc(prefix-min
// of Q); element k takes the min position iff Q_k < trunc(prefix-min before
// k) (the first element always does).  Likewise for the max.  NaN never takes
// a position and never enters a prefix (fminf / fmaxf skip it), as in the
// reference's comparisons.
//
// One block per region.  Thread t < nslot*L (nslot = blockDim / L) owns label
// l = t % L and the j-th (j = t / L) of nslot contiguous ranges of the
// members, so at each step the block reads nslot whole member rows.
//   pass 1: min and max of every (range, label), then of every label;
//   carry of (range, label): min / max over the earlier labels and over the
//           earlier ranges of its own label (exclusive prefix in scan order);
//   pass 2: each thread scans its range in order from its carry and keeps
//           the last member it takes;
//   the last label with a taker (and its last taker) gives the positions.
constexpr int kUcondBlock = 1024;
constexpr int kUcondBatch = 8;  // members whose loads are issued together
__global__ void __launch_bounds__(kUcondBlock) k_ucond_region(const float* __restrict__ Q, int L, const int* __restrict__ mstart,
                                                         const int* __restrict__ mpix, int* __restrict__ mpos) {
  __shared__ float rmn[kUcondBlock], rmx[kUcondBlock];  // per (range, label)
  __shared__ float lmn[kMaxLabels], lmx[kMaxLabels];    // per label
  __shared__ int bmn[kUcondBlock], bmx[kUcondBlock];    // per thread last takers
  const int r = blockIdx.x, t = threadIdx.x;
  const int nslot = blockDim.x / L, act = nslot * L;
  const int l = t % L, j = t / L;
  const int b = mstart[r], s = mstart[r + 1] - b;
  const int per = (s + nslot - 1) / nslot;
  const int m0 = t < act ? min(s, j * per) : s, m1 = min(s, m0 + per);
  // a batch of members' values, loads issued together (members past m1: NaN, skipped by fminf / fmaxf)
  auto batch = [&](int m, float (&q)[kUcondBatch]) {
    int px[kUcondBatch];
#pragma unroll
    for (int i = 0; i < kUcondBatch; ++i) px[i] = m + i < m1 ? mpix[b + m + i] : -1;
#pragma unroll
    for (int i = 0; i < kUcondBatch; ++i) q[i] = px[i] >= 0 ? Q[(size_t)px[i] * L + l] : NAN;
  };
  // pass 1
  float mn = INFINITY, mx = -INFINITY;
  for (int m = m0; m < m1; m += kUcondBatch) {
    float q[kUcondBatch];
    batch(m, q);
#pragma unroll
    for (int i = 0; i < kUcondBatch; ++i) {
      mn = fminf(mn, q[i]);
      mx = fmaxf(mx, q[i]);
    }
  }
  rmn[t] = mn;
  rmx[t] = mx;
  __syncthreads();
  if (t < L) {
    float a = INFINITY, c = -INFINITY;
    for (int k = 0; k < nslot; ++k) {
      a = fminf(a, rmn[k * L + t]);
      c = fmaxf(c, rmx[k * L + t]);
    }
    lmn[t] = a;
    lmx[t] = c;
  }
  __syncthreads();
  float emin = INFINITY, emax = -INFINITY;  // prefix before the range's first member
  if (t < act) {
    for (int k = 0; k < l; ++k) {
      emin = fminf(emin, lmn[k]);
      emax = fmaxf(emax, lmx[k]);
    }
    for (int k = 0; k < j; ++k) {
      emin = fminf(emin, rmn[k * L + l]);
      emax = fmaxf(emax, rmx[k * L + l]);
    }
  }
  // pass 2
  int best_min = -1, best_max = -1;
  for (int m = m0; m < m1; m += kUcondBatch) {
    float q[kUcondBatch];
    batch(m, q);
#pragma unroll
    for (int i = 0; i < kUcondBatch; ++i)
      if (m + i < m1) {
        // int state before the element: INT_MAX / INT_MIN initially, else trunc(prefix)
        const bool take_min = (emin == INFINITY) ? (2147483648.0f > q[i]) : ((float)(int)emin > q[i]);
        const bool take_max = (emax == -INFINITY) ? (-2147483648.0f < q[i]) : ((float)(int)emax < q[i]);
        if (take_min) best_min = m + i;
        if (take_max) best_max = m + i;
        emin = fminf(emin, q[i]);
        emax = fmaxf(emax, q[i]);
      }
  }
  bmn[t] = best_min;
  bmx[t] = best_max;
  __syncthreads();
  if (t == 0) {
    int lmin = 0, mmin = 0, lmax = 0, mmax = 0;
    bool fmin = false, fmax = false;
    for (int ll = L - 1; ll >= 0 && !(fmin && fmax); --ll) {
      int a = -1, c = -1;
      for (int k = 0; k < nslot; ++k) {
        a = max(a, bmn[k * L + ll]);
        c = max(c, bmx[k * L + ll]);
      }
      if (!fmin && a >= 0) {
        lmin = ll;
        mmin = a;
        fmin = true;
      }
      if (!fmax && c >= 0) {
        lmax = ll;
        mmax = c;
        fmax = true;
      }
    }
    // no taker at all leaves the first element (lab 0, first member)
    mpos[4 * r + 0] = mpix[b + mmin];
    mpos[4 * r + 1] = lmin;
    mpos[4 * r + 2] = mpix[b + mmax];
    mpos[4 * r + 3] = lmax;
  }
}

// Us(min) = c*e, then Us(max) = -c*e (the max write wins), regions of one layer
__global__ void __launch_bounds__(kBlock) k_ucond_write(const int* __restrict__ mpos, const float* __restrict__ e, const float* __restrict__ c,
                              int r0, int r1, int L, float* __restrict__ Us) {
  const int r = r0 + blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= r1) return;
  const float v = c[r] * e[r];
  Us[(size_t)mpos[4 * r + 0] * L + mpos[4 * r + 1]] = v;
  Us[(size_t)mpos[4 * r + 2] * L + mpos[4 * r + 3]] = -c[r] * e[r];
}

// sum_r e_r * max_l (max_members Q - min_members Q), one block per region
__global__ void __launch_bounds__(kBlock) k_ho_maxmin(const float* __restrict__ Q, int L, const int* __restrict__ mstart,
                            const int* __restrict__ mpix, const float* __restrict__ e, double* __restrict__ parts) {
  const int r = blockIdx.x;
  const int b = mstart[r], s = mstart[r + 1] - b;
  __shared__ float smax[kMaxLabels][4], smin[kMaxLabels][4];
  const int wid = threadIdx.x >> 6, lane = threadIdx.x & 63;
  for (int l = 0; l < L; ++l) {
    float mx = -INFINITY, mn = INFINITY;
    for (int m = threadIdx.x; m < s; m += blockDim.x) {
      const float q = Q[(size_t)mpix[b + m] * L + l];
      mx = fmaxf(mx, q);
      mn = fminf(mn, q);
    }
    for (int o = 32; o > 0; o >>= 1) {
      mx = fmaxf(mx, __shfl_xor(mx, o, 64));
      mn = fminf(mn, __shfl_xor(mn, o, 64));
    }
    if (lane == 0) {
      smax[l][wid] = mx;
      smin[l][wid] = mn;
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    float md = -INFINITY;
    for (int l = 0; l < L; ++l) {
      float mx = smax[l][0], mn = smin[l][0];
      for (int w = 1; w < (int)blockDim.x / 64; ++w) {
        mx = fmaxf(mx, smax[l][w]);
        mn = fminf(mn, smin[l][w]);
      }
      md = fmaxf(md, mx - mn);
    }
    parts[r] = (double)(e[r] * md);
  }
}

// assignment_energy_higher_order's region term (:856-867), bug-compatible:
// region r differs iff some loop index pix < s_r has lab[pix] != lab[first member]
__global__ void __launch_bounds__(kBlock) k_ho_assign(const int* __restrict__ lab, const int* __restrict__ mstart,
                            const int* __restrict__ mpix, const float* __restrict__ e, int R,
                            const int* __restrict__ fd, double* __restrict__ parts) {
  const int first_diff = *fd;
  double v[1] = {0.0};
  for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < R; r += gridDim.x * blockDim.x) {
    const int s = mstart[r + 1] - mstart[r];
    const int l0 = (int)(short)lab[mpix[mstart[r]]];
    const bool differ = l0 != (int)(short)lab[0] || first_diff < s;
    if (differ) v[0] += (double)e[r];
    v[0] += (double)((float)(R - 1) * e[r]);
  }
  block_sum<1>(v);
  if (threadIdx.x == 0) parts[blockIdx.x] = v[0];
}

// first pixel index whose label differs from lab[0] (N if none): the block's
// first such index by a wave and block minimum, then one atomic per block
__global__ void __launch_bounds__(kBlock) k_first_diff(const int* __restrict__ lab, int N, int* __restrict__ out) {
  __shared__ int wmin[kBlock / 64];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  int v = (i < N && (short)lab[i] != (short)lab[0]) ? i : INT_MAX;
  for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o, 64));
  if ((threadIdx.x & 63) == 0) wmin[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < kBlock / 64; ++w) v = min(v, wmin[w]);
    if (v != INT_MAX) atomicMin(out, v);
  }
}

// blocks of the (region, label) reductions: fixed by the sizes, so the sums
// are the same run to run
inline int ho_red_blocks(int R, int L) {
  return (int)std::max<long>(1, std::min<long>(((long)R * L + kBlock - 1) / kBlock, kRedBlocks));
}

}  // namespace

float* ho_hb(CRF::HOState* h) { return h->hb.get(); }
int ho_smax(CRF::HOState* h) { return h->smax; }

CRF::HOState* CRF::ho_state() {
  if (R_ == 0) return nullptr;
  HOState& h = *ho_;
  if (h.dirty) {
    // distinct sizes and size groups
    const int R = h.R;
    std::vector<int> order(R);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return h.h_size[a] < h.h_size[b]; });
    std::vector<int> usz, gstart, reg_grp(R);
    for (int i = 0; i < R; ++i) {
      const int sz = h.h_size[order[i]];
      if (usz.empty() || usz.back() != sz) {
        usz.push_back(sz);
        gstart.push_back(i);
      }
      reg_grp[order[i]] = (int)usz.size() - 1;
    }
    gstart.push_back(R);
    h.G = (int)usz.size();
    h.smax = usz.back();
    auto up = [&](DBuf<int>& d, const std::vector<int>& v) {
      d.resize(std::max<size_t>(v.size(), 1));
      DCRF_HIP(hipMemcpyAsync(d.get(), v.data(), v.size() * sizeof(int), hipMemcpyHostToDevice, s_));
    };
    up(h.usz, usz);
    up(h.grp_start, gstart);
    up(h.grp_reg, order);
    up(h.reg_grp, reg_grp);
    up(h.mstart, h.h_mstart);
    up(h.mpix, h.h_mpix);
    h.e.resize(R);
    h.c.resize(R);
    DCRF_HIP(hipMemcpyAsync(h.e.get(), h.h_e.data(), R * sizeof(float), hipMemcpyHostToDevice, s_));
    DCRF_HIP(hipMemcpyAsync(h.c.get(), h.h_c.data(), R * sizeof(float), hipMemcpyHostToDevice, s_));
    const size_t LR = (size_t)L_ * R, GL = (size_t)L_ * h.G;
    h.z.resize(LR);
    h.gz.resize(LR);
    h.sxz.resize(LR);
    h.cgz.resize(LR);
    h.nrQ.resize(LR);
    h.nrS.resize(LR);
    h.segB.resize(GL);
    h.prefB.resize(GL);
    h.hb.resize((size_t)h.smax * L_);
    h.mpos.resize(4 * (size_t)R);
    h.parts.resize(std::max(R, 4 * kRedBlocks));
    DCRF_HIP(hipStreamSynchronize(s_));
    h.dirty = false;
  }
  return &h;
}

void CRF::add_layer(const int* h_region, int R, const float* h_e, const float* h_c) {
  DCRF_REQUIRE(R > 0, "superpixel layer: R must be positive");
  if (!ho_) ho_ = std::make_unique<HOState>();
  HOState& h = *ho_;
  std::vector<std::vector<int>> mem(R);
  for (int i = 0; i < N_; ++i) {
    const int r = h_region[i];
    DCRF_REQUIRE(r >= 0 && r < R, "superpixel layer: region id out of range");
    mem[r].push_back(i);
  }
  if (h.h_mstart.empty()) h.h_mstart.push_back(0);
  h.h_layer_start.push_back(h.R);
  for (int r = 0; r < R; ++r) {
    if (mem[r].empty()) continue;  // regions without pixels are dropped
    h.h_size.push_back((int)mem[r].size());
    h.h_e.push_back(h_e[r]);
    h.h_c.push_back(h_c[r]);
    h.h_mpix.insert(h.h_mpix.end(), mem[r].begin(), mem[r].end());
    h.h_mstart.push_back((int)h.h_mpix.size());
    ++h.R;
  }
  h.nmem = (int)h.h_mpix.size();
  R_ = h.R;
  h.dirty = true;
  h.have_z = false;
}

// QP start: z = 1, grad_z, cond_grad_z, grad_y += RtoN((z-1).*c), energy term
void CRF::ho_qp_init(hipStream_t s, const float* Q, float* gy, double* sho) {
  HOState& h = *ho_;
  const int G = h.G, L = L_, R = h.R;
  k_fill<<<grid_for((size_t)R * L), kBlock, 0, s>>>(h.z.get(), (size_t)R * L, 1.f);
  h.have_z = true;
  k_ho_pixprefix<0><<<L, kPixPrefixBlock, 0, s>>>(Q, nullptr, L, h.usz.get(), G, h.prefB.get(), nullptr);
  const int nrb = ho_red_blocks(R, L);
  k_ho_gradz_energy<<<nrb, kBlock, 0, s>>>(h.z.get(), h.e.get(), h.reg_grp.get(), h.prefB.get(), R, L, 1, h.gz.get(),
                                           h.nrQ.get(), h.parts.get(), nullptr);
  k_ho_finalize<1><<<1, kBlock, 0, s>>>(h.parts.get(), nrb, sho + kH_E, nullptr);
  k_ho_groupsum_zm1<<<grid_for((size_t)G * L), kBlock, 0, s>>>(h.z.get(), h.e.get(), h.grp_start.get(),
                                                               h.grp_reg.get(), G, L, h.segB.get(), nullptr);
  k_ho_scan<false><<<L, kBlock, 0, s>>>(h.segB.get(), h.prefB.get(), G, L, nullptr);
  k_ho_bcast<<<grid_for((size_t)h.smax * L), kBlock, 0, s>>>(h.prefB.get(), h.usz.get(), G, h.smax, L, h.hb.get(),
                                                             nullptr);
  k_ho_add_hb<<<grid_for((size_t)h.smax * L), kBlock, 0, s>>>(gy, h.hb.get(), (size_t)h.smax * L);
  check_launch("ho_qp_init");
}

// after the argmin: cond_grad_z, sx_z, NtoR(sx) and the line-search terms
void CRF::ho_qp_direction(hipStream_t s, const float* Q, const int* lab, double* sho, int* fl) {
  HOState& h = *ho_;
  const int G = h.G, L = L_;
  k_ho_pixprefix<1><<<L, kPixPrefixBlock, 0, s>>>(Q, lab, L, h.usz.get(), G, h.prefB.get(), fl);
  const int nrb = ho_red_blocks(h.R, L);
  k_ho_direction<<<nrb, kBlock, 0, s>>>(h.z.get(), h.gz.get(), h.e.get(), h.reg_grp.get(), h.prefB.get(),
                                        h.nrQ.get(), h.R, L, h.cgz.get(), h.sxz.get(), h.parts.get(), fl);
  k_ho_finalize<4><<<1, kBlock, 0, s>>>(h.parts.get(), nrb, sho + kH_A, fl);
  check_launch("ho_qp_direction");
}

// after the step: z update and the new broadcast term RtoN((z-1).*c)
void CRF::ho_qp_zstep(hipStream_t s, double* sc, int* fl) {
  HOState& h = *ho_;
  const int G = h.G, L = L_;
  k_ho_zstep<<<grid_for((size_t)h.R * L), kBlock, 0, s>>>(h.z.get(), h.sxz.get(), h.R * L, sc, fl);
  k_ho_groupsum_zm1<<<grid_for((size_t)G * L), kBlock, 0, s>>>(h.z.get(), h.e.get(), h.grp_start.get(),
                                                               h.grp_reg.get(), G, L, h.segB.get(), fl);
  k_ho_scan<false><<<L, kBlock, 0, s>>>(h.segB.get(), h.prefB.get(), G, L, fl);
  k_ho_bcast<<<grid_for((size_t)h.smax * L), kBlock, 0, s>>>(h.prefB.get(), h.usz.get(), G, h.smax, L, h.hb.get(),
                                                             fl);
  check_launch("ho_qp_zstep");
}

// after the update of Q: NtoR(Q-1), grad_z and the energy term
void CRF::ho_qp_after(hipStream_t s, const float* Q, double* sho, int* fl) {
  HOState& h = *ho_;
  const int G = h.G, L = L_;
  k_ho_pixprefix<0><<<L, kPixPrefixBlock, 0, s>>>(Q, nullptr, L, h.usz.get(), G, h.prefB.get(), fl);
  const int nrb = ho_red_blocks(h.R, L);
  k_ho_gradz_energy<<<nrb, kBlock, 0, s>>>(h.z.get(), h.e.get(), h.reg_grp.get(), h.prefB.get(), h.R, L, 0,
                                           h.gz.get(), h.nrQ.get(), h.parts.get(), fl);
  k_ho_finalize<1><<<1, kBlock, 0, s>>>(h.parts.get(), nrb, sho + kH_E, fl);
  check_launch("ho_qp_after");
}

void CRF::ho_qp_finish(hipStream_t) {}

// the region term of compute_energy_higher_order (src/densecrf.cpp:974-975):
// sum(z .* c) + multiplySuperPixels((1 - z) .* c, Q - 1) with the z_labels_
// of the last QP; 0 without superpixel layers
double CRF::ho_energy_term(const float* Q) {
  HOState* hp = ho_state();
  if (!hp) return 0.0;
  HOState& h = *hp;
  DCRF_REQUIRE(h.have_z, "compute_energy_higher_order: z_labels_ are set by a QP run; run the QP first");
  hipStream_t s = s_;
  h.eterm.resize(1);
  k_ho_pixprefix<0><<<L_, kPixPrefixBlock, 0, s>>>(Q, nullptr, L_, h.usz.get(), h.G, h.prefB.get(), nullptr);
  const int nrb = ho_red_blocks(h.R, L_);
  k_ho_gradz_energy<<<nrb, kBlock, 0, s>>>(h.z.get(), h.e.get(), h.reg_grp.get(), h.prefB.get(), h.R, L_, 0,
                                           h.gz.get(), h.nrQ.get(), h.parts.get(), nullptr);
  k_ho_finalize<1><<<1, kBlock, 0, s>>>(h.parts.get(), nrb, h.eterm.get(), nullptr);
  check_launch("ho_energy_term");
  double e = 0;
  DCRF_HIP(hipMemcpyAsync(&e, h.eterm.get(), sizeof(double), hipMemcpyDeviceToHost, s));
  DCRF_HIP(hipStreamSynchronize(s));
  return e;
}

// z_labels_ after the last QP (L x R, region-major)
void CRF::z_labels(float* host) {
  HOState* h = ho_state();
  if (!h) return;
  DCRF_HIP(hipMemcpyAsync(host, h->z.get(), (size_t)h->R * L_ * sizeof(float), hipMemcpyDeviceToHost, s_));
  DCRF_HIP(hipStreamSynchronize(s_));
}

// computeUCondGrad (src/densecrf.cpp:803-830) into a zero-filled Us
void CRF::ucondgrad_dev(const float* Q, float* Us, hipStream_t s) {
  HOState& h = *ho_state();
  k_fill<<<grid_for((size_t)N_ * L_), kBlock, 0, s>>>(Us, (size_t)N_ * L_, 0.f);
  const int nl = (int)h.h_layer_start.size();
  k_ucond_region<<<h.R, kUcondBlock, 0, s>>>(Q, L_, h.mstart.get(), h.mpix.get(), h.mpos.get());
  for (int li = 0; li < nl; ++li) {  // later layers overwrite earlier ones, as in region order
    const int r0 = h.h_layer_start[li], r1 = li + 1 < nl ? h.h_layer_start[li + 1] : h.R;
    if (r1 > r0)
      k_ucond_write<<<grid_for(r1 - r0), kBlock, 0, s>>>(h.mpos.get(), h.e.get(), h.c.get(), r0, r1, L_, Us);
  }
  check_launch("ucondgrad");
}

void CRF::ucondgrad(const float* Q, float* Us) {
  DCRF_REQUIRE(R_ > 0, "computeUCondGrad: no superpixel layer");
  ucondgrad_dev(Q, Us, s_);
  DCRF_HIP(hipStreamSynchronize(s_));
}

// sum_r e_r * max_l(max - min) into *out (device double)
void CRF::ho_maxmin_dev(const float* Q, double* out, hipStream_t s) {
  HOState& h = *ho_state();
  k_ho_maxmin<<<h.R, kBlock, 0, s>>>(Q, L_, h.mstart.get(), h.mpix.get(), h.e.get(), h.parts.get());
  finalize_sums<1><<<1, kBlock, 0, s>>>(h.parts.get(), h.R, out);
  check_launch("ho_maxmin");
}

// region term of assignment_energy_higher_order into *out (device double)
void CRF::ho_assign_dev(const int* lab, double* out, int* scratch_int, hipStream_t s) {
  HOState& h = *ho_state();
  k_fill_int<<<1, 1, 0, s>>>(scratch_int, 1, N_);
  k_first_diff<<<grid_for(N_), kBlock, 0, s>>>(lab, N_, scratch_int);
  k_ho_assign<<<kRedBlocks, kBlock, 0, s>>>(lab, h.mstart.get(), h.mpix.get(), h.e.get(), h.R, scratch_int,
                                            h.parts.get());
  finalize_sums<1><<<1, kBlock, 0, s>>>(h.parts.get(), kRedBlocks, out);
  check_launch("ho_assign");
}

}  // namespace dcrf
