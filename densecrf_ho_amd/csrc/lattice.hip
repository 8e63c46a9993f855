// Permutohedral lattice on gfx950: build (elevation, hash, compaction,
// neighbours, splat gather lists) and the plain / ordered filters.
//
// Float arithmetic follows the reference operation by operation and the
// library is compiled with -ffp-contract=off, so the lattice tables are
// bit-identical to the reference's and the filters are too wherever the
// splat sums a lattice point in one chunk (sequential pixel order).
#include <hipcub/hipcub.hpp>

#include "lattice.hpp"
#include "tuner.hpp"

namespace dcrf {
namespace {

constexpr int kMaxD = 12;
constexpr int kChunk = 64;  // splat entries summed sequentially per chunk
constexpr int kOBatch = 8;  // ordered splat entries whose loads are issued together
constexpr int kOHist = 10;  // ordered splat histogram slots per label (bins 0..9)
constexpr int kOSBatch = 8;  // ordered splat entries whose loads are issued together (two 16-byte loads)
// entries of padding (pixel 0, weight 0) after the sorted entry lists, read by
// the last batch of the splats (at least kOBatch and kOSBatch)
constexpr int kSpPad = 16;
static_assert(kSpPad >= kOBatch && kSpPad >= kOSBatch, "splat batches read past the last entry");

// ---- packed keys: 4 int16 coordinates per 64-bit word ---------------------
template <int D>
struct PKey {
  static constexpr int W = (D + 3) / 4;
  uint64_t w[W];
};

template <int D>
__device__ inline PKey<D> pack(const int16_t* k) {
  PKey<D> p;
#pragma unroll
  for (int i = 0; i < PKey<D>::W; ++i) p.w[i] = 0;
#pragma unroll
  for (int i = 0; i < D; ++i) p.w[i / 4] |= (uint64_t)(uint16_t)k[i] << (16 * (i % 4));
  return p;
}
template <int D>
__device__ inline void unpack(const PKey<D>& p, int16_t* k) {
#pragma unroll
  for (int i = 0; i < D; ++i) k[i] = (int16_t)(uint16_t)(p.w[i / 4] >> (16 * (i % 4)));
}
template <int D>
__device__ inline bool key_eq(const PKey<D>& a, const PKey<D>& b) {
  bool eq = true;
#pragma unroll
  for (int i = 0; i < PKey<D>::W; ++i) eq = eq && (a.w[i] == b.w[i]);
  return eq;
}
template <int D>
__device__ inline uint32_t key_hash(const PKey<D>& k) {
  uint64_t x = 0x9E3779B97F4A7C15ull;
#pragma unroll
  for (int i = 0; i < PKey<D>::W; ++i) {
    x ^= k.w[i];
    x *= 0xbf58476d1ce4e5b9ull;
    x ^= x >> 31;
  }
  x *= 0x94d049bb133111ebull;
  x ^= x >> 29;
  return (uint32_t)x;
}

struct ElevConst {
  float scale[kMaxD];
  float invdp1, dp1;
};

// ---- Permutohedral::init, per pixel (src/permutohedral.cpp:203-288) --------
// Pixels at or past N (the padding of the last block of 4) take a zero
// feature vector and still produce keys, like the reference's SSE lanes.
// SCALAR: the portable init of a build without SSE (:364-446): no padding
// (Np = N), the nearer of floor and ceil with ties down, an int `sum` that
// adds each rd * down_factor in float and truncates, and the wrap-around of
// the barycentric coordinates in double.
template <int D, bool SCALAR>
__global__ void __launch_bounds__(kBlock) k_elevate(const float* __restrict__ feat, int N, int Np, ElevConst c, PKey<D>* __restrict__ ekeys,
                          int* __restrict__ rank_out, float* __restrict__ bary_out) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= Np) return;
  constexpr int D1 = D + 1;
  float f[D], elev[D1], rem0[D1], rk[D1], bc[D + 2];
#pragma unroll
  for (int j = 0; j < D; ++j) f[j] = i < N ? feat[(size_t)i * D + j] : 0.0f;
  float sm = 0.f;
#pragma unroll
  for (int j = D; j > 0; --j) {
    const float cf = f[j - 1] * c.scale[j - 1];
    const float jc = (float)j * cf;
    elev[j] = sm - jc;
    sm = sm + cf;
  }
  elev[0] = sm;
  float sum = 0.f;
  if constexpr (SCALAR) {
    int isum = 0;
#pragma unroll
    for (int j = 0; j < D1; ++j) {
      const float v = c.invdp1 * elev[j];
      const float up = ceilf(v) * c.dp1, down = floorf(v) * c.dp1;
      const int rd = (up - elev[j] < elev[j] - down) ? (int)(short)up : (int)(short)down;
      rem0[j] = (float)rd;
      isum = (int)((float)isum + (float)rd * c.invdp1);
    }
    sum = (float)isum;
  } else {
#pragma unroll
    for (int j = 0; j < D1; ++j) {
      float v = c.invdp1 * elev[j];
      v = __builtin_rintf(v);  // round half to even, as _mm_cvtps_epi32
      rem0[j] = v * c.dp1;
      sum = sum + v;
    }
  }
#pragma unroll
  for (int j = 0; j < D1; ++j) rk[j] = 0.f;
#pragma unroll
  for (int a = 0; a < D; ++a) {
    const float da = elev[a] - rem0[a];
#pragma unroll
    for (int b = a + 1; b < D1; ++b) {
      const float db = elev[b] - rem0[b];
      const float cc = (da < db) ? 1.f : 0.f;
      rk[a] = rk[a] + cc;
      rk[b] = rk[b] + (1.f - cc);
    }
  }
#pragma unroll
  for (int j = 0; j < D1; ++j) {
    rk[j] = rk[j] + sum;
    const float add = rk[j] < 0.f ? c.dp1 : 0.f;
    const float sub = rk[j] >= c.dp1 ? c.dp1 : 0.f;
    const float as = add - sub;
    rk[j] = rk[j] + as;
    rem0[j] = rem0[j] + as;
  }
#pragma unroll
  for (int j = 0; j < D + 2; ++j) bc[j] = 0.f;
#pragma unroll
  for (int j = 0; j < D1; ++j) {
    const float v = (elev[j] - rem0[j]) * c.invdp1;
    const int p = (int)((float)D - rk[j]);
    // p is a permutation of 0..D, so each slot sees at most one + and one -
#pragma unroll
    for (int q = 0; q < D1; ++q) {
      if (q == p) bc[q] = bc[q] + v;
      if (q == p + 1) bc[q] = bc[q] - v;
    }
    if (p + 1 == D1) bc[D1] = bc[D1] - v;
  }
  if constexpr (SCALAR)
    bc[0] = (float)((double)bc[0] + (1.0 + (double)bc[D + 1]));
  else
    bc[0] = bc[0] + (1.f + bc[D + 1]);
  int rki[D1];
#pragma unroll
  for (int j = 0; j < D1; ++j) rki[j] = (int)rk[j];
#pragma unroll
  for (int r = 0; r < D1; ++r) {
    int16_t key[D];
#pragma unroll
    for (int j = 0; j < D; ++j) {
      const int cn = (rki[j] <= D - r) ? r : r - D1;  // canonical simplex (:183-188)
      key[j] = (int16_t)(int)(rem0[j] + (float)cn);
    }
    const size_t e = (size_t)i * D1 + r;
    ekeys[e] = pack<D>(key);
    rank_out[e] = rki[r];
    bary_out[e] = bc[r];
  }
}

// Open-addressing insert with CAS.  Every table slot ends up holding the
// smallest entry index carrying its key (the key's first occurrence).
template <int D>
__global__ void __launch_bounds__(kBlock) k_insert(const PKey<D>* __restrict__ ekeys, int E, int* table, uint32_t mask,
                         int* __restrict__ slot_of, int* __restrict__ err) {
  const int e = blockIdx.x * kBlock + threadIdx.x;
  if (e >= E) return;
  const PKey<D> k = ekeys[e];
  uint32_t h = key_hash<D>(k) & mask;
  for (uint32_t probes = 0; probes <= mask; ++probes) {
    int cur = __hip_atomic_load(&table[h], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (cur < 0) {
      const int prev = atomicCAS(&table[h], -1, e);
      if (prev < 0) {
        slot_of[e] = (int)h;
        return;
      }
      cur = prev;
    }
    if (key_eq<D>(ekeys[cur], k)) {
      atomicMin(&table[h], e);
      slot_of[e] = (int)h;
      return;
    }
    h = (h + 1) & mask;
  }
  atomicExch(err, 1);  // table full: cannot happen with capacity >= 2E
}

__global__ void __launch_bounds__(kBlock) k_mark_first(const int* __restrict__ table, const int* __restrict__ slot_of, int E,
                             int* __restrict__ flag) {
  const int e = blockIdx.x * kBlock + threadIdx.x;
  if (e >= E) return;
  flag[e] = table[slot_of[e]] == e ? 1 : 0;
}

template <int D>
__global__ void __launch_bounds__(kBlock) k_assign(const int* __restrict__ table, const int* __restrict__ slot_of, const int* __restrict__ pos,
                         const int* __restrict__ flag, const PKey<D>* __restrict__ ekeys, int E,
                         int* __restrict__ offset, PKey<D>* __restrict__ ukeys) {
  const int e = blockIdx.x * kBlock + threadIdx.x;
  if (e >= E) return;
  offset[e] = pos[table[slot_of[e]]];
  if (flag[e]) ukeys[pos[e]] = ekeys[e];
}

template <int D>
__device__ inline int lookup(const PKey<D>& k, const int* __restrict__ table, uint32_t mask,
                             const PKey<D>* __restrict__ ekeys, const int* __restrict__ pos) {
  uint32_t h = key_hash<D>(k) & mask;
  for (uint32_t probes = 0; probes <= mask; ++probes) {
    const int cur = table[h];
    if (cur < 0) return -1;
    if (key_eq<D>(ekeys[cur], k)) return pos[cur];
    h = (h + 1) & mask;
  }
  return -1;
}

// neighbours along each of the d+1 lattice axes (src/permutohedral.cpp:317-330)
template <int D>
__global__ void __launch_bounds__(kBlock) k_neighbors(const PKey<D>* __restrict__ ukeys, int M, const int* __restrict__ table, uint32_t mask,
                            const PKey<D>* __restrict__ ekeys, const int* __restrict__ pos, int* __restrict__ nbr,
                            int16_t* __restrict__ keys16) {
  const size_t t = (size_t)blockIdx.x * kBlock + threadIdx.x;
  if (t >= (size_t)M * (D + 1)) return;
  const int j = (int)(t / M), i = (int)(t % M);
  int16_t key[D], n1[D], n2[D];
  unpack<D>(ukeys[i], key);
#pragma unroll
  for (int q = 0; q < D; ++q) {
    n1[q] = (int16_t)(key[q] - 1);
    n2[q] = (int16_t)(key[q] + 1);
  }
  if (j < D) {
#pragma unroll
    for (int q = 0; q < D; ++q)
      if (q == j) {
        n1[q] = (int16_t)(key[q] + D);
        n2[q] = (int16_t)(key[q] - D);
      }
  }
  nbr[t * 2 + 0] = lookup<D>(pack<D>(n1), table, mask, ekeys, pos);
  nbr[t * 2 + 1] = lookup<D>(pack<D>(n2), table, mask, ekeys, pos);
  if (j == 0)
#pragma unroll
    for (int q = 0; q < D; ++q) keys16[(size_t)i * D + q] = key[q];
}

// ---- splat gather lists -------------------------------------------------------
__global__ void __launch_bounds__(kBlock) k_count(const int* __restrict__ offset, int E, int* __restrict__ cnt) {
  const int e = blockIdx.x * kBlock + threadIdx.x;
  if (e < E) atomicAdd(&cnt[offset[e]], 1);
}
__global__ void __launch_bounds__(kBlock) k_iota(int* __restrict__ v, int n) {
  const int e = blockIdx.x * kBlock + threadIdx.x;
  if (e < n) v[e] = e;
}
__global__ void __launch_bounds__(kBlock) k_fill_sp(const int* __restrict__ ent, const float* __restrict__ bary, int E, int d1,
                          int* __restrict__ sp_pix, float* __restrict__ sp_w) {
  const int t = blockIdx.x * kBlock + threadIdx.x;
  if (t >= E + kSpPad) return;
  if (t >= E) {  // padding entries
    sp_pix[t] = 0;
    sp_w[t] = 0.f;
    return;
  }
  const int e = ent[t];
  sp_pix[t] = e / d1;
  sp_w[t] = bary[e];
}
// chunks per point (at least one so that every row of the value buffer is
// written, also for points that only padding lanes produced)
__global__ void __launch_bounds__(kBlock) k_nchunks(const int* __restrict__ cnt, int M, int* __restrict__ nch, int* __restrict__ multi_flag,
                          int* __restrict__ pn) {
  const int p = blockIdx.x * kBlock + threadIdx.x;
  if (p >= M) return;
  int n = (cnt[p] + kChunk - 1) / kChunk;
  if (n < 1) n = 1;
  nch[p] = n;
  multi_flag[p] = n > 1 ? 1 : 0;
  pn[p] = n > 1 ? n : 0;  // partial-sum slots are only needed for multi-chunk points
}
__global__ void __launch_bounds__(kBlock) k_chpoint(const int* __restrict__ chstart, const int* __restrict__ mpos,
                          const int* __restrict__ mflag, int M, int* __restrict__ chpoint, int* __restrict__ multi) {
  const int p = blockIdx.x * kBlock + threadIdx.x;
  if (p >= M) return;
  for (int c = chstart[p]; c < chstart[p + 1]; ++c) chpoint[c] = p;
  if (mflag[p]) multi[mpos[p]] = p;
}

// ---- filter kernels ---------------------------------------------------------
// Threads are mapped so that a wave covers floor(64/V) rows of V values
// (V<=64) or one row with a stride-64 loop (V>64).
struct RowMap {
  int gpw;  // rows per wave
  __device__ inline void at(int V, long& row, int& k0, int& kstep) const {
    at_block(V, (long)xcd_block(), row, k0, kstep);
  }
  // same mapping for an explicit (virtual) block index
  __device__ inline void at_block(int V, long blk, long& row, int& k0, int& kstep) const {
    const long gt = blk * kBlock + threadIdx.x;
    const long wave = gt >> 6;
    const int lane = (int)(gt & 63);
    if (V <= 64) {
      const int g = lane / V;
      row = g < gpw ? wave * gpw + g : -1;
      k0 = lane - g * V;
      kstep = V;
    } else {
      row = wave;
      k0 = lane;
      kstep = 64;
    }
  }
};
inline RowMap rowmap(int V) { return RowMap{V <= 64 ? 64 / V : 1}; }
inline unsigned rows_grid(long rows, int V) {
  const long gpw = V <= 64 ? 64 / V : 1;
  const long waves = (rows + gpw - 1) / gpw;
  return grid_for((size_t)waves * 64);
}

template <SplatIn KIND>
__device__ inline float splat_in(const SplatArgs& a, int pix, int k, int V) {
  if (KIND == SplatIn::kPlain) return a.in[(size_t)pix * V + k];
  if (KIND == SplatIn::kScaled) return a.in[(size_t)pix * V + k] * a.norm[pix];
  if (KIND == SplatIn::kSxScaled) {
    const float s = (a.lab[pix] == k ? 1.f : 0.f) - a.in[(size_t)pix * V + k];
    return s * a.norm[pix];
  }
  if (KIND == SplatIn::kSx) return (a.lab[pix] == k ? 1.f : 0.f) - a.in[(size_t)pix * V + k];
  return a.lab[pix] == k ? 0.f : 1.f;  // kNotOnehot
}

// values[p] = sum over the point's entries, in pixel order, of w * in(pixel)
template <SplatIn KIND>
__global__ void __launch_bounds__(kBlock) k_splat(SplatArgs a, int V, RowMap rm, const int* __restrict__ sp_pix, const float* __restrict__ sp_w,
                        const int* __restrict__ seg, const int* __restrict__ chstart, const int* __restrict__ chpoint,
                        const int* __restrict__ pstart, int nchunks, float* __restrict__ values,
                        float* __restrict__ part) {
  long c;
  int k0, ks;
  rm.at(V, c, k0, ks);
  if (c < 0 || c >= nchunks) return;
  const int p = chpoint[c];
  const int c0 = chstart[p], nc = chstart[p + 1] - c0;
  const int b = seg[p] + (int)(c - c0) * kChunk;
  const int e = min(seg[p + 1], b + kChunk);
  float* dst = nc == 1 ? values + (size_t)(p + 1) * V : part + (size_t)(pstart[p] + (c - c0)) * V;
  for (int k = k0; k < V; k += ks) {
    float acc = 0.f;
    int t = b;
    for (; t + 4 <= e; t += 4) {
      const int p0 = sp_pix[t], p1 = sp_pix[t + 1], p2 = sp_pix[t + 2], p3 = sp_pix[t + 3];
      const float w0 = sp_w[t], w1 = sp_w[t + 1], w2 = sp_w[t + 2], w3 = sp_w[t + 3];
      const float x0 = splat_in<KIND>(a, p0, k, V), x1 = splat_in<KIND>(a, p1, k, V);
      const float x2 = splat_in<KIND>(a, p2, k, V), x3 = splat_in<KIND>(a, p3, k, V);
      acc = acc + w0 * x0;
      acc = acc + w1 * x1;
      acc = acc + w2 * x2;
      acc = acc + w3 * x3;
    }
    for (; t < e; ++t) acc = acc + sp_w[t] * splat_in<KIND>(a, sp_pix[t], k, V);
    dst[k] = acc;
  }
}

// points whose entries span several chunks: sum the chunk partials in order
__global__ void __launch_bounds__(kBlock) k_combine(int V, RowMap rm, const int* __restrict__ multi, int nmulti, const int* __restrict__ chstart,
                          const int* __restrict__ pstart, const float* __restrict__ part, float* __restrict__ values) {
  long m;
  int k0, ks;
  rm.at(V, m, k0, ks);
  if (m < 0 || m >= nmulti) return;
  const int p = multi[m];
  const int c0 = pstart[p], c1 = c0 + chstart[p + 1] - chstart[p];
  for (int k = k0; k < V; k += ks) {
    float acc = part[(size_t)c0 * V + k];
    for (int c = c0 + 1; c < c1; ++c) acc = acc + part[(size_t)c * V + k];
    values[(size_t)(p + 1) * V + k] = acc;
  }
}

// one blur axis: new = old + 0.5 * (n1 + n2); SEQ evaluates the sum in double
// and rounds once (seqCompute, :518), otherwise float (sseCompute, :1035)
template <bool SEQ>
__global__ void __launch_bounds__(kBlock) k_blur(int V, RowMap rm, int M, const int* __restrict__ nbr, const float* __restrict__ src,
                       float* __restrict__ dst) {
  long i;
  int k0, ks;
  rm.at(V, i, k0, ks);
  if (i < 0 || i >= M) return;
  const int n1 = nbr[2 * i] + 1, n2 = nbr[2 * i + 1] + 1;
  for (int k = k0; k < V; k += ks) {
    const float o = src[(size_t)(i + 1) * V + k];
    const float s = src[(size_t)n1 * V + k] + src[(size_t)n2 * V + k];
    dst[(size_t)(i + 1) * V + k] = SEQ ? (float)((double)o + 0.5 * (double)s) : o + 0.5f * s;
  }
}


// slice_store with its two reads (norm, previous output) issued earlier by the caller
struct SlicePre {
  float norm, prev;
};
__device__ inline SlicePre slice_pre(const SliceArgs& o, int pix, int k, int V) {
  return SlicePre{o.norm ? o.norm[pix] : 1.f, o.accumulate ? o.out[(size_t)pix * V + k] : 0.f};
}
__device__ inline void slice_store_pre(const SliceArgs& o, int pix, int k, int V, float x, const SlicePre& pre) {
  if (o.norm) x = x * pre.norm;
  if (o.apply_scale) x = o.scale * x;
  if (o.post_mult != 1.f) x = o.post_mult * x;
  o.out[(size_t)pix * V + k] = o.accumulate ? pre.prev + x : x;
}
__device__ inline void slice_store(const SliceArgs& o, int pix, int k, int V, float x) {
  if (o.norm) x = x * o.norm[pix];
  if (o.apply_scale) x = o.scale * x;
  if (o.post_mult != 1.f) x = o.post_mult * x;
  float* dst = o.out + (size_t)pix * V + k;
  *dst = o.accumulate ? *dst + x : x;
}

// barycentric gather: SEQ uses (w*v)*alpha (:533), otherwise (w*alpha)*v (:1048)
template <bool SEQ>
__global__ void __launch_bounds__(kBlock) k_slice(int V, RowMap rm, int N, int d1, const int* __restrict__ offset, const float* __restrict__ bary,
                        const float* __restrict__ values, float alpha, SliceArgs o) {
  long pix;
  int k0, ks;
  rm.at(V, pix, k0, ks);
  if (pix < 0 || pix >= N) return;
  for (int k = k0; k < V; k += ks) {
    float acc = 0.f;
    for (int j = 0; j < d1; ++j) {
      const int off = offset[(size_t)pix * d1 + j] + 1;
      const float w = bary[(size_t)pix * d1 + j];
      const float v = values[(size_t)off * V + k];
      if (SEQ)
        acc = acc + (w * v) * alpha;
      else
        acc = acc + (w * alpha) * v;
    }
    slice_store(o, (int)pix, k, V, acc);
  }
}

// Storage of the ordered filter's lattice rows: float, or bfloat16 bits
// (round to nearest even); all arithmetic is float.
__device__ inline uint16_t f2bf(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__device__ inline float ld_v(const float* p, size_t i) { return p[i]; }
__device__ inline float ld_v(const uint16_t* p, size_t i) { return __uint_as_float((uint32_t)p[i] << 16); }
// two adjacent row values from a 4-byte (float) or 2-byte (bfloat16) aligned
// position: one 8-byte load of floats, two 2-byte loads of bfloat16
struct __attribute__((packed, aligned(4))) FloatPair {
  float x, y;
};
__device__ inline float2 ld_pair(const float* p, size_t i) {
  const FloatPair v = *reinterpret_cast<const FloatPair*>(p + i);
  return make_float2(v.x, v.y);
}
__device__ inline float2 ld_pair(const uint16_t* p, size_t i) { return make_float2(ld_v(p, i), ld_v(p, i + 1)); }
__device__ inline void st_v(float* p, size_t i, float v) { p[i] = v; }
__device__ inline void st_v(uint16_t* p, size_t i, float v) { p[i] = f2bf(v); }
// four consecutive row values: float4, or four bfloat16 in a uint2
__device__ inline float4 ld_quad(const float* p, size_t q) { return reinterpret_cast<const float4*>(p)[q]; }
__device__ inline float4 ld_quad(const uint16_t* p, size_t q) {
  const uint2 v = reinterpret_cast<const uint2*>(p)[q];
  return make_float4(__uint_as_float(v.x << 16), __uint_as_float(v.x & 0xffff0000u), __uint_as_float(v.y << 16),
                     __uint_as_float(v.y & 0xffff0000u));
}
__device__ inline void st_quad(float* p, size_t q, float4 v) { reinterpret_cast<float4*>(p)[q] = v; }
// streaming store (not kept in L2 for reuse): the blur's output rows
typedef float F4v __attribute__((ext_vector_type(4)));
typedef uint32_t U2v __attribute__((ext_vector_type(2)));
__device__ inline void st_quad_nt(float* p, size_t q, float4 v) {
  __builtin_nontemporal_store(F4v{v.x, v.y, v.z, v.w}, reinterpret_cast<F4v*>(p) + q);
}
__device__ inline void st_quad_nt(uint16_t* p, size_t q, float4 v) {
  const U2v o{(uint32_t)f2bf(v.x) | ((uint32_t)f2bf(v.y) << 16), (uint32_t)f2bf(v.z) | ((uint32_t)f2bf(v.w) << 16)};
  __builtin_nontemporal_store(o, reinterpret_cast<U2v*>(p) + q);
}
__device__ inline void st_quad(uint16_t* p, size_t q, float4 v) {
  uint2 o;
  o.x = (uint32_t)f2bf(v.x) | ((uint32_t)f2bf(v.y) << 16);
  o.y = (uint32_t)f2bf(v.z) | ((uint32_t)f2bf(v.w) << 16);
  reinterpret_cast<uint2*>(p)[q] = o;
}


// Bin-edge flags of the ordered filter, one per part of the lattice (one
// image, or each image of a union) plus flag[parts] = "any part".  A part
// whose flag is 0 takes the upper pass from T - LO[c+1] and skips its own.
// Bin-edge bookkeeping of one ordered-filter call, per part (image) of the
// lattice.  A (pixel, label) value on a bin edge has bin_up = bin_lo - 1; its
// (label, bin_lo) pair is its "combo".  A part with at most kMaxCombo combos
// and kPartEdges edge values takes the upper pass as Tb - LO[c+1] + D[c+1],
// D the filtered weight of its edge values of combo (label, c+1)
// ("correction" mode 1..kMaxCombo, the number of combos); a part with more
// has its upper pass splatted and blurred in full (mode kMaxCombo + 1).  Every
// decision is per part, so a part of a union lattice is filtered exactly as
// the image alone.  Per part p, at flag[p * kP_STRIDE + kP_*]: edge flag,
// mode, combo codes (label * 16 + bin, in code order), listed count, and a
// bitmap of the codes (64 labels x 16 bins); then two global flags (some part
// in correction mode, some part in full mode).
constexpr int kMaxCombo = 8, kPartEdges = 512;
constexpr int kP_EDGE = 0, kP_MODE = 1, kP_COMBO = 2, kP_NLIST = kP_COMBO + kMaxCombo, kP_BITS = kP_NLIST + 1,
              kP_STRIDE = kP_BITS + 32;
constexpr int kG_CORR = 0, kG_FULL = 1, kG_COUNT = 2;
inline size_t edge_ints(int parts) { return (size_t)parts * kP_STRIDE + kG_COUNT; }
struct EdgeMap {
  const int* flag;
  const uint16_t* pimg;   // part of each pixel (null: one part)
  const uint16_t* ppimg;  // part of each lattice point (null: one part)
  int parts;
  __device__ inline int ppart(long pix) const { return pimg ? pimg[pix] : 0; }
  __device__ inline int qpart(long p) const { return ppimg ? ppimg[p] : 0; }
  __device__ inline int mode(int part) const { return flag[part * kP_STRIDE + kP_MODE]; }
  __device__ inline int combo(int part, int c) const { return flag[part * kP_STRIDE + kP_COMBO + c]; }
  __device__ inline bool any_corr() const { return flag[parts * kP_STRIDE + kG_CORR] != 0; }
  __device__ inline bool any_full() const { return flag[parts * kP_STRIDE + kG_FULL] != 0; }
  __device__ inline bool point_full(long p) const { return mode(qpart(p)) > kMaxCombo; }
};

// k_combine of the lower pass, or of the upper pass (upper != 0) for the
// points of parts with bin edges only: V values of partial rows and point rows
// of stride RS
template <class T>
__global__ void __launch_bounds__(kBlock) k_combine_if(int V, int RS, RowMap rm, const int* __restrict__ multi,
                                                       int nmulti, const int* __restrict__ chstart,
                                                       const int* __restrict__ pstart, const float* __restrict__ part,
                                                       const int* __restrict__ orow, T* __restrict__ values,
                                                       EdgeMap edge, int upper) {
  if (upper && !edge.any_full()) return;
  const long nvb = (long)((nmulti + rm.gpw - 1) / rm.gpw * 64 + kBlock - 1) / kBlock;
  for (long vb = xcd_block(); vb < nvb; vb += gridDim.x) {
    long m;
    int k0, ks;
    rm.at_block(V, vb, m, k0, ks);
    if (m < 0 || m >= nmulti) continue;
    const int p = multi[m];
    if (upper && !edge.point_full(p)) continue;
    const int c0 = pstart[p], c1 = c0 + chstart[p + 1] - chstart[p];
    for (int k = k0; k < V; k += ks) {
      float acc = part[(size_t)c0 * RS + k];
      for (int c = c0 + 1; c < c1; ++c) acc = acc + part[(size_t)c * RS + k];
      st_v(values, (size_t)orow[p] * RS + k, acc);
    }
  }
}
__global__ void __launch_bounds__(kBlock) k_fill_f(float* __restrict__ p, size_t n, float v) {
  const size_t t = (size_t)blockIdx.x * kBlock + threadIdx.x;
  if (t < n) p[t] = v;
}

// ---- wide-row variants (V >= kStageMinV) ----------------------------------------
constexpr int kStageMinV = 8;

// One lane per (chunk, quad of values): each entry's four input values are
// one 16-byte load at 4-byte alignment (rows of V floats; the last quad of a
// row, when V % 4 != 0, loads its values one by one).  Entries are read in
// batches of kOBatch, the next batch's entry loads in flight during this
// batch's sums, and added in pixel order as the per-value loop does.
struct __attribute__((packed, aligned(4))) Float4u {
  float x, y, z, w;
};
struct __attribute__((packed, aligned(4))) Int4u {
  int x, y, z, w;
};
template <SplatIn KIND>
__device__ inline float4 splat_in4(const SplatArgs& a, int pix, int k, int V, bool whole) {
  float4 x = make_float4(0.f, 0.f, 0.f, 0.f);
  if (KIND != SplatIn::kNotOnehot) {
    const float* r = a.in + (size_t)pix * V + k;
    if (whole) {
      const Float4u u = *reinterpret_cast<const Float4u*>(r);
      x = make_float4(u.x, u.y, u.z, u.w);
    } else {
      x.x = r[0];
      if (k + 1 < V) x.y = r[1];
      if (k + 2 < V) x.z = r[2];
    }
  }
  if (KIND == SplatIn::kPlain) return x;
  if (KIND == SplatIn::kScaled) {
    const float n = a.norm[pix];
    return make_float4(x.x * n, x.y * n, x.z * n, x.w * n);
  }
  const int lb = a.lab[pix];
  auto oh = [&](int kk) { return lb == k + kk ? 1.f : 0.f; };
  if (KIND == SplatIn::kNotOnehot) return make_float4(1.f - oh(0), 1.f - oh(1), 1.f - oh(2), 1.f - oh(3));
  float4 d = make_float4(oh(0) - x.x, oh(1) - x.y, oh(2) - x.z, oh(3) - x.w);
  if (KIND == SplatIn::kSxScaled) {
    const float n = a.norm[pix];
    d = make_float4(d.x * n, d.y * n, d.z * n, d.w * n);
  }
  return d;
}
template <SplatIn KIND>
__global__ void __launch_bounds__(kBlock) k_splat_quad(SplatArgs a, int V, long nlanes,
                                                       const int* __restrict__ sp_pix, const float* __restrict__ sp_w,
                                                       const int* __restrict__ seg, const int* __restrict__ chstart,
                                                       const int* __restrict__ chpoint, const int* __restrict__ pstart,
                                                       float* __restrict__ values, float* __restrict__ part) {
  const int nq = (V + 3) >> 2;
  for (long g = (long)xcd_block() * kBlock + threadIdx.x; g < nlanes; g += (long)gridDim.x * kBlock) {
    const int c = (int)(g / nq), q = (int)(g - (long)c * nq);
    const int k = 4 * q;
    const bool whole = k + 4 <= V;
    const int p = chpoint[c];
    const int c0 = chstart[p], nc = chstart[p + 1] - c0;
    const int b = seg[p] + (c - c0) * kChunk;
    const int e = min(seg[p + 1], b + kChunk);
    float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
    int px[kOBatch];
    float w[kOBatch];
#pragma unroll
    for (int i = 0; i < kOBatch; ++i) {
      px[i] = sp_pix[b + i];
      w[i] = sp_w[b + i];
    }
    for (int t0 = b; t0 < e; t0 += kOBatch) {
      float4 x[kOBatch];
      float wc[kOBatch];
#pragma unroll
      for (int i = 0; i < kOBatch; ++i) {
        x[i] = splat_in4<KIND>(a, px[i], k, V, whole);
        wc[i] = w[i];
      }
      if (t0 + kOBatch < e)
#pragma unroll
        for (int i = 0; i < kOBatch; ++i) {
          px[i] = sp_pix[t0 + kOBatch + i];
          w[i] = sp_w[t0 + kOBatch + i];
        }
#pragma unroll
      for (int i = 0; i < kOBatch; ++i)
        if (t0 + i < e) {
          acc.x = acc.x + wc[i] * x[i].x;
          acc.y = acc.y + wc[i] * x[i].y;
          acc.z = acc.z + wc[i] * x[i].z;
          acc.w = acc.w + wc[i] * x[i].w;
        }
    }
    float* dst = (nc == 1 ? values + (size_t)(p + 1) * V : part + (size_t)(pstart[p] + (c - c0)) * V) + k;
    if (whole) {
      *reinterpret_cast<Float4u*>(dst) = Float4u{acc.x, acc.y, acc.z, acc.w};
    } else {
      dst[0] = acc.x;
      if (k + 1 < V) dst[1] = acc.y;
      if (k + 2 < V) dst[2] = acc.z;
    }
  }
}

// One lane per (pixel, quad of values) (V >= 8, non-SEQ rounding): each
// vertex row's four values are one 16-byte load at 4-byte alignment (the last
// quad of a row, when V % 4 != 0, loads its values one by one).  When the
// pixel has at most as many vertices as quads, lane j of the pixel's group
// loads vertex j's (offset, weight) and the group shares them by shuffles.
__device__ inline float4 ld_row4(const float* r, int k, int V, bool whole) {
  if (whole) {
    const Float4u u = *reinterpret_cast<const Float4u*>(r);
    return make_float4(u.x, u.y, u.z, u.w);
  }
  float4 x = make_float4(r[0], 0.f, 0.f, 0.f);
  if (k + 1 < V) x.y = r[1];
  if (k + 2 < V) x.z = r[2];
  return x;
}
__global__ void __launch_bounds__(kBlock) k_slice_quad(int V, long nlanes, int N, int d1,
                                                       const int* __restrict__ offset, const float* __restrict__ bary,
                                                       const float* __restrict__ values, float alpha, SliceArgs o) {
  const int nq = (V + 3) >> 2;
  const long g = (long)blockIdx.x * kBlock + threadIdx.x;
  const long gc = g < nlanes ? g : nlanes - 1;  // lanes past the end mirror the last one and store nothing
  const int pix = (int)(gc / nq), q = (int)(gc - (long)pix * nq);
  const int k = 4 * q;
  const bool whole = k + 4 <= V;
  // a pixel's lanes are consecutive; its group may straddle two waves only
  // when nq does not divide 64, so the shared loads need the whole group in
  // this wave: lanes whose group starts before this wave's lane 0 or ends
  // after lane 63 load their own vertices
  const int lane = (int)(threadIdx.x & 63);
  const int gl0 = lane - q;
  const bool shared = d1 <= nq && gl0 >= 0 && gl0 + nq <= 64;
  int myo = 0;
  float myw = 0.f;
  if (q < d1) {
    myo = offset[(size_t)pix * d1 + q] + 1;
    myw = bary[(size_t)pix * d1 + q];
  }
  float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
  for (int j = 0; j < d1; ++j) {
    int off;
    float w;
    const int so = __shfl(myo, gl0 + j), sw = __float_as_int(__shfl(myw, gl0 + j));
    if (shared) {
      off = so;
      w = __int_as_float(sw);
    } else {
      off = offset[(size_t)pix * d1 + j] + 1;
      w = bary[(size_t)pix * d1 + j];
    }
    const float4 v = ld_row4(values + (size_t)off * V + k, k, V, whole);
    const float wa = w * alpha;
    acc.x = acc.x + wa * v.x;
    acc.y = acc.y + wa * v.y;
    acc.z = acc.z + wa * v.z;
    acc.w = acc.w + wa * v.w;
  }
  if (g >= nlanes) return;
  float x[4] = {acc.x, acc.y, acc.z, acc.w};
  const float nm = o.norm ? o.norm[pix] : 1.f;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    if (o.norm) x[i] = x[i] * nm;
    if (o.apply_scale) x[i] = o.scale * x[i];
    if (o.post_mult != 1.f) x[i] = o.post_mult * x[i];
  }
  float* dst = o.out + (size_t)pix * V + k;
  if (whole) {
    if (o.accumulate) {
      const Float4u u = *reinterpret_cast<const Float4u*>(dst);
      x[0] = u.x + x[0];
      x[1] = u.y + x[1];
      x[2] = u.z + x[2];
      x[3] = u.w + x[3];
    }
    *reinterpret_cast<Float4u*>(dst) = Float4u{x[0], x[1], x[2], x[3]};
  } else {
    for (int i = 0; i < 3; ++i)
      if (k + i < V) dst[i] = o.accumulate ? dst[i] + x[i] : x[i];
  }
}

// one blur axis over rows of V >= 8 floats, one lane per (row, quad of
// values), the arithmetic of k_blur<false>
__global__ void __launch_bounds__(kBlock) k_blur_quad(int V, FastDiv divNQ, uint32_t nl,
                                                      const int* __restrict__ nbr, const float* __restrict__ src,
                                                      float* __restrict__ dst) {
  const uint32_t nq = divNQ.d, stride = gridDim.x * kBlock;
  for (uint32_t t = xcd_block() * kBlock + threadIdx.x; t < nl; t += stride) {
    const uint32_t i = divNQ.div(t);
    const int k = 4 * (int)(t - i * nq);
    const bool whole = k + 4 <= V;
    const int n1 = nbr[2 * i] + 1, n2 = nbr[2 * i + 1] + 1;
    const float4 o = ld_row4(src + (size_t)(i + 1) * V + k, k, V, whole);
    const float4 a = ld_row4(src + (size_t)n1 * V + k, k, V, whole);
    const float4 b = ld_row4(src + (size_t)n2 * V + k, k, V, whole);
    const float r[4] = {o.x + 0.5f * (a.x + b.x), o.y + 0.5f * (a.y + b.y), o.z + 0.5f * (a.z + b.z),
                        o.w + 0.5f * (a.w + b.w)};
    float* d = dst + (size_t)(i + 1) * V + k;
    if (whole) {
      *reinterpret_cast<Float4u*>(d) = Float4u{r[0], r[1], r[2], r[3]};
    } else {
      for (int j = 0; j < 3; ++j)
        if (k + j < V) d[j] = r[j];
    }
  }
}

// slice with the pixel's d+1 (offset, weight) pairs staged in LDS (V >= d+1)
template <bool SEQ>
__global__ void __launch_bounds__(kBlock) k_slice_staged(int V, RowMap rm, int N, int d1, const int* __restrict__ offset,
                               const float* __restrict__ bary, const float* __restrict__ values, float alpha,
                               SliceArgs o) {
  __shared__ int soff[kBlock];
  __shared__ float sbw[kBlock];
  long pix;
  int k0, ks;
  rm.at(V, pix, k0, ks);
  const bool act = pix >= 0 && pix < N;
  const int gb = (threadIdx.x & ~63) + ((threadIdx.x & 63) / V) * V;
  if (act && k0 < d1) {
    soff[gb + k0] = offset[(size_t)pix * d1 + k0] + 1;
    sbw[gb + k0] = bary[(size_t)pix * d1 + k0];
  }
  __syncthreads();
  if (!act) return;
  float acc = 0.f;
#pragma unroll 4
  for (int j = 0; j < d1; ++j) {
    const float w = sbw[gb + j];
    const float v = values[(size_t)soff[gb + j] * V + k0];
    if (SEQ)
      acc = acc + (w * v) * alpha;
    else
      acc = acc + (w * alpha) * v;
  }
  slice_store(o, (int)pix, k0, V, acc);
}

// ---- ordered filter ------------------------------------------------------------
// bins of src/permutohedral.cpp:548-589
__device__ inline int bin_lo(float q) { return min((int)floorf(q * 10.f), 9); }
__device__ inline int bin_up(float q) { return max((int)floor(((double)q - 1e-9) * 10.0), 0); }
__device__ inline float rescaled(const float* Q, size_t idx, const float* mm) {
  float q = Q[idx] - mm[0];
  if (mm[1] > 0.f) q = q / mm[1];
  return q;
}

// Ordered filter state, label-major rows of orow_stride(L) values per lattice
// point.  The lower pass adds w to bins 0..bin_lo (:548-569), so its bin 0
// holds every entry's weight: it equals the point's total weight Tb, which is
// splatted and blurred with the same float operations in the same order, and
// is not stored.  LO rows hold bins 1..9 of every label, label-major (bin b of
// label k at k*9 + b - 1) or bin-major (at (b-1)*L + k), chosen per lattice.
// The upper pass (+w to bins bin_up..9) equals Tb - LO[bin+1] whenever
// bin_up == bin_lo for every (pixel, label); when some q of a part sits on a
// bin edge (bin_up == bin_lo - 1) the upper pass of that part is splatted and
// blurred in full into UP rows holding bins 0..8 (bin 9 is Tb), bin b at k*9 + b
// (label-major) or b*L + k (bin-major).
constexpr int kOBins = 9;      // stored bins per label
// Row layout per lattice kernel specialisation: label-major for the d = 2
// kernels (three vertices per pixel: the slice's one 8-byte load per vertex
// pays), bin-major otherwise (with six vertices the label-major slice touches
// every line of a row per vertex and runs slower than two masked 4-byte loads
// on bin-major rows).  The splat and the slice of one lattice use the same
// specialisation, so they agree.
constexpr bool label_major(int tag) { return tag == 2; }

// Static per lattice, for the ordered slice: each (pixel, vertex) as
// {row of the point in the ordered state (int bits), barycentric weight, the
// point's total weight, 0}
__global__ void __launch_bounds__(kBlock) k_ovtx(const int* __restrict__ offset, const float* __restrict__ bary,
                                                 const float* __restrict__ Tb, const int* __restrict__ orow, long E,
                                                 float4* __restrict__ vtx) {
  const long t = (long)blockIdx.x * kBlock + threadIdx.x;
  if (t >= E) return;
  const int p = offset[t];
  vtx[t] = make_float4(__int_as_float(orow[p]), bary[t], Tb[p + 1], 0.f);
}

// Morton code of each point's key (bit b of coordinate j at bit b*d + j,
// min(64/d, 30) bits per coordinate, offset to be non-negative; wider keys
// wrap, which only loosens the locality)
__global__ void __launch_bounds__(kBlock) k_morton(const int16_t* __restrict__ keys, int M, int d,
                                                   unsigned long long* __restrict__ code, int* __restrict__ ids) {
  const int p = blockIdx.x * kBlock + threadIdx.x;
  if (p >= M) return;
  const int bits = min(64 / d, 30);
  unsigned long long c = 0;
  for (int j = 0; j < d; ++j) {
    const uint32_t v = (uint32_t)((int)keys[(size_t)p * d + j] + (1 << (bits - 1)));
    for (int b = 0; b < bits; ++b) c |= (unsigned long long)((v >> b) & 1u) << (b * d + j);
  }
  code[p] = c;
  ids[p] = p;
}
// rows from the sorted order: orow[order[r]] = r + 1; the part of each row
__global__ void __launch_bounds__(kBlock) k_orow(const int* __restrict__ order, int M,
                                                 const uint16_t* __restrict__ ppimg, int* __restrict__ orow,
                                                 uint16_t* __restrict__ oppart) {
  const int r = blockIdx.x * kBlock + threadIdx.x;
  if (r >= M) return;
  orow[order[r]] = r + 1;
  if (ppimg) oppart[r] = ppimg[order[r]];
}
// part of each point in sorted order: out[r] = part[order[r]]
__global__ void __launch_bounds__(kBlock) k_gather_u16(const uint16_t* __restrict__ part,
                                                       const int* __restrict__ order, int M,
                                                       uint16_t* __restrict__ out) {
  const int r = blockIdx.x * kBlock + threadIdx.x;
  if (r < M) out[r] = part[order[r]];
}
// blur neighbours of row r along every axis, as rows (row - 1, -1 if absent)
__global__ void __launch_bounds__(kBlock) k_onbr(const int* __restrict__ nbr, const int* __restrict__ order,
                                                 const int* __restrict__ orow, int M, int d1,
                                                 int* __restrict__ onbr) {
  const long t = (long)blockIdx.x * kBlock + threadIdx.x;
  if (t >= (long)M * d1) return;
  const int j = (int)(t / M), r = (int)(t - (long)j * M);
  const int i = order[r];
  for (int h = 0; h < 2; ++h) {
    const int n = nbr[((size_t)j * M + i) * 2 + h];
    onbr[((size_t)j * M + r) * 2 + h] = n >= 0 ? orow[n] - 1 : -1;
  }
}

// One lane per (point chunk, quad of labels): each entry's bins of the four
// labels are one 4-byte word of its pixel's bins row (rows of LP bytes, LP a
// multiple of 8; labels past L are row padding, computed and dropped).  The
// chunk's entries are read in batches of kOSBatch (all of a batch's loads in
// flight together, the next batch's entry loads in flight during this batch's
// sums; the entry lists hold kOSBatch entries of padding, pixel 0 with weight
// 0, and entries past the chunk's end get weight 0).
// Each lane keeps a histogram of its chunk's weights by bin, per label, in its
// own LDS slots: per (entry, label) the slot of the entry's bin is read,
// added to and written back, in entry order (the slots are the lane's own, and
// a wave's LDS operations execute in order, so each read sees the previous
// write).  The
// cumulative rows are the histogram's suffix sums (lower pass: bins 1..9 get
// the weights of bins >= b) or prefix sums (upper pass: bins 0..8 get the
// weights of bins <= b), taken once per chunk.  Slot (kk, c) of thread t is
// word (kk * 10 + c) * kBlock + t: the 32 lanes of each half-wave hit 32
// different banks whatever their bins.
template <int TAG, class T>  // TAG: lattice dim, +100 for the upper pass
__global__ void __launch_bounds__(kBlock) k_osplat(const uint8_t* __restrict__ bins, int L, int LP, int RS,
                                                   long nlanes, const int* __restrict__ sp_pix,
                                                   const float* __restrict__ sp_w, const int* __restrict__ seg,
                                                   const int* __restrict__ chstart, const int* __restrict__ chpoint,
                                                   const int* __restrict__ pstart, int nchunks,
                                                   const int* __restrict__ orow, T* __restrict__ values,
                                                   float* __restrict__ part, EdgeMap edge) {
  constexpr bool kUpper = TAG >= 100;
  constexpr bool lmaj = label_major(TAG % 100);
  if (kUpper && !edge.any_full()) return;
  __shared__ float hist[4 * kOHist * kBlock];
  float* hl = hist + threadIdx.x;  // this lane's slots, stride kBlock
#pragma unroll
  for (int i = 0; i < 4 * kOHist; ++i) hl[i * kBlock] = 0.f;
  const int nq = (L + 3) >> 2;
  for (long g = (long)xcd_block() * kBlock + threadIdx.x; g < nlanes; g += (long)gridDim.x * kBlock) {
    const int c = (int)(g / nq), q = (int)(g - (long)c * nq);
    const int p = chpoint[c];
    if (kUpper && !edge.point_full(p)) continue;
    const int c0 = chstart[p], nc = chstart[p + 1] - c0;
    const int b = seg[p] + (c - c0) * kChunk;
    const int e = min(seg[p + 1], b + kChunk);
    static_assert(kOSBatch == 8, "entries load as two 16-byte words");
    int px[kOSBatch];
    float w[kOSBatch];
    // kOSBatch consecutive entries: two 16-byte loads per list (4-byte aligned)
    auto entries = [&](int t) {
      const Int4u p0 = *reinterpret_cast<const Int4u*>(sp_pix + t), p1 = *reinterpret_cast<const Int4u*>(sp_pix + t + 4);
      const Float4u w0 = *reinterpret_cast<const Float4u*>(sp_w + t), w1 = *reinterpret_cast<const Float4u*>(sp_w + t + 4);
      px[0] = p0.x, px[1] = p0.y, px[2] = p0.z, px[3] = p0.w, px[4] = p1.x, px[5] = p1.y, px[6] = p1.z, px[7] = p1.w;
      w[0] = w0.x, w[1] = w0.y, w[2] = w0.z, w[3] = w0.w, w[4] = w1.x, w[5] = w1.y, w[6] = w1.z, w[7] = w1.w;
    };
    entries(b);
    for (int t0 = b; t0 < e; t0 += kOSBatch) {
      uint32_t bw[kOSBatch];
      float wc[kOSBatch];
#pragma unroll
      for (int i = 0; i < kOSBatch; ++i) {
        bw[i] = *reinterpret_cast<const uint32_t*>(bins + (size_t)px[i] * LP + 4 * q);
        wc[i] = t0 + i < e ? w[i] : 0.f;
      }
      if (t0 + kOSBatch < e) entries(t0 + kOSBatch);  // the next batch's entries load during this batch's sums
#pragma unroll
      for (int i = 0; i < kOSBatch; ++i) {
        const float wi = wc[i];
#pragma unroll
        for (int kk = 0; kk < 4; ++kk) {
          const int c = min(kUpper ? (int)((bw[i] >> (8 * kk + 4)) & 15u) : (int)((bw[i] >> (8 * kk)) & 15u),
                            kOHist - 1);
          float* slot = &hl[(kk * kOHist + c) * kBlock];
          *slot = *slot + wi;
        }
      }
    }
    float acc[4][kOBins];
#pragma unroll
    for (int kk = 0; kk < 4; ++kk) {
      float h[kOHist];
#pragma unroll
      for (int c = 0; c < kOHist; ++c) {
        h[c] = hl[(kk * kOHist + c) * kBlock];
        hl[(kk * kOHist + c) * kBlock] = 0.f;  // the lane's next chunk starts from zero
      }
      float sum = kUpper ? h[0] : h[kOHist - 1];
      if (kUpper) {  // bin r (slot r, r = 0..8): weights of bins 0..r
        acc[kk][0] = sum;
#pragma unroll
        for (int r = 1; r < kOBins; ++r) acc[kk][r] = sum = sum + h[r];
      } else {  // bin r + 1 (slot r): weights of bins r + 1..9
        acc[kk][kOBins - 1] = sum;
#pragma unroll
        for (int r = kOBins - 2; r >= 0; --r) acc[kk][r] = sum = sum + h[r + 1];
      }
    }
    // label-major rows: bin slot r of label k at k * kOBins + r, so the
    // lane's four labels are 36 consecutive values (nine aligned quads);
    // bin-major rows: at r * L + k, so the lane's four labels of one bin are
    // one 16-byte store at 4-byte alignment (float rows; bfloat16 rows and a
    // row's last, partial quad store value by value)
    auto slot = [&](int f) { return acc[f / kOBins][f % kOBins]; };
    const bool quads = lmaj && 4 * q + 4 <= L;
    const bool bquads = !lmaj && 4 * q + 4 <= L;
    auto bm_quads = [&](float* dst) {
#pragma unroll
      for (int r = 0; r < kOBins; ++r)
        *reinterpret_cast<Float4u*>(dst + (size_t)r * L + 4 * q) = Float4u{acc[0][r], acc[1][r], acc[2][r], acc[3][r]};
    };
    auto at = [&](int kk, int r) { return lmaj ? (size_t)(4 * q + kk) * kOBins + r : (size_t)r * L + 4 * q + kk; };
    if (nc == 1) {  // the point's row
      const size_t r0 = (size_t)orow[p] * RS;
      if (quads) {
#pragma unroll
        for (int i = 0; i < kOBins; ++i)
          st_quad(values, r0 / 4 + (size_t)kOBins * q + i,
                  make_float4(slot(4 * i), slot(4 * i + 1), slot(4 * i + 2), slot(4 * i + 3)));
      } else if (std::is_same<T, float>::value && bquads) {
        bm_quads(reinterpret_cast<float*>(values) + r0);
      } else {
#pragma unroll
        for (int kk = 0; kk < 4; ++kk)
          if (4 * q + kk < L)
#pragma unroll
            for (int r = 0; r < kOBins; ++r) st_v(values, r0 + at(kk, r), acc[kk][r]);
      }
      if (q == 0)
        for (int x = kOBins * L; x < RS; ++x) st_v(values, r0 + x, 0.f);  // row padding
    } else {  // a chunk partial (float, rows of stride RS), summed by k_combine_if
      float* dst = part + (size_t)(pstart[p] + (c - c0)) * RS;
      if (quads) {
#pragma unroll
        for (int i = 0; i < kOBins; ++i)
          st_quad(dst, (size_t)kOBins * q + i,
                  make_float4(slot(4 * i), slot(4 * i + 1), slot(4 * i + 2), slot(4 * i + 3)));
      } else if (bquads) {
        bm_quads(dst);
      } else {
#pragma unroll
        for (int kk = 0; kk < 4; ++kk)
          if (4 * q + kk < L)
#pragma unroll
            for (int r = 0; r < kOBins; ++r) dst[at(kk, r)] = acc[kk][r];
      }
    }
  }
}

// one blur axis over the ordered state: one lane per (point, quad of the
// row), over the lower rows; when some part is in correction mode also the
// correction rows (kMaxCombo floats, two quads per point, in their own
// buffers); when some part is in full mode the upper rows (half_q quads
// further) of that part's points (a point's neighbours are in its own part)
template <class T>
__device__ inline void blur_quad(const int* __restrict__ nbr, const T* __restrict__ src, T* __restrict__ dst,
                                 size_t base, uint32_t i, uint32_t R4, uint32_t q, bool nt) {
  const int n1 = nbr[2 * i] + 1, n2 = nbr[2 * i + 1] + 1;
  const float4 o = ld_quad(src, base + (size_t)(i + 1) * R4 + q);
  const float4 a = ld_quad(src, base + (size_t)n1 * R4 + q);
  const float4 b = ld_quad(src, base + (size_t)n2 * R4 + q);
  float4 r;
  r.x = o.x + 0.5f * (a.x + b.x);
  r.y = o.y + 0.5f * (a.y + b.y);
  r.z = o.z + 0.5f * (a.z + b.z);
  r.w = o.w + 0.5f * (a.w + b.w);
  if (nt)
    st_quad_nt(dst, base + (size_t)(i + 1) * R4 + q, r);
  else
    st_quad(dst, base + (size_t)(i + 1) * R4 + q, r);
}
template <int TAG, class T>
__global__ void __launch_bounds__(kBlock) k_blur2(int R4, FastDiv divR4, uint32_t nq, size_t half_q, uint32_t M,
                                                  const int* __restrict__ nbr, const T* __restrict__ src,
                                                  T* __restrict__ dst, const float* __restrict__ dsrc,
                                                  float* __restrict__ ddst, EdgeMap edge, int nt) {
  const uint32_t stride = gridDim.x * kBlock, t0 = xcd_block() * kBlock + threadIdx.x;
  for (uint32_t t = t0; t < nq; t += stride) {  // lower rows
    const uint32_t i = divR4.div(t);
    blur_quad(nbr, src, dst, 0, i, (uint32_t)R4, t - i * (uint32_t)R4, nt);
  }
  if (edge.any_corr())  // correction rows: two quads per point
    for (uint32_t t = t0; t < 2 * M; t += stride) blur_quad(nbr, dsrc, ddst, 0, t >> 1, 2u, t & 1u, false);
  if (edge.any_full())  // upper rows of the points of parts in full mode
    for (uint32_t t = t0; t < nq; t += stride) {
      const uint32_t i = divR4.div(t);
      if (edge.point_full(i)) blur_quad(nbr, src, dst, half_q, i, (uint32_t)R4, t - i * (uint32_t)R4, nt);
    }
}

// bins of every (pixel, label) of the rescaled field, packed lo | up << 4,
// in rows of LP bytes per pixel; values on a bin edge (bin_up != bin_lo) set
// their part's flag and combo bit and are listed (the first kPartEdges of
// each part)
// A lattice that is the union of several images' lattices rescales each
// image with its own {min, max}: pimg[pixel] picks the pair mm[2*img].
__global__ void __launch_bounds__(kBlock) k_obin(const float* __restrict__ Q, const float* __restrict__ mm, int N,
                                                 int L, FastDiv divNQ, uint8_t* __restrict__ bins, int LP,
                                                 int* __restrict__ edge, const uint16_t* __restrict__ pimg,
                                                 int2* __restrict__ elist) {
  // one lane per (pixel, quad of labels): four bins, one 4-byte store (padding labels get 0)
  const uint32_t nq = divNQ.d, n = (uint32_t)N * nq;
  for (uint32_t g = blockIdx.x * kBlock + threadIdx.x; g < n; g += gridDim.x * kBlock) {
    const uint32_t pix = divNQ.div(g), q = g - pix * nq;
    const int img = pimg ? pimg[pix] : 0;
    const int k0 = 4 * (int)q;
    // the quad's four values in one 16-byte load (4-byte aligned rows of L floats)
    float4 v4;
    if (k0 + 4 <= L) {
      const Float4u u = *reinterpret_cast<const Float4u*>(Q + (size_t)pix * L + k0);
      v4 = make_float4(u.x, u.y, u.z, u.w);
    } else {
      const float* r = Q + (size_t)pix * L + k0;
      v4 = make_float4(r[0], k0 + 1 < L ? r[1] : 0.f, k0 + 2 < L ? r[2] : 0.f, 0.f);
    }
    const float vq[4] = {v4.x, v4.y, v4.z, v4.w};
    const float m0 = mm[2 * img], m1 = mm[2 * img + 1];
    uint32_t word = 0;
#pragma unroll
    for (int kk = 0; kk < 4; ++kk) {
      const int k = k0 + kk;
      if (k >= L) break;
      float qv = vq[kk] - m0;  // rescaled() on the loaded value
      if (m1 > 0.f) qv = qv / m1;
      const int lo = bin_lo(qv), up = bin_up(qv);
      if (lo != up) {  // on a bin edge: flag its part and its combo, and list it
        int* P = edge + (size_t)img * kP_STRIDE;
        if (P[kP_EDGE] == 0) atomicOr(P + kP_EDGE, 1);
        const int code = k * 16 + lo;
        const int bwd = (int)(1u << (code & 31));
        if ((P[kP_BITS + (code >> 5)] & bwd) == 0) atomicOr(P + kP_BITS + (code >> 5), bwd);
        const int at = atomicAdd(P + kP_NLIST, 1);
        if (at < kPartEdges) elist[(size_t)img * kPartEdges + at] = make_int2((int)pix, code);
      }
      word |= (uint32_t)(lo | (up << 4)) << (8 * kk);
    }
    *reinterpret_cast<uint32_t*>(bins + (size_t)pix * LP + 4 * q) = word;
  }
}

// One block per part after k_obin: the part's combos in code order (label,
// then bin) and its mode; in correction mode each listed edge value's vertex
// weights are added to its combo's entry of dfix (64-bit fixed point in 2^-40
// units: integer atomics give the same sums in any order)
constexpr double kDFix = 1099511627776.0;  // 2^40
__global__ void __launch_bounds__(1024) k_ocombo(int* __restrict__ edge, int parts, int d1,
                                                 const int2* __restrict__ elist, const int* __restrict__ offset,
                                                 const float* __restrict__ bary, const int* __restrict__ orow,
                                                 unsigned long long* __restrict__ dfix) {
  __shared__ int wsum[16];
  const int part = blockIdx.x;
  int* P = edge + (size_t)part * kP_STRIDE;
  if (!P[kP_EDGE]) return;  // mode stays 0
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
  const bool bit = ((uint32_t)P[kP_BITS + (t >> 5)] >> (t & 31)) & 1u;
  const unsigned long long bal = __ballot(bit);
  if (lane == 0) wsum[wv] = __popcll(bal);
  __syncthreads();
  int before = 0, total = 0;
  for (int w = 0; w < 16; ++w) {
    if (w < wv) before += wsum[w];
    total += wsum[w];
  }
  const int pos = before + __popcll(bal & ((1ull << lane) - 1ull));
  if (bit && pos < kMaxCombo) P[kP_COMBO + pos] = t;
  const int nl = P[kP_NLIST];
  const bool corr = total <= kMaxCombo && nl <= kPartEdges;
  if (t == 0) {
    P[kP_MODE] = corr ? total : kMaxCombo + 1;
    atomicOr(edge + (size_t)parts * kP_STRIDE + (corr ? kG_CORR : kG_FULL), 1);
  }
  if (!corr) return;
  __syncthreads();  // combo codes written
  for (int i = t; i < nl; i += blockDim.x) {
    const int2 ev = elist[(size_t)part * kPartEdges + i];
    int ci = 0;
    for (int c = 0; c < total; ++c)
      if (P[kP_COMBO + c] == ev.y) ci = c;
    for (int j = 0; j < d1; ++j) {
      const int of = orow[offset[(size_t)ev.x * d1 + j]];
      const unsigned long long v = (unsigned long long)llrint((double)bary[(size_t)ev.x * d1 + j] * kDFix);
      atomicAdd(dfix + (size_t)of * kMaxCombo + ci, v);
    }
  }
}

// correction mode: dv = dfix as float, and dfix cleared for the next call
__global__ void __launch_bounds__(kBlock) k_odconv(const int* __restrict__ edge, int parts, size_t n,
                                                   unsigned long long* __restrict__ dfix, float* __restrict__ dv) {
  if (!edge[(size_t)parts * kP_STRIDE + kG_CORR]) return;
  for (size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (size_t)gridDim.x * kBlock) {
    const unsigned long long v = dfix[i];
    dv[i] = (float)((double)v * (1.0 / kDFix));
    if (v) dfix[i] = 0;
  }
}

// lower slice then upper slice (:638-650, :686-696), then out (+)= scale*x.
// Every vertex's loads are issued before the sums, which keep the
// reference's order (all lower terms, then all upper terms).  TAG 2 and 5
// fix the vertex count at compile time; TAG 0 reads it from d1.
template <int TAG, class T>
__global__ void __launch_bounds__(kBlock) k_oslice(const uint8_t* __restrict__ bins, int L, int LP, int RS, RowMap rm,
                                                   int N, int d1, const float4* __restrict__ vtx,
                                                   const T* __restrict__ lo_v, const T* __restrict__ up_v,
                                                   const float* __restrict__ dv, EdgeMap edge, float alpha,
                                                   SliceArgs o) {
  constexpr bool lmaj = label_major(TAG);
  constexpr int kD1 = TAG == 2 ? 3 : TAG == 5 ? 6 : 0;
  constexpr int NV = kD1 ? kD1 : kMaxD + 1;
  const int nv = kD1 ? kD1 : d1;
  long pix;
  int k0, ks;
  rm.at(L, pix, k0, ks);
  if (pix < 0 || pix >= N) return;
  const SlicePre pre = slice_pre(o, (int)pix, k0, L);
  const int bb = bins[(size_t)pix * LP + k0];
  float4 vt[NV];
  if (L >= nv) {  // lane j of the pixel's label group loads vertex j's record; the group shares them by shuffles
    float4 mine = make_float4(0.f, 0.f, 0.f, 0.f);
    if (k0 < nv) mine = vtx[(size_t)pix * nv + k0];
    const int g0 = (int)(threadIdx.x & 63) - k0;
#pragma unroll
    for (int j = 0; j < NV; ++j)
      if (j < nv) vt[j] = make_float4(__shfl(mine.x, g0 + j), __shfl(mine.y, g0 + j), __shfl(mine.z, g0 + j), 0.f);
  } else {
#pragma unroll
    for (int j = 0; j < NV; ++j)
      if (j < nv) vt[j] = vtx[(size_t)pix * nv + j];
  }
  const int part = edge.ppart(pix);
  const int mode = edge.mode(part);
  const int lo = bb & 15, up = bb >> 4;
  float a[NV], u[NV];
  const bool full = mode > kMaxCombo;
  if constexpr (lmaj) {
    // Label-major rows: LO[b] (b = 1..9) of label k at k * kOBins + b - 1.
    // The slice needs LO[lo] (bin 0 is Tb) and LO[up + 1] (bin 10 is 0), and
    // up is lo or lo - 1, so one pair of adjacent slots LO[i0 + 1], LO[i0 + 2]
    // holds both: one 8-byte load per vertex.
    const int i0 = min(max(lo, 1), 8) - 1;
    const size_t ip = (size_t)k0 * kOBins + i0;
    const bool a_hi = lo - 1 - i0 == 1;  // LO[lo] is the pair's second slot
    const bool u_hi = up - i0 == 1;      // LO[up + 1] is the pair's second slot
#pragma unroll
    for (int j = 0; j < NV; ++j)
      if (j < nv) {
        const float2 pr = ld_pair(lo_v, (size_t)__float_as_int(vt[j].x) * RS + ip);
        a[j] = a_hi ? pr.y : pr.x;
        u[j] = u_hi ? pr.y : pr.x;
      }
    if (full && up < 9)  // the upper rows were splatted and blurred in full: UP[up] (bins 0..8)
#pragma unroll
      for (int j = 0; j < NV; ++j)
        if (j < nv) u[j] = ld_v(up_v, (size_t)__float_as_int(vt[j].x) * RS + (size_t)k0 * kOBins + up);
  } else {
    // Bin-major rows: LO[b] of label k at (b - 1) * L + k (UP[b] at b * L + k).
    // A lane loads only what it uses: most labels have lo = 0 and read just
    // their bin-1 slot, one run of L values per row.
    const int ilo = (max(lo, 1) - 1) * L + k0;  // LO[lo]
    const int iup = min(up, 8) * L + k0;        // LO[up + 1], or UP[up]
#pragma unroll
    for (int j = 0; j < NV; ++j) {
      a[j] = 0.f;
      u[j] = 0.f;
    }
    if (lo > 0)
#pragma unroll
      for (int j = 0; j < NV; ++j)
        if (j < nv) a[j] = ld_v(lo_v, (size_t)__float_as_int(vt[j].x) * RS + ilo);
    if (up < 9)
#pragma unroll
      for (int j = 0; j < NV; ++j)
        if (j < nv) u[j] = ld_v(lo_v, (size_t)__float_as_int(vt[j].x) * RS + iup);
    if (full)  // the upper rows were splatted and blurred in full: UP[up]
#pragma unroll
      for (int j = 0; j < NV; ++j)
        if (j < nv) u[j] = ld_v(up_v, (size_t)__float_as_int(vt[j].x) * RS + iup);
  }
  int ci = -1;  // correction row of combo (k0, up + 1), if any
  if (mode >= 1 && mode <= kMaxCombo && up < 9)
    for (int c = 0; c < mode; ++c)
      if (edge.combo(part, c) == k0 * 16 + up + 1) ci = c;
  float lv[NV], uv[NV];
#pragma unroll
  for (int j = 0; j < NV; ++j) {
    if (j < nv) {
      const float t = vt[j].z;
      lv[j] = lo == 0 ? t : a[j];
      uv[j] = full ? (up == 9 ? t : u[j]) : t - (up < 9 ? u[j] : 0.f);
      if (ci >= 0) uv[j] = uv[j] + dv[(size_t)__float_as_int(vt[j].x) * kMaxCombo + ci];
    }
  }
  float acc = 0.f;
#pragma unroll
  for (int j = 0; j < NV; ++j)
    if (j < nv) acc = acc + lv[j] * (-vt[j].y * alpha);
#pragma unroll
  for (int j = 0; j < NV; ++j)
    if (j < nv) acc = acc + uv[j] * (vt[j].y * alpha);
  slice_store_pre(o, (int)pix, k0, L, acc, pre);
}

// values per lattice point row of the ordered state (bins 1..9 or 0..8 of L
// labels), padded to a multiple of 4 for the blur's quads
inline int orow_stride(int L) { return (kOBins * L + 3) & ~3; }
// bytes per pixel row of the bins (a multiple of 8, read as uint2)
inline int obin_stride(int L) { return (L + 7) & ~7; }

__global__ void k_zero_int(int* p) { *p = 0; }
__global__ void k_zero_ints(int* p, int n) {
  for (int i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0;
}

// n floats from each of gridDim.y pointers set to zero (one launch for the
// zero rows of several buffers)
struct ZeroRows {
  float* p[4];
};
__global__ void __launch_bounds__(kBlock) k_zero_rows(ZeroRows z, size_t n) {
  const size_t t = (size_t)blockIdx.x * kBlock + threadIdx.x;
  if (t < n) z.p[blockIdx.y][t] = 0.f;
}


float lattice_alpha(int d) { return 1.0f / (1 + powf(2.0f, (float)-d)); }

// ---- host-side build, templated on the feature dimension ---------------------
struct BuildOut {
  DBuf<int>* offset;
  DBuf<int>* rank;
  DBuf<float>* bary;
  DBuf<int16_t>* keys;
  DBuf<int>* nbr;
  int M;
};

void scan_excl(const int* in, int* out, int n, hipStream_t s) {
  size_t tmp = 0;
  DCRF_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp, in, out, n, s));
  DBuf<char> t(tmp + 1);
  DCRF_HIP(hipcub::DeviceScan::ExclusiveSum(t.get(), tmp, in, out, n, s));
  DCRF_HIP(hipStreamSynchronize(s));  // t is released on return
}

template <int D>
int build_tables(const float* feat, int N, int Np, bool scalar, hipStream_t s, DBuf<int>& offset, DBuf<int>& rank,
                 DBuf<float>& bary, DBuf<int16_t>& keys, DBuf<int>& nbr) {
  constexpr int D1 = D + 1;
  const int E = Np * D1;
  ElevConst c{};
  const float inv_std = (float)(std::sqrt(2.0 / 3.0) * D1);  // :191
  for (int i = 0; i < D; ++i) c.scale[i] = (float)(1.0 / std::sqrt((double)((i + 2) * (i + 1))) * (double)inv_std);
  c.invdp1 = 1.0f / D1;
  c.dp1 = (float)D1;

  DBuf<PKey<D>> ekeys(E);
  offset.resize(E);
  rank.resize(E);
  bary.resize(E);
  if (scalar)
    k_elevate<D, true><<<grid_for(Np), kBlock, 0, s>>>(feat, N, Np, c, ekeys.get(), rank.get(), bary.get());
  else
    k_elevate<D, false><<<grid_for(Np), kBlock, 0, s>>>(feat, N, Np, c, ekeys.get(), rank.get(), bary.get());
  check_launch("k_elevate");

  uint32_t cap = 1024;
  while (cap < 2u * (uint32_t)E) cap <<= 1;
  DBuf<int> table(cap), slot_of(E), flag(E), pos(E), err(1);
  DCRF_HIP(hipMemsetAsync(table.get(), 0xff, (size_t)cap * sizeof(int), s));
  DCRF_HIP(hipMemsetAsync(err.get(), 0, sizeof(int), s));
  k_insert<D><<<grid_for(E), kBlock, 0, s>>>(ekeys.get(), E, table.get(), cap - 1, slot_of.get(), err.get());
  check_launch("k_insert");
  k_mark_first<<<grid_for(E), kBlock, 0, s>>>(table.get(), slot_of.get(), E, flag.get());
  check_launch("k_mark_first");
  scan_excl(flag.get(), pos.get(), E, s);
  int last_pos = 0, last_flag = 0, herr = 0;
  DCRF_HIP(hipMemcpyAsync(&last_pos, pos.get() + E - 1, sizeof(int), hipMemcpyDeviceToHost, s));
  DCRF_HIP(hipMemcpyAsync(&last_flag, flag.get() + E - 1, sizeof(int), hipMemcpyDeviceToHost, s));
  DCRF_HIP(hipMemcpyAsync(&herr, err.get(), sizeof(int), hipMemcpyDeviceToHost, s));
  DCRF_HIP(hipStreamSynchronize(s));
  if (herr) throw Error(kInternal, "lattice hash table overflow");
  const int M = last_pos + last_flag;

  DBuf<PKey<D>> ukeys(M);
  k_assign<D><<<grid_for(E), kBlock, 0, s>>>(table.get(), slot_of.get(), pos.get(), flag.get(), ekeys.get(), E,
                                             offset.get(), ukeys.get());
  check_launch("k_assign");
  keys.resize((size_t)M * D);
  nbr.resize((size_t)D1 * M * 2);
  k_neighbors<D><<<grid_for((size_t)M * D1), kBlock, 0, s>>>(ukeys.get(), M, table.get(), cap - 1, ekeys.get(),
                                                              pos.get(), nbr.get(), keys.get());
  check_launch("k_neighbors");
  DCRF_HIP(hipStreamSynchronize(s));
  return M;
}

using BuildFn = int (*)(const float*, int, int, bool, hipStream_t, DBuf<int>&, DBuf<int>&, DBuf<float>&,
                        DBuf<int16_t>&, DBuf<int>&);
BuildFn build_fn(int d) {
  switch (d) {
    case 1: return build_tables<1>;
    case 2: return build_tables<2>;
    case 3: return build_tables<3>;
    case 4: return build_tables<4>;
    case 5: return build_tables<5>;
    case 6: return build_tables<6>;
    case 7: return build_tables<7>;
    case 8: return build_tables<8>;
    case 9: return build_tables<9>;
    case 10: return build_tables<10>;
    case 11: return build_tables<11>;
    case 12: return build_tables<12>;
    default: return nullptr;
  }
}

}  // namespace

std::atomic<int>& reference_path() {
  static std::atomic<int> path{0};
  return path;
}

Lattice::Lattice(hipStream_t s) : stream_(s) {}
Lattice::~Lattice() = default;

void Lattice::build(const float* d_feat, int N, int d) {
  DCRF_REQUIRE(N > 0, "lattice: N must be positive");
  BuildFn fn = build_fn(d);
  DCRF_REQUIRE(fn != nullptr, "lattice: feature dimension must be in [1, 12]");
  N_ = N;
  d_ = d;
  scalar_ = reference_path().load() != 0;
  Np_ = scalar_ ? N : (N + 3) / 4 * 4;  // the SSE build pads to blocks of four pixels
  const int d1 = d + 1;
  hipStream_t s = stream_;
  M_ = fn(d_feat, N, Np_, scalar_, s, offset_, rank_, bary_, keys_, nbr_);

  // splat gather lists: entries of real pixels sorted by (point, pixel)
  const int E = N * d1, M = M_;
  DBuf<int> cnt(M + 1), iota(E), ids_sorted(E), ent_sorted(E);
  DCRF_HIP(hipMemsetAsync(cnt.get(), 0, (M + 1) * sizeof(int), s));
  k_count<<<grid_for(E), kBlock, 0, s>>>(offset_.get(), E, cnt.get());
  check_launch("k_count");
  seg_.resize(M + 1);
  scan_excl(cnt.get(), seg_.get(), M + 1, s);
  k_iota<<<grid_for(E), kBlock, 0, s>>>(iota.get(), E);
  check_launch("k_iota");
  int bits = 1;
  while ((1 << bits) <= M) ++bits;
  {
    size_t tmp = 0;
    DCRF_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp, offset_.get(), ids_sorted.get(), iota.get(),
                                                ent_sorted.get(), E, 0, bits, s));
    DBuf<char> t(tmp + 1);
    DCRF_HIP(hipcub::DeviceRadixSort::SortPairs(t.get(), tmp, offset_.get(), ids_sorted.get(), iota.get(),
                                                ent_sorted.get(), E, 0, bits, s));
    DCRF_HIP(hipStreamSynchronize(s));
  }
  sp_pix_.resize(E + kSpPad);  // padding read by the splats' last batch: pixel 0, weight 0
  sp_w_.resize(E + kSpPad);
  k_fill_sp<<<grid_for(E + kSpPad), kBlock, 0, s>>>(ent_sorted.get(), bary_.get(), E, d1, sp_pix_.get(),
                                                     sp_w_.get());
  check_launch("k_fill_sp");

  DBuf<int> nch(M + 1), mflag(M + 1), mpos(M + 1), pn(M + 1);
  DCRF_HIP(hipMemsetAsync(nch.get(), 0, (M + 1) * sizeof(int), s));
  DCRF_HIP(hipMemsetAsync(mflag.get(), 0, (M + 1) * sizeof(int), s));
  DCRF_HIP(hipMemsetAsync(pn.get(), 0, (M + 1) * sizeof(int), s));
  k_nchunks<<<grid_for(M), kBlock, 0, s>>>(cnt.get(), M, nch.get(), mflag.get(), pn.get());
  check_launch("k_nchunks");
  chstart_.resize(M + 1);
  pstart_.resize(M + 1);
  scan_excl(nch.get(), chstart_.get(), M + 1, s);
  scan_excl(mflag.get(), mpos.get(), M + 1, s);
  scan_excl(pn.get(), pstart_.get(), M + 1, s);
  DCRF_HIP(hipMemcpyAsync(&nchunks_, chstart_.get() + M, sizeof(int), hipMemcpyDeviceToHost, s));
  DCRF_HIP(hipMemcpyAsync(&nmulti_, mpos.get() + M, sizeof(int), hipMemcpyDeviceToHost, s));
  DCRF_HIP(hipMemcpyAsync(&npart_, pstart_.get() + M, sizeof(int), hipMemcpyDeviceToHost, s));
  DCRF_HIP(hipStreamSynchronize(s));
  chpoint_.resize(nchunks_);
  multi_.resize(nmulti_ > 0 ? nmulti_ : 1);
  k_chpoint<<<grid_for(M), kBlock, 0, s>>>(chstart_.get(), mpos.get(), mflag.get(), M, chpoint_.get(), multi_.get());
  check_launch("k_chpoint");
  DCRF_HIP(hipStreamSynchronize(s));
}

void Lattice::ensure_values(size_t floats_per_row, bool zero_row0) {
  const size_t need = (size_t)(M_ + 1) * floats_per_row;
  if (need > vcap_) {
    vA_.resize(need);
    vB_.resize(need);
    vcap_ = need;
  }
  // row 0 stands for absent neighbours and must read as zero; rows 1..M are
  // fully rewritten by every splat.  A kernel (not a memset) keeps this
  // capturable in a HIP graph.
  if (zero_row0) {
    k_zero_rows<<<dim3(grid_for(floats_per_row), 2), kBlock, 0, stream_>>>(
        ZeroRows{{vA_.get(), vB_.get(), nullptr, nullptr}}, floats_per_row);
    check_launch("k_zero_rows");
  }
  if (nmulti_ > 0) part_.resize((size_t)npart_ * floats_per_row);
}

void Lattice::filter(const SplatArgs& in, const SliceArgs& out, int V, bool reverse) {
  DCRF_REQUIRE(V > 0, "filter: value size must be positive");
  ensure_values(V);
  hipStream_t s = stream_;
  const RowMap rm = rowmap(V);
  const std::string tag = prof_on() ? "[d=" + std::to_string(d_) + ",V=" + std::to_string(V) + "]" : std::string();
  const bool seq = V <= 2 || scalar_;  // seqCompute; every value size in the portable build
  const float* src = splat_blur(in, V, reverse, seq, tag);
  const float alpha = lattice_alpha(d_);
  KTimer tm("slice" + tag, s);
  if (!seq && V >= kStageMinV && V <= 64) {
    const long nl = (long)N_ * ((V + 3) / 4);
    if (nl > 0)
      k_slice_quad<<<grid_for((size_t)nl), kBlock, 0, s>>>(V, nl, N_, d_ + 1, offset_.get(), bary_.get(), src, alpha,
                                                           out);
  } else if (V >= d_ + 1 && V <= 64) {
    if (seq)
      k_slice_staged<true><<<rows_grid(N_, V), kBlock, 0, s>>>(V, rm, N_, d_ + 1, offset_.get(), bary_.get(), src,
                                                               alpha, out);
    else
      k_slice_staged<false><<<rows_grid(N_, V), kBlock, 0, s>>>(V, rm, N_, d_ + 1, offset_.get(), bary_.get(), src,
                                                                alpha, out);
  } else if (seq)
    k_slice<true><<<rows_grid(N_, V), kBlock, 0, s>>>(V, rm, N_, d_ + 1, offset_.get(), bary_.get(), src, alpha, out);
  else
    k_slice<false><<<rows_grid(N_, V), kBlock, 0, s>>>(V, rm, N_, d_ + 1, offset_.get(), bary_.get(), src, alpha, out);
  check_launch("k_slice");
}

// splat of `in` into the value rows and the d+1 blur axes (reversed order
// when `reverse`; `seq`: the blur sums in double and rounds once, as
// seqCompute does); returns the rows holding the result
const float* Lattice::splat_blur(const SplatArgs& in, int V, bool reverse, bool seq, const std::string& tag) {
  hipStream_t s = stream_;
  const RowMap rm = rowmap(V);
  float* src = vA_.get();
  float* dst = vB_.get();
  {
  KTimer tm("splat" + tag, s);
  const bool stage = V >= kStageMinV && V <= 64;
  switch (in.kind) {
#define DCRF_SPLAT(K)                                                                                             \
  case K:                                                                                                         \
    if (stage)                                                                                                    \
      k_splat_quad<K><<<grid_for((size_t)nchunks_ * ((V + 3) / 4), kBlock, 16384), kBlock, 0, s>>>(             \
          in, V, (long)nchunks_ * ((V + 3) / 4), sp_pix_.get(), sp_w_.get(), seg_.get(), chstart_.get(),         \
          chpoint_.get(), pstart_.get(), src, part_.get());                                                      \
    else                                                                                                          \
      k_splat<K><<<rows_grid(nchunks_, V), kBlock, 0, s>>>(in, V, rm, sp_pix_.get(), sp_w_.get(), seg_.get(),   \
                                                           chstart_.get(), chpoint_.get(), pstart_.get(),        \
                                                           nchunks_, src, part_.get());                          \
    break;
    DCRF_SPLAT(SplatIn::kPlain)
    DCRF_SPLAT(SplatIn::kScaled)
    DCRF_SPLAT(SplatIn::kSxScaled)
    DCRF_SPLAT(SplatIn::kNotOnehot)
    DCRF_SPLAT(SplatIn::kSx)
#undef DCRF_SPLAT
  }
  check_launch("k_splat");
  if (nmulti_ > 0) {
    k_combine<<<rows_grid(nmulti_, V), kBlock, 0, s>>>(V, rm, multi_.get(), nmulti_, chstart_.get(), pstart_.get(),
                                                       part_.get(), src);
    check_launch("k_combine");
  }
  }
  const long bl = (long)M_ * ((V + 3) / 4);  // quad-blur lanes
  DCRF_REQUIRE(bl < (1L << 31), "filter: lattice too large for one blur launch");
  for (int jj = 0; jj <= d_; ++jj) {
    const int j = reverse ? d_ - jj : jj;
    const int* nb = nbr_.get() + (size_t)j * M_ * 2;
    KTimer tm("blur" + tag, s);
    if (seq)
      k_blur<true><<<rows_grid(M_, V), kBlock, 0, s>>>(V, rm, M_, nb, src, dst);
    else if (V >= kStageMinV && bl > 0)
      k_blur_quad<<<grid_for((size_t)bl, kBlock, 16384), kBlock, 0, s>>>(V, FastDiv((uint32_t)((V + 3) / 4)),
                                                                         (uint32_t)bl, nb, src, dst);
    else
      k_blur<false><<<rows_grid(M_, V), kBlock, 0, s>>>(V, rm, M_, nb, src, dst);
    check_launch("k_blur");
    std::swap(src, dst);
  }
  return src;
}

// total splatted weight per point, blurred with the ordered filter's float
// arithmetic (:631); the upper pass of the ordered filter is T - LO[c+1]
void Lattice::ensure_total_weight() {
  if (have_tb_) return;
  hipStream_t s = stream_;
  ensure_values(1);
  DBuf<float> ones(N_);
  k_fill_f<<<grid_for(N_), kBlock, 0, s>>>(ones.get(), (size_t)N_, 1.f);
  SplatArgs a;
  a.kind = SplatIn::kPlain;
  a.in = ones.get();
  const RowMap rm = rowmap(1);
  float* src = vA_.get();
  float* dst = vB_.get();
  k_splat<SplatIn::kPlain><<<rows_grid(nchunks_, 1), kBlock, 0, s>>>(a, 1, rm, sp_pix_.get(), sp_w_.get(), seg_.get(),
                                                                     chstart_.get(), chpoint_.get(), pstart_.get(),
                                                                     nchunks_, src, part_.get());
  if (nmulti_ > 0)
    k_combine<<<rows_grid(nmulti_, 1), kBlock, 0, s>>>(1, rm, multi_.get(), nmulti_, chstart_.get(), pstart_.get(),
                                                       part_.get(), src);
  for (int j = 0; j <= d_; ++j) {
    k_blur<false><<<rows_grid(M_, 1), kBlock, 0, s>>>(1, rm, M_, nbr_.get() + (size_t)j * M_ * 2, src, dst);
    std::swap(src, dst);
  }
  tb_.resize(M_ + 1);
  DCRF_HIP(hipMemcpyAsync(tb_.get(), src, (M_ + 1) * sizeof(float), hipMemcpyDeviceToDevice, s));
  edge_.resize(edge_ints(nparts_));
  DCRF_HIP(hipStreamSynchronize(s));
  check_launch("total weight");
  have_tb_ = true;
}

// static vertex records of the ordered slice, built at the lattice's first
// ordered call (after the total weight)
void Lattice::ensure_ordered_tables() {
  if (have_otab_) return;
  hipStream_t s = stream_;
  // row order: each part's points by the Morton code of their keys.  One
  // device-wide radix sort by code; for a union lattice a second, stable sort
  // by part then groups the parts, keeping Morton order inside each (a
  // segmented sort gives each part one block: tens of ms at 1080p)
  {
    const int M = M_;
    DBuf<unsigned long long> code(std::max(M, 1)), code_s(std::max(M, 1));
    DBuf<int> ids(std::max(M, 1)), order(std::max(M, 1));
    k_morton<<<grid_for(M), kBlock, 0, s>>>(keys_.get(), M, d_, code.get(), ids.get());
    const int code_bits = std::min(64, std::min(64 / d_, 30) * d_);
    size_t tmp = 0;
    DCRF_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp, code.get(), code_s.get(), ids.get(), order.get(), M, 0,
                                                code_bits, s));
    const int nseg = part_pt_off_.empty() ? 1 : (int)part_pt_off_.size() - 1;
    int part_bits = 0;
    while ((1 << part_bits) < nseg) ++part_bits;
    DBuf<uint16_t> pk, pk_s;
    DBuf<int> order2;
    size_t tmp2 = 0;
    if (nseg > 1) {
      pk.resize(M);
      pk_s.resize(M);
      order2.resize(M);
      DCRF_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp2, pk.get(), pk_s.get(), order.get(), order2.get(), M,
                                                  0, part_bits, s));
    }
    DBuf<char> t(std::max(tmp, tmp2) + 1);
    DCRF_HIP(hipcub::DeviceRadixSort::SortPairs(t.get(), tmp, code.get(), code_s.get(), ids.get(), order.get(), M, 0,
                                                code_bits, s));
    if (nseg > 1) {
      k_gather_u16<<<grid_for(M), kBlock, 0, s>>>(ppimg_.get(), order.get(), M, pk.get());
      DCRF_HIP(hipcub::DeviceRadixSort::SortPairs(t.get(), tmp2, pk.get(), pk_s.get(), order.get(), order2.get(), M,
                                                  0, part_bits, s));
      std::swap(order, order2);
    }
    orow_.resize(std::max(M, 1));
    if (ppimg_.n) oppart_.resize(std::max(M, 1));
    k_orow<<<grid_for(M), kBlock, 0, s>>>(order.get(), M, ppimg_.n ? ppimg_.get() : nullptr, orow_.get(),
                                          oppart_.n ? oppart_.get() : nullptr);
    onbr_.resize((size_t)(d_ + 1) * M * 2 + 1);
    k_onbr<<<grid_for((size_t)M * (d_ + 1)), kBlock, 0, s>>>(nbr_.get(), order.get(), orow_.get(), M, d_ + 1,
                                                             onbr_.get());
    check_launch("ordered row order");
    DCRF_HIP(hipStreamSynchronize(s));  // temporaries are released on return
  }
  const long E = (long)N_ * (d_ + 1);
  vtx_.resize(std::max<long>(E, 1));
  k_ovtx<<<grid_for(E), kBlock, 0, s>>>(offset_.get(), bary_.get(), tb_.get(), orow_.get(), E, vtx_.get());
  check_launch("k_ovtx");
  // bin-edge correction: the edge list, fixed-point sums (kept at zero between
  // calls) and two float row buffers for the blur (row 0 stays zero)
  elist_.resize((size_t)nparts_ * kPartEdges);
  const size_t nd = (size_t)(M_ + 1) * kMaxCombo;
  dfix_.resize(nd);
  dvA_.resize(nd);
  dvB_.resize(nd);
  DCRF_HIP(hipMemsetAsync(dfix_.get(), 0, nd * sizeof(unsigned long long), s));
  DCRF_HIP(hipMemsetAsync(dvA_.get(), 0, nd * sizeof(float), s));
  DCRF_HIP(hipMemsetAsync(dvB_.get(), 0, nd * sizeof(float), s));
  DCRF_HIP(hipStreamSynchronize(s));
  have_otab_ = true;
}

// Float rows take the coded filter (ordered_coded.hip), whose results equal
// the dense rows' bit for bit; DCRF_ORDERED_DENSE=1 selects the dense rows.
// bfloat16 rows are dense.
void Lattice::ordered(const float* Q, const float* d_minmax, int L, const SliceArgs& out, OrderedBins* shared,
                      bool fresh) {
  const char* dense = getenv("DCRF_ORDERED_DENSE");
  if (bf16_)
    ordered_t<uint16_t>(Q, d_minmax, L, out, shared, fresh);
  else if (dense && dense[0] == '1')
    ordered_t<float>(Q, d_minmax, L, out, shared, fresh);
  else
    ordered_coded(Q, d_minmax, L, out, shared, fresh);
}

// T: storage of the lattice rows (float, or bfloat16 bits)
template <class T>
void Lattice::ordered_t(const float* Q, const float* d_minmax, int L, const SliceArgs& out, OrderedBins* shared,
                        bool fresh) {
  DCRF_REQUIRE(L > 0, "ordered filter: label count must be positive");
  DCRF_REQUIRE(L <= 64, "ordered filter: at most 64 labels");
  ensure_total_weight();
  ensure_ordered_tables();  // before k_obin, which lists edge values into elist_
  const int row = orow_stride(L);
  ensure_values(2 * (size_t)row, false);  // LO and UP halves of each buffer (float-sized; bf16 rows use half)
  const size_t half = (size_t)(M_ + 1) * row;  // elements of T
  hipStream_t s = stream_;
  const RowMap rm = rowmap(L);
  T* srcL = reinterpret_cast<T*>(vA_.get());
  T* dstL = reinterpret_cast<T*>(vB_.get());
  T* srcU = srcL + half;
  T* dstU = dstL + half;
  // row 0 of each half is the zero row (row * sizeof(T) bytes, a whole number of floats)
  const size_t zf = (size_t)row * sizeof(T) / sizeof(float);
  k_zero_rows<<<dim3(grid_for(zf), 4), kBlock, 0, s>>>(
      ZeroRows{{reinterpret_cast<float*>(srcL), reinterpret_cast<float*>(dstL), reinterpret_cast<float*>(srcU),
                reinterpret_cast<float*>(dstU)}},
      zf);
  const std::string tag = prof_on() ? "[d=" + std::to_string(d_) + ",L=" + std::to_string(L) + "]" : std::string();
  const int LP = obin_stride(L);
  if (shared) {
    shared->bins.resize((size_t)N_ * LP);
    shared->edge.resize(edge_ints(nparts_));
    shared->elist.resize((size_t)nparts_ * kPartEdges);
  } else {
    bins_.resize((size_t)N_ * LP);
  }
  uint8_t* bins = shared ? shared->bins.get() : bins_.get();
  int* edge = shared ? shared->edge.get() : edge_.get();
  int2* elist = shared ? shared->elist.get() : elist_.get();
  const EdgeMap em{edge, pimg_.n ? pimg_.get() : nullptr, ppimg_.n ? ppimg_.get() : nullptr, nparts_};
  // the blur works on rows of the ordered state: row neighbours, the part of each row
  const EdgeMap emr{edge, em.pimg, oppart_.n ? oppart_.get() : nullptr, nparts_};
  if (!shared || fresh) {
    KTimer tm("obin" + tag, s);
    k_zero_ints<<<1, kBlock, 0, s>>>(edge, (int)edge_ints(nparts_));
    const uint32_t nq = (uint32_t)(L + 3) / 4;
    k_obin<<<grid_for((size_t)N_ * nq, kBlock, 16384), kBlock, 0, s>>>(Q, d_minmax, N_, L, FastDiv(nq), bins, LP,
                                                                        edge, pimg_.n ? pimg_.get() : nullptr,
                                                                        elist);
  }
  {  // bin edges: combos and mode (the same for every lattice of a group);
     // in correction mode this lattice's correction rows
    KTimer tm("oedge" + tag, s);
    k_ocombo<<<nparts_, 1024, 0, s>>>(edge, nparts_, d_ + 1, elist, offset_.get(), bary_.get(), orow_.get(),
                                      dfix_.get());
    const size_t nd = (size_t)(M_ + 1) * kMaxCombo;
    k_odconv<<<grid_for(nd, kBlock, 2048), kBlock, 0, s>>>(edge, nparts_, nd, dfix_.get(), dvA_.get());
    check_launch("k_ocombo");
  }
  float* dsrc = dvA_.get();
  float* ddst = dvB_.get();
  {
    KTimer tm("osplat" + tag, s);
    for (int upper = 0; upper < 2; ++upper) {
      T* dst_half = upper ? srcU : srcL;
      const long nlanes = (long)nchunks_ * ((L + 3) / 4);
      const unsigned nvb = grid_for((size_t)nlanes);
      auto launch = [&](auto kern, unsigned grid) {
        kern<<<grid, kBlock, 0, s>>>(bins, L, LP, row, nlanes, sp_pix_.get(), sp_w_.get(), seg_.get(),
                                     chstart_.get(), chpoint_.get(), pstart_.get(), nchunks_, orow_.get(), dst_half,
                                     part_.get(), em);
      };
      auto run = [&](auto kl, auto ku) {
        if (upper) {
          launch(ku, std::min(nvb, 2048u));
        } else {  // the lower pass writes every row: tuned grid of its grid-stride loop
          const unsigned g = Tuner::get().pick(d_ == 5 ? "osplat5" : "osplat", nvb, grid_candidates(nvb), s,
                                               [&](unsigned gg) { launch(kl, gg); });
          launch(kl, g);
        }
      };
      if (d_ == 5)
        run(k_osplat<5, T>, k_osplat<105, T>);
      else if (d_ == 2)
        run(k_osplat<2, T>, k_osplat<102, T>);
      else
        run(k_osplat<0, T>, k_osplat<100, T>);
      if (nmulti_ > 0) {
        const int pv = kOBins * L;  // values of a partial row
        const RowMap rw = rowmap(pv);
        const unsigned g = upper ? std::min(rows_grid(nmulti_, pv), 1024u) : rows_grid(nmulti_, pv);
        k_combine_if<T><<<g, kBlock, 0, s>>>(pv, row, rw, multi_.get(), nmulti_, chstart_.get(), pstart_.get(),
                                             part_.get(), orow_.get(), dst_half, em, upper);
      }
    }
    check_launch("k_osplat");
  }
  const int R4 = row / 4;
  const size_t nquads = (size_t)M_ * R4;
  DCRF_REQUIRE(nquads < (1u << 31), "ordered filter: lattice too large for one blur launch");
  const FastDiv divR4((uint32_t)R4);
  const size_t half_q = half / 4;  // quads from a lower row to its upper row
  // the last axis streams its rows out (non-temporal stores): the slice that
  // reads them then finds no dirty lines of the blur to write back in L2
  auto blur = [&](auto kern, const int* nb, const T* a, T* b, int nt) {
    const unsigned full = grid_for(nquads);
    // src -> dst, the same bits for any grid: tuned on the lower rows
    const unsigned g = Tuner::get().pick(d_ == 5 ? "oblur5" : "oblur", M_, grid_candidates(full), s, [&](unsigned gg) {
      kern<<<gg, kBlock, 0, s>>>(R4, divR4, (uint32_t)nquads, half_q, (uint32_t)M_, nb, a, b, dsrc, ddst, emr, nt);
    });
    kern<<<g, kBlock, 0, s>>>(R4, divR4, (uint32_t)nquads, half_q, (uint32_t)M_, nb, a, b, dsrc, ddst, emr, nt);
  };
  for (int j = 0; j <= d_; ++j) {  // forward order only (:621, :669)
    const int* nb = onbr_.get() + (size_t)j * M_ * 2;
    {
      KTimer tm("oblur" + tag, s);
      const int nt = j == d_;
      if (d_ == 2)
        blur(k_blur2<2, T>, nb, srcL, dstL, nt);
      else if (d_ == 5)
        blur(k_blur2<5, T>, nb, srcL, dstL, nt);
      else
        blur(k_blur2<0, T>, nb, srcL, dstL, nt);
    }
    check_launch("k_blur2");
    std::swap(srcL, dstL);
    std::swap(srcU, dstU);
    std::swap(dsrc, ddst);
  }
  KTimer tm("oslice" + tag, s);
  auto slice = [&](auto kern) {
    kern<<<rows_grid(N_, L), kBlock, 0, s>>>(bins, L, LP, row, rm, N_, d_ + 1, vtx_.get(), srcL, srcU, dsrc, em,
                                             lattice_alpha(d_), out);
  };
  if (d_ == 2)
    slice(k_oslice<2, T>);
  else if (d_ == 5)
    slice(k_oslice<5, T>);
  else
    slice(k_oslice<0, T>);
  check_launch("k_oslice");
}

namespace {
// dst[i] = src[i] + add (add only to entries >= 0 when skip_neg)
__global__ void __launch_bounds__(kBlock) k_copy_add(const int* __restrict__ src, int* __restrict__ dst, size_t n,
                                                     int add, int skip_neg) {
  const size_t t = (size_t)blockIdx.x * kBlock + threadIdx.x;
  if (t >= n) return;
  const int v = src[t];
  dst[t] = (skip_neg && v < 0) ? v : v + add;
}
__global__ void __launch_bounds__(kBlock) k_fill_u16(uint16_t* __restrict__ p, size_t n, uint16_t v) {
  const size_t t = (size_t)blockIdx.x * kBlock + threadIdx.x;
  if (t < n) p[t] = v;
}
}  // namespace

std::unique_ptr<Lattice> Lattice::concat(const std::vector<const Lattice*>& parts, hipStream_t s) {
  DCRF_REQUIRE(!parts.empty() && parts.size() <= 65535, "lattice union: 1..65535 parts");
  const int d = parts[0]->d_, d1 = d + 1;
  long N = 0, M = 0, E = 0, nch = 0, nmul = 0, npart = 0;
  for (const Lattice* p : parts) {
    DCRF_REQUIRE(p->d_ == d, "lattice union: parts differ in feature dimension");
    N += p->N_;
    M += p->M_;
    E += (long)p->N_ * d1;
    nch += p->nchunks_;
    nmul += p->nmulti_;
    npart += p->npart_;
  }
  DCRF_REQUIRE(N * d1 < (1L << 31) && M < (1L << 30), "lattice union: too large for 32-bit indices");
  auto u = std::make_unique<Lattice>(s);
  u->N_ = (int)N;
  u->Np_ = (int)N;
  u->scalar_ = parts[0]->scalar_;
  for (const Lattice* p : parts)
    DCRF_REQUIRE(p->scalar_ == u->scalar_, "lattice union: parts built for different reference builds");
  u->d_ = d;
  u->M_ = (int)M;
  u->nchunks_ = (int)nch;
  u->nmulti_ = (int)nmul;
  u->npart_ = (int)npart;
  u->offset_.resize(E);
  u->bary_.resize(E);
  u->nbr_.resize((size_t)d1 * M * 2);
  u->keys_.resize((size_t)M * d);
  u->part_pt_off_.assign(1, 0);
  u->sp_pix_.resize(E + kSpPad);  // padding read by the splats' last batch: pixel 0, weight 0
  u->sp_w_.resize(E + kSpPad);
  DCRF_HIP(hipMemsetAsync(u->sp_pix_.get() + E, 0, kSpPad * sizeof(int), s));
  DCRF_HIP(hipMemsetAsync(u->sp_w_.get() + E, 0, kSpPad * sizeof(float), s));
  u->seg_.resize(M + 1);
  u->chstart_.resize(M + 1);
  u->pstart_.resize(M + 1);
  u->chpoint_.resize(std::max(nch, 1L));
  u->multi_.resize(std::max(nmul, 1L));
  u->pimg_.resize(N);
  u->ppimg_.resize(M + 1);
  u->nparts_ = (int)parts.size();
  long pb = 0, mb = 0, eb = 0, cb = 0, ub = 0, qb = 0;
  for (size_t b = 0; b < parts.size(); ++b) {
    const Lattice& p = *parts[b];
    const size_t e = (size_t)p.N_ * d1, m = p.M_;
    auto add = [&](const int* src, int* dst, size_t n, long a, int skip) {
      if (n) k_copy_add<<<grid_for(n), kBlock, 0, s>>>(src, dst, n, (int)a, skip);
    };
    add(p.offset_.get(), u->offset_.get() + eb, e, mb, 0);
    DCRF_HIP(hipMemcpyAsync(u->bary_.get() + eb, p.bary_.get(), e * sizeof(float), hipMemcpyDeviceToDevice, s));
    for (int j = 0; j < d1; ++j)
      add(p.nbr_.get() + (size_t)j * m * 2, u->nbr_.get() + ((size_t)j * M + mb) * 2, m * 2, mb, 1);
    add(p.sp_pix_.get(), u->sp_pix_.get() + eb, e, pb, 0);
    DCRF_HIP(hipMemcpyAsync(u->sp_w_.get() + eb, p.sp_w_.get(), e * sizeof(float), hipMemcpyDeviceToDevice, s));
    // per-point starts: M entries each, the union's final entry is written last
    add(p.seg_.get(), u->seg_.get() + mb, m, eb, 0);
    add(p.chstart_.get(), u->chstart_.get() + mb, m, cb, 0);
    add(p.pstart_.get(), u->pstart_.get() + mb, m, qb, 0);
    add(p.chpoint_.get(), u->chpoint_.get() + cb, p.nchunks_, mb, 0);
    if (p.nmulti_ > 0) add(p.multi_.get(), u->multi_.get() + ub, p.nmulti_, mb, 0);
    k_fill_u16<<<grid_for(p.N_), kBlock, 0, s>>>(u->pimg_.get() + pb, p.N_, (uint16_t)b);
    if (m) k_fill_u16<<<grid_for(m), kBlock, 0, s>>>(u->ppimg_.get() + mb, m, (uint16_t)b);
    if (m)
      DCRF_HIP(hipMemcpyAsync(u->keys_.get() + (size_t)mb * d, p.keys_.get(), (size_t)m * d * sizeof(int16_t),
                              hipMemcpyDeviceToDevice, s));
    u->part_pt_off_.push_back((int)(mb + m));
    pb += p.N_;
    mb += m;
    eb += e;
    cb += p.nchunks_;
    ub += p.nmulti_;
    qb += p.npart_;
  }
  const int tail[3] = {(int)E, (int)nch, (int)npart};
  DCRF_HIP(hipMemcpyAsync(u->seg_.get() + M, &tail[0], sizeof(int), hipMemcpyHostToDevice, s));
  DCRF_HIP(hipMemcpyAsync(u->chstart_.get() + M, &tail[1], sizeof(int), hipMemcpyHostToDevice, s));
  DCRF_HIP(hipMemcpyAsync(u->pstart_.get() + M, &tail[2], sizeof(int), hipMemcpyHostToDevice, s));
  check_launch("lattice union");
  DCRF_HIP(hipStreamSynchronize(s));
  return u;
}

void Lattice::export_tables(int* offset, int* rank, float* bary, short* keys, int* nbr) const {
  const size_t E = (size_t)Np_ * (d_ + 1);
  if (offset) DCRF_HIP(hipMemcpyAsync(offset, offset_.get(), E * sizeof(int), hipMemcpyDeviceToHost, stream_));
  if (rank) DCRF_HIP(hipMemcpyAsync(rank, rank_.get(), E * sizeof(int), hipMemcpyDeviceToHost, stream_));
  if (bary) DCRF_HIP(hipMemcpyAsync(bary, bary_.get(), E * sizeof(float), hipMemcpyDeviceToHost, stream_));
  if (keys)
    DCRF_HIP(hipMemcpyAsync(keys, keys_.get(), (size_t)M_ * d_ * sizeof(int16_t), hipMemcpyDeviceToHost, stream_));
  if (nbr)
    DCRF_HIP(hipMemcpyAsync(nbr, nbr_.get(), (size_t)(d_ + 1) * M_ * 2 * sizeof(int), hipMemcpyDeviceToHost, stream_));
  DCRF_HIP(hipStreamSynchronize(stream_));
}

}  // namespace dcrf
