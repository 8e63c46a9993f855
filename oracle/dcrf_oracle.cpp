// CPU oracle for the dense-CRF hot path.  TEST INFRASTRUCTURE ONLY.
//
// This file restates, in plain single-threaded C++, the algorithm of the
// reference (thwjoy/densecrf_ho) for the permutohedral-lattice filter, the
// ordered ("upper minus lower") filter, the Frank-Wolfe QP, the proximal LP and
// the Pn higher-order terms.  It exists to CHECK the HIP engine: only tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.  The
// product path never calls it.
//
// Two instantiations:
//   * float  - follows the reference's float arithmetic step by step (SSE2
//              lattice semantics, Eigen 3.2.7 SSE2 reduction order for the
//              short column sums), so it can be pinned against the compiled
//              reference;
//   * double - the same control flow with every field in FP64 on the same
//              (float-built) lattice structure: the FP64 oracle.
//
// Each function cites the reference file:line it restates.  Pinning status of
// every function is listed in the header comment of tests/test_oracle.py.

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

namespace orc {

// Deliberate mistakes for the tests that show the fixtures pin the oracle
// (orc_set_mutation); 0 = faithful restatement.
//   1 no padding lanes   2 round half away from zero   3 upper bin without 1e-9
//   4 float blur for V<=2   5 QP step without the factor 2   6 last argmin
//   7 superpixel sums over member ids   8 ucondgrad without int truncation
//   9 restricted ordered filter blurring every point
static int g_mut = 0;
// 1: the reference as built where __SSE__ is not defined: the portable init
// (src/permutohedral.cpp:334-486) and seqCompute for every value size
// (:1059-1064); 0: the SSE2 build, compiled on x86-64
static int g_scalar = 0;

// ---------------------------------------------------------------------------
// Reduction order of Eigen 3.2.7 on SSE2 (include/Eigen/src/Core/Redux.h:193-245,
// arch/SSE/PacketMath.h:423-431): two packet accumulators over the aligned
// span, a horizontal add (a0+a2)+(a1+a3), then the unaligned head and tail.
// `astart` is the index of the first 16-byte-aligned element.
// ---------------------------------------------------------------------------
template <class F>
static float eigen_sum_f(int size, int astart, F val) {
  const int ps = 4;
  if (astart > size) astart = size;
  const int as2 = ((size - astart) / (2 * ps)) * (2 * ps);
  const int as1 = ((size - astart) / ps) * ps;
  const int ae2 = astart + as2, ae = astart + as1;
  float res;
  if (as1) {
    float p0[4], p1[4];
    for (int q = 0; q < 4; ++q) p0[q] = val(astart + q);
    if (as1 > ps) {
      for (int q = 0; q < 4; ++q) p1[q] = val(astart + ps + q);
      for (int i = astart + 2 * ps; i < ae2; i += 2 * ps)
        for (int q = 0; q < 4; ++q) {
          p0[q] = p0[q] + val(i + q);
          p1[q] = p1[q] + val(i + ps + q);
        }
      for (int q = 0; q < 4; ++q) p0[q] = p0[q] + p1[q];
      if (ae > ae2)
        for (int q = 0; q < 4; ++q) p0[q] = p0[q] + val(ae2 + q);
    }
    float t0 = p0[0] + p0[2], t1 = p0[1] + p0[3];
    res = t0 + t1;
    for (int i = 0; i < astart; ++i) res = res + val(i);
    for (int i = ae; i < size; ++i) res = res + val(i);
  } else {
    res = val(0);
    for (int i = 1; i < size; ++i) res = res + val(i);
  }
  return res;
}

template <class F>
static double eigen_sum_d(long size, F val) {
  const long ps = 2;
  const long as2 = (size / (2 * ps)) * (2 * ps);
  const long as1 = (size / ps) * ps;
  double res;
  if (size == 0) return 0.0;
  if (as1) {
    double p0[2] = {val(0), val(1)}, p1[2];
    if (as1 > ps) {
      p1[0] = val(2);
      p1[1] = val(3);
      for (long i = 2 * ps; i < as2; i += 2 * ps) {
        p0[0] += val(i);
        p0[1] += val(i + 1);
        p1[0] += val(i + 2);
        p1[1] += val(i + 3);
      }
      p0[0] += p1[0];
      p0[1] += p1[1];
      if (as1 > as2) {
        p0[0] += val(as2);
        p0[1] += val(as2 + 1);
      }
    }
    res = p0[0] + p0[1];
    for (long i = as1; i < size; ++i) res += val(i);
  } else {
    res = val(0);
    for (long i = 1; i < size; ++i) res += val(i);
  }
  return res;
}

// first aligned element of column `col` of an L-row float matrix whose storage
// is 16-byte aligned (include/Eigen/src/Core/util/Memory.h:467-489).
static inline int col_astart(long col, int L) { return (int)((4 - ((col * L) & 3)) & 3); }

// Sum of a column: float mimics Eigen, double is plain.
template <class T> struct ColSum;
template <> struct ColSum<float> {
  static float run(const float* c, int L, long col) {
    return eigen_sum_f(L, std::min(col_astart(col, L), L), [&](int i) { return c[i]; });
  }
  static float run_aligned(const float* c, int L) {
    return eigen_sum_f(L, 0, [&](int i) { return c[i]; });
  }
  static float dot_aligned(const float* a, const float* b, int L) {
    return eigen_sum_f(L, 0, [&](int i) { return a[i] * b[i]; });
  }
};
template <> struct ColSum<double> {
  static double run(const double* c, int L, long) {
    double s = 0;
    for (int i = 0; i < L; ++i) s += c[i];
    return s;
  }
  static double run_aligned(const double* c, int L) { return run(c, L, 0); }
  static double dot_aligned(const double* a, const double* b, int L) {
    double s = 0;
    for (int i = 0; i < L; ++i) s += a[i] * b[i];
    return s;
  }
};

// Sum of a long aligned vector (qp_values.sum(), z.sum()).
static float vec_sum(const float* v, long n) {
  if (n == 0) return 0.f;
  return eigen_sum_f((int)n, 0, [&](int i) { return v[i]; });
}
static double vec_sum(const double* v, long n) {
  double s = 0;
  for (long i = 0; i < n; ++i) s += v[i];
  return s;
}

// dotProduct (src/eigen_utils.cpp:68-74): cast both to double, multiply, sum.
template <class T>
static double dot(const T* a, const T* b, long n) {
  return eigen_sum_d(n, [&](long i) { return (double)a[i] * (double)b[i]; });
}

// ---------------------------------------------------------------------------
// Permutohedral lattice (src/permutohedral.cpp:56-333, SSE2 path).
// ---------------------------------------------------------------------------
struct KeyHash {
  int d = 0;
  std::vector<int16_t> keys;  // id-major
  std::vector<int> table;
  size_t mask = 0;
  int count = 0;
  void reset(int dd, size_t expect) {
    d = dd;
    size_t cap = 16;
    while (cap < 2 * expect + 16) cap <<= 1;
    table.assign(cap, -1);
    mask = cap - 1;
    keys.clear();
    keys.reserve(expect * d);
    count = 0;
  }
  uint64_t h(const int16_t* k) const {
    uint64_t x = 0x9E3779B97F4A7C15ull;
    for (int i = 0; i < d; ++i) {
      x ^= (uint16_t)k[i];
      x *= 0xff51afd7ed558ccdull;
      x ^= x >> 29;
    }
    return x;
  }
  void grow() {
    std::vector<int> nt(table.size() * 2, -1);
    mask = nt.size() - 1;
    for (int e = 0; e < count; ++e) {
      size_t s = h(&keys[(size_t)e * d]) & mask;
      while (nt[s] >= 0) s = (s + 1) & mask;
      nt[s] = e;
    }
    table.swap(nt);
  }
  // ids are assigned in insertion order, as in HashTable::find (:100-128)
  int find(const int16_t* k, bool create) {
    if (create && 2 * (size_t)(count + 1) >= table.size()) grow();
    size_t s = h(k) & mask;
    while (true) {
      int e = table[s];
      if (e < 0) {
        if (!create) return -1;
        keys.insert(keys.end(), k, k + d);
        table[s] = count;
        return count++;
      }
      if (std::memcmp(&keys[(size_t)e * d], k, sizeof(int16_t) * d) == 0) return e;
      s = (s + 1) & mask;
    }
  }
};

struct Lattice {
  int N = 0, d = 0, M = 0, Npad = 0;
  std::vector<int> offset, rank;  // Npad*(d+1)
  std::vector<float> bary;        // Npad*(d+1)
  std::vector<int16_t> keys;      // M*d
  std::vector<int> nbr;           // ((j*M+i)*2 + {0:n1, 1:n2})

  // Permutohedral::init, SSE2 path (src/permutohedral.cpp:152-333).  Pixels are
  // processed in blocks of 4 lanes; a lane past N sees a zero feature vector and
  // still inserts its d+1 keys (:203-208, :273-288).
  void init(const float* feat, int N_, int d_) {
    N = N_;
    d = d_;
    Npad = (g_mut == 1 || g_scalar) ? N : (N + 3) / 4 * 4;  // the portable init has no lanes
    const int d1 = d + 1;
    offset.assign((size_t)Npad * d1, 0);
    rank.assign((size_t)Npad * d1, 0);
    bary.assign((size_t)Npad * d1, 0.f);
    std::vector<int16_t> canonical(d1 * d1);
    for (int i = 0; i <= d; ++i) {  // :183-188
      for (int j = 0; j <= d - i; ++j) canonical[i * d1 + j] = (int16_t)i;
      for (int j = d - i + 1; j <= d; ++j) canonical[i * d1 + j] = (int16_t)(i - d1);
    }
    const float inv_std = (float)(std::sqrt(2.0 / 3.0) * d1);  // :191
    std::vector<float> scale(d);
    for (int i = 0; i < d; ++i)  // :193-194
      scale[i] = (float)(1.0 / std::sqrt((double)((i + 2) * (i + 1))) * (double)inv_std);
    const float invdp1 = 1.0f / d1, dp1 = (float)d1;
    KeyHash ht;
    ht.reset(d, (size_t)N * d1 / 4 + 16);
    std::vector<float> f(d), elev(d1), rem0(d1), rk(d1), bc(d + 2);
    std::vector<int16_t> key(d1);
    for (int k = 0; k < Npad; ++k) {
      for (int j = 0; j < d; ++j) f[j] = k < N ? feat[(size_t)k * d + j] : 0.0f;
      float sm = 0.f;  // elevate (:213-219)
      for (int j = d; j > 0; --j) {
        float cf = f[j - 1] * scale[j - 1];
        float jc = (float)j * cf;
        elev[j] = sm - jc;
        sm = sm + cf;
      }
      elev[0] = sm;
      float sum = 0.f;  // round to the 0-coloured simplex (:222-232)
      if (g_scalar) {
        // portable init (:386-402): the nearer of floor and ceil, ties down;
        // `int sum += rd2 * down_factor` adds in float and truncates
        int isum = 0;
        for (int i = 0; i <= d; ++i) {
          const float v = invdp1 * elev[i];
          const float up = std::ceil(v) * dp1, down = std::floor(v) * dp1;
          const int rd = (up - elev[i] < elev[i] - down) ? (int)(short)up : (int)(short)down;
          rem0[i] = (float)rd;
          isum = (int)((float)isum + (float)rd * invdp1);
        }
        sum = (float)isum;
      } else {
        for (int i = 0; i <= d; ++i) {
          float v = invdp1 * elev[i];
          v = g_mut == 2 ? std::round(v) : std::nearbyint(v);
          rem0[i] = v * dp1;
          sum = sum + v;
        }
      }
      for (int i = 0; i <= d; ++i) rk[i] = 0.f;  // rank (:235-245)
      for (int i = 0; i < d; ++i) {
        float di = elev[i] - rem0[i];
        for (int j = i + 1; j <= d; ++j) {
          float dj = elev[j] - rem0[j];
          float c = (di < dj) ? 1.f : 0.f;
          rk[i] = rk[i] + c;
          rk[j] = rk[j] + (1.f - c);
        }
      }
      for (int i = 0; i <= d; ++i) {  // wrap (:248-254)
        rk[i] = rk[i] + sum;
        float add = rk[i] < 0.f ? dp1 : 0.f;
        float sub = rk[i] >= dp1 ? dp1 : 0.f;
        float as = add - sub;
        rk[i] = rk[i] + as;
        rem0[i] = rem0[i] + as;
      }
      for (int i = 0; i < d + 2; ++i) bc[i] = 0.f;  // barycentric (:257-275)
      for (int i = 0; i <= d; ++i) {
        float v = (elev[i] - rem0[i]) * invdp1;
        int p = (int)((float)d - rk[i]);
        bc[p] = bc[p] + v;
        bc[p + 1] = bc[p + 1] - v;
      }
      if (g_scalar)
        bc[0] = (float)((double)bc[0] + (1.0 + (double)bc[d + 1]));  // :438, a double expression
      else
        bc[0] = bc[0] + (1.f + bc[d + 1]);
      for (int r = 0; r <= d; ++r) {  // keys (:280-287)
        for (int i = 0; i < d; ++i)
          key[i] = (int16_t)(int)(rem0[i] + (float)canonical[r * d1 + (int)rk[i]]);
        offset[(size_t)k * d1 + r] = ht.find(key.data(), true);
        rank[(size_t)k * d1 + r] = (int)rk[r];
        bary[(size_t)k * d1 + r] = bc[r];
      }
    }
    M = ht.count;
    keys = ht.keys;
    nbr.assign((size_t)d1 * M * 2, -1);  // neighbours (:317-330)
    std::vector<int16_t> n1(d1), n2(d1);
    for (int j = 0; j <= d; ++j)
      for (int i = 0; i < M; ++i) {
        const int16_t* kk = &keys[(size_t)i * d];
        for (int q = 0; q < d; ++q) {
          n1[q] = (int16_t)(kk[q] - 1);
          n2[q] = (int16_t)(kk[q] + 1);
        }
        if (j < d) {
          n1[j] = (int16_t)(kk[j] + d);
          n2[j] = (int16_t)(kk[j] - d);
        }
        nbr[((size_t)j * M + i) * 2 + 0] = ht.find(n1.data(), false);
        nbr[((size_t)j * M + i) * 2 + 1] = ht.find(n2.data(), false);
      }
  }

  // Permutohedral::compute (src/permutohedral.cpp:1079-1087): V<=2 takes
  // seqCompute (:488-540), otherwise sseCompute (:998-1058).  Row 0 of the
  // value buffer is the zero pad for absent neighbours.
  void filter_f(const float* in, float* out, int V, bool reverse) const {
    const int d1 = d + 1;
    const bool seq = V <= 2 || g_scalar;  // seqCompute (every value size in the portable build)
    std::vector<float> a((size_t)(M + 2) * V, 0.f), b((size_t)(M + 2) * V, 0.f);
    float* values = a.data();
    float* nv = b.data();
    for (int i = 0; i < N; ++i)
      for (int j = 0; j <= d; ++j) {
        int o = offset[(size_t)i * d1 + j] + 1;
        float w = bary[(size_t)i * d1 + j];
        for (int k = 0; k < V; ++k) values[(size_t)o * V + k] = values[(size_t)o * V + k] + w * in[(size_t)i * V + k];
      }
    for (int jj = 0; jj <= d; ++jj) {
      int j = reverse ? d - jj : jj;
      for (int i = 0; i < M; ++i) {
        const float* ov = values + (size_t)(i + 1) * V;
        float* nw = nv + (size_t)(i + 1) * V;
        int n1 = nbr[((size_t)j * M + i) * 2] + 1, n2 = nbr[((size_t)j * M + i) * 2 + 1] + 1;
        const float* v1 = values + (size_t)n1 * V;
        const float* v2 = values + (size_t)n2 * V;
        for (int k = 0; k < V; ++k) {
          float s = v1[k] + v2[k];
          if (seq && g_mut != 4)
            nw[k] = (float)((double)ov[k] + 0.5 * (double)s);  // :518 (double expression)
          else
            nw[k] = ov[k] + 0.5f * s;  // :1035
        }
      }
      std::swap(values, nv);
    }
    const float alpha = 1.0f / (1 + std::pow(2.0f, (float)-d));
    for (int i = 0; i < N; ++i) {
      float* o_ = out + (size_t)i * V;
      std::vector<float> acc(V, 0.f);
      for (int j = 0; j <= d; ++j) {
        int o = offset[(size_t)i * d1 + j] + 1;
        float w = bary[(size_t)i * d1 + j];
        if (seq) {
          for (int k = 0; k < V; ++k) acc[k] = acc[k] + (w * values[(size_t)o * V + k]) * alpha;  // :533
        } else {
          float wa = w * alpha;  // :1048
          for (int k = 0; k < V; ++k) acc[k] = acc[k] + wa * values[(size_t)o * V + k];
        }
      }
      for (int k = 0; k < V; ++k) o_[k] = acc[k];
    }
  }

  // FP64 filter on the same lattice structure.
  void filter_d(const double* in, double* out, int V, bool reverse) const {
    const int d1 = d + 1;
    std::vector<double> a((size_t)(M + 2) * V, 0.0), b((size_t)(M + 2) * V, 0.0);
    double* values = a.data();
    double* nv = b.data();
    for (int i = 0; i < N; ++i)
      for (int j = 0; j <= d; ++j) {
        int o = offset[(size_t)i * d1 + j] + 1;
        double w = bary[(size_t)i * d1 + j];
        for (int k = 0; k < V; ++k) values[(size_t)o * V + k] += w * in[(size_t)i * V + k];
      }
    for (int jj = 0; jj <= d; ++jj) {
      int j = reverse ? d - jj : jj;
      for (int i = 0; i < M; ++i) {
        int n1 = nbr[((size_t)j * M + i) * 2] + 1, n2 = nbr[((size_t)j * M + i) * 2 + 1] + 1;
        for (int k = 0; k < V; ++k)
          nv[(size_t)(i + 1) * V + k] =
              values[(size_t)(i + 1) * V + k] + 0.5 * (values[(size_t)n1 * V + k] + values[(size_t)n2 * V + k]);
      }
      std::swap(values, nv);
    }
    const double alpha = 1.0 / (1.0 + std::pow(2.0, -d));
    for (int i = 0; i < N; ++i) {
      for (int k = 0; k < V; ++k) out[(size_t)i * V + k] = 0;
      for (int j = 0; j <= d; ++j) {
        int o = offset[(size_t)i * d1 + j] + 1;
        double wa = (double)bary[(size_t)i * d1 + j] * alpha;
        for (int k = 0; k < V; ++k) out[(size_t)i * V + k] += wa * values[(size_t)o * V + k];
      }
    }
  }

  // Bin rules of the ordered filter (src/permutohedral.cpp:548-589):
  // lower bin  = min(floor(float(q*10)), 9)
  // upper bin  = max(floor((double(q) - 1e-9) * 10), 0)
  // Permutohedral::gradient (src/permutohedral.cpp:1095-1178): df (N x d) of
  // a^T K b; splat in float, blur sums in double and rounds once, slice
  // through the rank-ordered vertices in float
  void gradient(const float* a, const float* b, int V, float* df) const {
    const int d1 = d + 1;
    std::vector<float> values((size_t)(M + 2) * V), nv((size_t)(M + 2) * V);
    std::fill(df, df + (size_t)N * d, 0.f);
    std::vector<float> scale(d);
    const float inv_std = (float)(std::sqrt(2.0 / 3.0) * d1);
    for (int i = 0; i < d; ++i) scale[i] = (float)(1.0 / std::sqrt(double((i + 2) * (i + 1))) * inv_std);
    const float alpha = 1.0f / (1 + std::pow(2.0f, (float)-d)) / d1;
    for (int dir = 0; dir < 2; ++dir) {
      std::fill(values.begin(), values.end(), 0.f);
      std::fill(nv.begin(), nv.end(), 0.f);
      const float* x = dir ? b : a;
      for (int i = 0; i < N; ++i)
        for (int j = 0; j <= d; ++j) {
          const int o = offset[(size_t)i * d1 + j] + 1;
          const float w = bary[(size_t)i * d1 + j];
          for (int k = 0; k < V; ++k) values[(size_t)o * V + k] += w * x[(size_t)i * V + k];
        }
      float* cur = values.data();
      float* nxt = nv.data();
      for (int jj = 0; jj <= d; ++jj) {
        const int j = dir ? d - jj : jj;
        for (int i = 0; i < M; ++i) {
          const int n1 = nbr[((size_t)j * M + i) * 2] + 1, n2 = nbr[((size_t)j * M + i) * 2 + 1] + 1;
          for (int k = 0; k < V; ++k)
            nxt[(size_t)(i + 1) * V + k] =
                (float)(cur[(size_t)(i + 1) * V + k] + 0.5 * (cur[(size_t)n1 * V + k] + cur[(size_t)n2 * V + k]));
        }
        std::swap(cur, nxt);
      }
      const float* y = dir ? a : b;
      std::vector<float> ra((size_t)d1 * V), sm(V);
      for (int i = 0; i < N; ++i) {
        std::fill(ra.begin(), ra.end(), 0.f);
        for (int j = 0; j <= d; ++j) {
          const int r0 = d - rank[(size_t)i * d1 + j];
          const int r1 = r0 + 1 > d ? 0 : r0 + 1;
          const int o0 = offset[(size_t)i * d1 + r0] + 1, o1 = offset[(size_t)i * d1 + r1] + 1;
          for (int k = 0; k < V; ++k) {
            ra[(size_t)j * V + k] += alpha * cur[(size_t)o0 * V + k];
            ra[(size_t)j * V + k] -= alpha * cur[(size_t)o1 * V + k];
          }
        }
        for (int k = 0; k < V; ++k) sm[k] = ra[k];
        for (int j = 1; j <= d; ++j) {
          float grad = 0;
          for (int k = 0; k < V; ++k) {
            const float v = scale[j - 1] * (sm[k] - j * ra[(size_t)j * V + k]);
            grad += y[(size_t)i * V + k] * v;
            sm[k] += ra[(size_t)j * V + k];
          }
          df[(size_t)i * d + j - 1] += grad;
        }
      }
    }
  }

  static int bin_lo(float q) { return std::min((int)std::floor(q * 10), 9); }
  static int bin_up(float q) { return std::max((int)std::floor(((double)q - (g_mut == 3 ? 0.0 : 1e-9)) * 10), 0); }

  // seqCompute_upper_minus_lower_ord (src/permutohedral.cpp:590-701); T is the
  // accumulation type, the bins always come from the float q.
  template <class T>
  void ordered(const float* q, T* out, int V) const {
    const int d1 = d + 1, R = 10;
    const T alpha = (T)(1.0f / (1 + std::pow(2.0f, (float)-d)));
    const T alpha_d = std::is_same<T, double>::value ? (T)(1.0 / (1.0 + std::pow(2.0, -d))) : alpha;
    std::vector<T> a((size_t)(M + 2) * V * R), b((size_t)(M + 2) * V * R);
    for (int pass = 0; pass < 2; ++pass) {
      const bool upper = pass == 1;
      std::fill(a.begin(), a.end(), (T)0);
      std::fill(b.begin(), b.end(), (T)0);
      T* values = a.data();
      T* nv = b.data();
      for (int i = 0; i < N; ++i)  // splat (:609-618, :657-666)
        for (int j = 0; j <= d; ++j) {
          int o = offset[(size_t)i * d1 + j] + 1;
          T w = (T)bary[(size_t)i * d1 + j];
          for (int k = 0; k < V; ++k) {
            float qq = q[(size_t)i * V + k];
            T* s = values + ((size_t)o * V + k) * R;
            if (upper) {
              for (int c = bin_up(qq); c < R; ++c) s[c] = s[c] + w;
            } else {
              for (int c = 0, e = bin_lo(qq); c <= e; ++c) s[c] = s[c] + w;
            }
          }
        }
      for (int j = 0; j <= d; ++j) {  // blur (:620-636)
        for (int i = 0; i < M; ++i) {
          int n1 = nbr[((size_t)j * M + i) * 2] + 1, n2 = nbr[((size_t)j * M + i) * 2 + 1] + 1;
          for (int k = 0; k < V; ++k) {
            const T* ov = values + ((size_t)(i + 1) * V + k) * R;
            const T* v1 = values + ((size_t)n1 * V + k) * R;
            const T* v2 = values + ((size_t)n2 * V + k) * R;
            T* nw = nv + ((size_t)(i + 1) * V + k) * R;
            for (int c = 0; c < R; ++c) nw[c] = ov[c] + (T)0.5 * (v1[c] + v2[c]);
          }
        }
        std::swap(values, nv);
      }
      for (int i = 0; i < N; ++i) {  // slice (:638-650, :686-696)
        if (!upper)
          for (int k = 0; k < V; ++k) out[(size_t)i * V + k] = 0;
        for (int j = 0; j <= d; ++j) {
          int o = offset[(size_t)i * d1 + j] + 1;
          float wf = bary[(size_t)i * d1 + j];
          T coef;
          if (std::is_same<T, float>::value)
            coef = upper ? (T)(wf * (float)alpha) : (T)(-wf * (float)alpha);
          else
            coef = upper ? (T)wf * alpha_d : -(T)wf * alpha_d;
          for (int k = 0; k < V; ++k) {
            float qq = q[(size_t)i * V + k];
            int c = upper ? bin_up(qq) : bin_lo(qq);
            out[(size_t)i * V + k] = out[(size_t)i * V + k] + values[((size_t)o * V + k) * R + c] * coef;
          }
        }
      }
    }
  }
};

// seqCompute_upper_minus_lower_ord_restricted (src/permutohedral.cpp:702-880):
// the ordered filter of the pixels pI (ascending) with every other pixel
// splatted from ext (full N x V field); only lattice points touched by pI are
// blurred.  State kept between calls as in the reference object: the splat of
// the pixels outside pI (one per pass) and the active points, both written by
// a call with store.  out (|pI| x V) is added to, as in the reference.
struct Restricted {
  std::vector<float> stored[2];
  std::vector<int> active;
  int V = 0;
  bool have = false;
  void run(const Lattice& lat, const float* rin, const int* pI, int nI, const float* ext, int V_, bool store,
           float* out) {
    const int d = lat.d, d1 = d + 1, M = lat.M, N = lat.N, R = 10;
    if (!store && (!have || V_ != V)) throw std::runtime_error("restricted: no stored values for this value size");
    V = V_;
    const size_t rows = (size_t)(M + 2) * V * R;
    const float alpha = 1.0f / (1 + std::pow(2.0f, (float)-d));
    std::vector<float> a(rows), b(rows);
    std::vector<int> inI(N, -1);
    for (int ii = 0; ii < nI; ++ii) inI[pI[ii]] = ii;
    if (store) {
      std::vector<char> act(M, 0);
      active.clear();
      for (int ii = 0; ii < nI; ++ii)  // :752-758, in pI order
        for (int j = 0; j <= d; ++j) {
          int o = lat.offset[(size_t)pI[ii] * d1 + j];
          if (!act[o]) {
            act[o] = 1;
            active.push_back(o);
          }
        }
    }
    for (int pass = 0; pass < 2; ++pass) {
      const bool upper = pass == 1;
      std::fill(a.begin(), a.end(), 0.f);
      std::fill(b.begin(), b.end(), 0.f);
      if (store) {  // :725-737 / :806-818: the pixels outside pI, from ext
        for (int i = 0; i < N; ++i) {
          if (inI[i] >= 0) continue;
          for (int j = 0; j <= d; ++j) {
            int o = lat.offset[(size_t)i * d1 + j] + 1;
            float w = lat.bary[(size_t)i * d1 + j];
            for (int k = 0; k < V; ++k) {
              float qq = ext[(size_t)i * V + k];
              float* s = &a[((size_t)o * V + k) * R];
              if (upper)
                for (int c = Lattice::bin_up(qq); c < R; ++c) s[c] = s[c] + w;
              else
                for (int c = 0, e = Lattice::bin_lo(qq); c <= e; ++c) s[c] = s[c] + w;
            }
          }
        }
        stored[pass] = a;
      } else {
        a = stored[pass];  // :761-763 / :833-835
      }
      for (int ii = 0; ii < nI; ++ii) {  // the pixels of pI, from rin
        const int i = pI[ii];
        for (int j = 0; j <= d; ++j) {
          int o = lat.offset[(size_t)i * d1 + j] + 1;
          float w = lat.bary[(size_t)i * d1 + j];
          for (int k = 0; k < V; ++k) {
            float qq = rin[(size_t)ii * V + k];
            float* s = &a[((size_t)o * V + k) * R];
            if (upper)
              for (int c = Lattice::bin_up(qq); c < R; ++c) s[c] = s[c] + w;
            else
              for (int c = 0, e = Lattice::bin_lo(qq); c <= e; ++c) s[c] = s[c] + w;
          }
        }
      }
      float* values = a.data();
      float* nv = b.data();
      std::vector<int> all;
      if (g_mut == 9)
        for (int i = 0; i < M; ++i) all.push_back(i);
      for (int j = 0; j <= d; ++j) {  // :778-793: active points only; the others keep what the buffer holds
        for (int i : (g_mut == 9 ? all : active)) {
          int n1 = lat.nbr[((size_t)j * M + i) * 2] + 1, n2 = lat.nbr[((size_t)j * M + i) * 2 + 1] + 1;
          for (int k = 0; k < V; ++k) {
            const float* ov = values + ((size_t)(i + 1) * V + k) * R;
            const float* v1 = values + ((size_t)n1 * V + k) * R;
            const float* v2 = values + ((size_t)n2 * V + k) * R;
            float* nw = nv + ((size_t)(i + 1) * V + k) * R;
            for (int c = 0; c < R; ++c) nw[c] = ov[c] + 0.5f * (v1[c] + v2[c]);
          }
        }
        std::swap(values, nv);
      }
      for (int ii = 0; ii < nI; ++ii) {  // :796-807 / :866-878
        const int i = pI[ii];
        for (int j = 0; j <= d; ++j) {
          int o = lat.offset[(size_t)i * d1 + j] + 1;
          float w = lat.bary[(size_t)i * d1 + j];
          const float coef = upper ? w * alpha : -w * alpha;
          for (int k = 0; k < V; ++k) {
            float qq = rin[(size_t)ii * V + k];
            int c = upper ? Lattice::bin_up(qq) : Lattice::bin_lo(qq);
            out[(size_t)ii * V + k] = out[(size_t)ii * V + k] + values[((size_t)o * V + k) * R + c] * coef;
          }
        }
      }
    }
    if (store) have = true;
  }
};

// Brute-force filters of PairwisePotential (before the compatibility):
// apply_bf (src/pairwise.cpp:368-383) and apply_upper_minus_lower_bf /
// _bf_ord (:294-353).  K(c, b) = exp(-|f_c - f_b|^2): the squared norm of the
// float difference vector in Eigen 3.2.7's SSE2 reduction order for a fresh
// (aligned) VectorXf (include/Eigen/src/Core/Redux.h:193-245, predux of
// arch/SSE/PacketMath.h:423-427), exp in double (the C library's exp of the
// promoted value), sums in float with the double term added before rounding.
static float sq_norm_eigen(const float* t, int n) {
  if (n < 4) {
    float r = t[0];
    for (int i = 1; i < n; ++i) r = r + t[i];
    return r;
  }
  const int a1 = n / 4 * 4, a2 = n / 8 * 8;
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
  float r = (p0[0] + p0[2]) + (p0[1] + p0[3]);
  for (int i = a1; i < n; ++i) r = r + t[i];
  return r;
}
static double bf_kernel(const float* a, const float* b, int d) {
  float t[64];
  for (int k = 0; k < d; ++k) {
    const float x = a[k] - b[k];
    t[k] = x * x;
  }
  return ::exp((double)-sq_norm_eigen(t, d));
}
static void bf_apply(const float* f, int N, int d, const float* Q, int L, float* out) {
  std::fill(out, out + (size_t)N * L, 0.f);
  for (int l = 0; l < L; ++l)
    for (int c = 0; c < N; ++c)
      for (int b = 0; b < N; ++b)
        out[(size_t)c * L + l] = (float)((double)out[(size_t)c * L + l] +
                                         bf_kernel(&f[(size_t)c * d], &f[(size_t)b * d], d) * Q[(size_t)b * L + l]);
}
// ord: split by value Q(l, ind(l, .)) instead of position; out stays sorted
static void bf_uml(const float* f, int N, int d, const int* ind, const float* Q, int L, float* out) {
  for (int l = 0; l < L; ++l)
    for (int c = 0; c < N; ++c) {
      const int pc = ind[(size_t)c * L + l];
      float acc = 0.f;
      for (int pass = 0; pass < 2; ++pass)
        for (int b = 0; b < N; ++b) {
          const int pb = ind[(size_t)b * L + l];
          bool take;
          if (Q) {
            const float qb = Q[(size_t)pb * L + l], qc = Q[(size_t)pc * L + l];
            take = pass == 0 ? !(qb > qc) : !(qb < qc);
          } else {
            take = pass == 0 ? b > c : b < c;
          }
          if (!take) continue;
          const double K = bf_kernel(&f[(size_t)pc * d], &f[(size_t)pb * d], d);
          acc = (float)(pass == 0 ? (double)acc + K : (double)acc - K);
        }
      out[(size_t)c * L + l] = acc;
    }
}

// PairwisePotential::apply_upper_minus_lower_dc (src/pairwise.cpp:241-293)
// before the compatibility: per label, the features in the order ind(l, .),
// a lattice on them, and the divide-and-conquer of
// apply_upper_minus_lower_sorted_slice: blocks of at most 30 pixels by brute
// force (alpha 1/0.6 in double), larger ones split in two halves (sharing the
// middle pixel when odd) plus Permutohedral::compute_upper_minus_lower_dc
// (src/permutohedral.cpp:881-997) between them.  out (N x L) stays in the
// sorted order, as in the reference.
struct DcLattice {
  const Lattice& lat;
  std::vector<float> values, nv;
  std::vector<char> act;
  explicit DcLattice(const Lattice& l) : lat(l), values(l.M + 2), nv(l.M + 2), act(l.M) {}
  // compute_upper_minus_lower_dc(out, low, middle_low, middle_high, high)
  void dc(float* out, int low, int ml, int mh, int high) {
    const int d = lat.d, d1 = d + 1;
    std::fill(values.begin(), values.end(), 0.f);
    std::fill(nv.begin(), nv.end(), 0.f);
    std::fill(act.begin(), act.end(), 0);
    std::vector<int> list;
    for (int i = mh; i < high; ++i)
      for (int j = 0; j <= d; ++j) {
        const int o = lat.offset[(size_t)i * d1 + j];
        values[o + 1] += lat.bary[(size_t)i * d1 + j];
        if (!act[o]) {
          act[o] = 1;
          list.push_back(o);
        }
      }
    for (int i = low; i < ml; ++i)
      for (int j = 0; j <= d; ++j) {
        const int o = lat.offset[(size_t)i * d1 + j];
        if (!act[o]) {
          act[o] = 1;
          list.push_back(o);
        }
      }
    const float alpha = 1.0f / (1 + std::pow(2.0f, (float)-d));
    auto blur = [&]() -> float* {
      float* cur = values.data();
      float* nxt = nv.data();
      for (int j = 0; j <= d; ++j) {
        for (int i : list) {
          const int n1 = lat.nbr[((size_t)j * lat.M + i) * 2], n2 = lat.nbr[((size_t)j * lat.M + i) * 2 + 1];
          nxt[i + 1] = (float)(cur[i + 1] + 0.5 * (cur[n1 + 1] + cur[n2 + 1]));
        }
        std::swap(cur, nxt);
      }
      return cur;
    };
    float* v = blur();
    for (int i = low; i < ml; ++i)
      for (int j = 0; j <= d; ++j) {
        const int o = lat.offset[(size_t)i * d1 + j] + 1;
        out[i] += lat.bary[(size_t)i * d1 + j] * v[o] * alpha;
      }
    std::fill(values.begin(), values.end(), 0.f);
    std::fill(nv.begin(), nv.end(), 0.f);
    for (int i = low; i < ml; ++i)
      for (int j = 0; j <= d; ++j) values[lat.offset[(size_t)i * d1 + j] + 1] += lat.bary[(size_t)i * d1 + j];
    v = blur();
    for (int i = mh; i < high; ++i)
      for (int j = 0; j <= d; ++j) {
        const int o = lat.offset[(size_t)i * d1 + j] + 1;
        out[i] -= lat.bary[(size_t)i * d1 + j] * v[o] * alpha;
      }
  }
};
static void dc_sorted_slice(DcLattice& L, const float* fs, int d, float* out, int mn, int mx) {
  const int size = mx - mn;
  if (size <= 30) {
    const double alpha = 1.0 / 0.6;
    for (int c = mn; c < mx; ++c)
      for (int b = mn; b < c; ++b) out[c] = (float)(out[c] - bf_kernel(&fs[(size_t)c * d], &fs[(size_t)b * d], d) * alpha);
    for (int c = mn; c < mx; ++c)
      for (int b = c + 1; b < mx; ++b)
        out[c] = (float)(out[c] + bf_kernel(&fs[(size_t)c * d], &fs[(size_t)b * d], d) * alpha);
    return;
  }
  int ml, mh;
  if (size % 2 == 0) {
    ml = mh = mn + size / 2;
  } else {
    ml = (int)std::floor(mn + size / 2.0);
    mh = ml + 1;
  }
  dc_sorted_slice(L, fs, d, out, mn, mh);
  dc_sorted_slice(L, fs, d, out, ml, mx);
  L.dc(out, mn, ml, mh, mx);
}
static void uml_dc(const float* f, int N, int d, const int* ind, int Lb, float* out) {
  std::vector<float> fs((size_t)N * d), col(N);
  for (int l = 0; l < Lb; ++l) {
    for (int j = 0; j < N; ++j)
      for (int k = 0; k < d; ++k) fs[(size_t)j * d + k] = f[(size_t)ind[(size_t)j * Lb + l] * d + k];
    Lattice lat;
    lat.init(fs.data(), N, d);
    DcLattice dl(lat);
    std::fill(col.begin(), col.end(), 0.f);
    dc_sorted_slice(dl, fs.data(), d, col.data(), 0, N);
    for (int j = 0; j < N; ++j) out[(size_t)j * Lb + l] = col[j];
  }
}

// ---------------------------------------------------------------------------
// Dense CRF with Potts pairwise terms and Pn superpixel layers.
// ---------------------------------------------------------------------------
struct Pairwise {
  Lattice lat;
  std::vector<float> feat;
  int d = 0;
  float w = 1.f;
  int ntype = 3;               // NormalizationType: 0 none, 1 before, 2 after, 3 symmetric
  std::vector<float> norm_f;   // float norm (DenseKernel::initLattice, src/pairwise.cpp:40-57)
  std::vector<double> norm_d;  // same rounding then widened
  // label compatibility (include/labelcompatibility.h:51-92): 0 Potts (w),
  // 1 Diagonal (cp = v[L]), 2 Matrix (cp = W[L*L], already (m + m^T)/2 as the
  // MatrixCompatibility constructor stores it, src/labelcompatibility.cpp:94-97)
  int ctype = 0;
  std::vector<float> cp;
  // the NO_NORMALIZATION copy always carries PottsCompatibility(parameters()(0))
  // taken when the term is added (src/densecrf.cpp:161-168)
  float nn_w = 1.f;
};

// LabelCompatibility::apply on one pixel's L-vector x -> y
// (src/labelcompatibility.cpp:49-52 Potts, :81-84 Diagonal, :100-102 Matrix;
// W is symmetric, so applyTranspose (:103-105) is the same map)
template <class T>
static void compat_apply(const Pairwise& p, int L, const T* x, T* y) {
  if (p.ctype == 0) {
    for (int l = 0; l < L; ++l) y[l] = -(T)p.w * x[l];
  } else if (p.ctype == 1) {
    for (int l = 0; l < L; ++l) y[l] = (T)p.cp[l] * x[l];
  } else {
    for (int l = 0; l < L; ++l) {
      T s = 0;
      for (int j = 0; j < L; ++j) s += (T)p.cp[(size_t)j * L + l] * x[j];
      y[l] = s;
    }
  }
}

// number of parameters of a compatibility (LabelCompatibility::parameters,
// src/labelcompatibility.cpp:61-65, :88-90, :109-115)
static int compat_nparams(const Pairwise& p, int L) {
  return p.ctype == 0 ? 1 : p.ctype == 1 ? L : L * (L + 1) / 2;
}

// LabelCompatibility::gradient(b, Q) accumulated into g (FP64):
// Potts -sum(b o Q) (:72-76), Diagonal row sums of b o Q (:91-93), Matrix the
// packed upper triangle of b Q^T with off-diagonal entries doubled (:122-129)
template <class T>
static void compat_grad_accum(const Pairwise& p, int N, int L, const T* b, const T* Q, double* g) {
  if (p.ctype == 0) {
    double s = 0;
    for (size_t i = 0; i < (size_t)N * L; ++i) s += (double)b[i] * (double)Q[i];
    g[0] -= s;
  } else if (p.ctype == 1) {
    for (int i = 0; i < N; ++i)
      for (int l = 0; l < L; ++l) g[l] += (double)b[(size_t)i * L + l] * (double)Q[(size_t)i * L + l];
  } else {
    std::vector<double> G((size_t)L * L, 0.0);  // G(a, c) = sum_i b(a,i) Q(c,i), at a*L + c
    for (int i = 0; i < N; ++i)
      for (int a = 0; a < L; ++a)
        for (int c = 0; c < L; ++c) G[(size_t)a * L + c] += (double)b[(size_t)i * L + a] * (double)Q[(size_t)i * L + c];
    for (int i = 0, k = 0; i < L; ++i)
      for (int j = i; j < L; ++j, ++k) g[k] += G[(size_t)i * L + j] + (i != j ? G[(size_t)j * L + i] : 0.0);
  }
}

template <class T>
struct CRF {
  int N, L;
  std::vector<T> U;  // L*N, label-contiguous
  std::vector<Pairwise*> pw;
  // superpixels (global region index across layers, in layer order)
  int R = 0;
  std::vector<std::vector<int>> members;
  std::vector<T> e_r;  // exp_of_superpixels_
  std::vector<T> c_r;  // constants_
  std::vector<T> z_labels;  // L*R after QP

  CRF(int n, int l) : N(n), L(l), U((size_t)n * l, (T)0) {}

  // DenseKernel::filter (src/pairwise.cpp:58-75) + LabelCompatibility::apply
  // (src/pairwise.cpp:203-213).  `no_norm` selects the NO_NORMALIZATION copy
  // built by addPairwiseEnergy (src/densecrf.cpp:164-168).
  void apply(int k, const T* in, T* out, bool no_norm = false, bool transpose = false) const {
    kernel(k, in, out, no_norm, transpose);
    const Pairwise& p = *pw[k];
    if (no_norm) {
      for (size_t i = 0; i < (size_t)N * L; ++i) out[i] = -(T)p.nn_w * out[i];
      return;
    }
    for (int i = 0; i < N; ++i) {
      T y[64];
      compat_apply(p, L, out + (size_t)i * L, y);
      std::copy(y, y + L, out + (size_t)i * L);
    }
  }
  // DenseKernel::apply / applyTranspose alone: the normalised filter without
  // the compatibility (src/pairwise.cpp:76-87)
  void kernel(int k, const T* in, T* out, bool no_norm = false, bool transpose = false) const {
    const Pairwise& p = *pw[k];
    const size_t NL = (size_t)N * L;
    std::vector<T> tmp(NL);
    const int nt = no_norm ? 0 : p.ntype;
    const bool pre = nt == 3 || (nt == 1 && !transpose) || (nt == 2 && transpose);
    const bool post = nt == 3 || (nt == 1 && transpose) || (nt == 2 && !transpose);
    for (int i = 0; i < N; ++i)
      for (int l = 0; l < L; ++l) {
        T x = in[(size_t)i * L + l];
        tmp[(size_t)i * L + l] = pre ? x * nrm(p, i) : x;
      }
    filt(p.lat, tmp.data(), tmp.data(), L, transpose);
    for (int i = 0; i < N; ++i)
      for (int l = 0; l < L; ++l) {
        T x = tmp[(size_t)i * L + l];
        if (post) x = x * nrm(p, i);
        out[(size_t)i * L + l] = x;
      }
  }
  static T nrm(const Pairwise& p, int i);
  static void filt(const Lattice& lat, const T* in, T* out, int V, bool rev);

  // PairwisePotential::apply_upper_minus_lower_ord (src/pairwise.cpp:214-228)
  void apply_ord(int k, const float* q01, T* out) const {
    for (size_t i = 0; i < (size_t)N * L; ++i)
      if (q01[i] > 1 || q01[i] < 0) throw std::runtime_error("Value out of bounds");
    pw[k]->lat.ordered<T>(q01, out, L);
    for (int i = 0; i < N; ++i) {
      T y[64];
      compat_apply(*pw[k], L, out + (size_t)i * L, y);
      std::copy(y, y + L, out + (size_t)i * L);
    }
  }

  // PairwisePotential::gradient(b, Q) = compatibility.gradient(b, K Q)
  // (src/pairwise.cpp:390-395), FP64 accumulation into g
  void pairwise_gradient(int k, const T* b, const T* Q, double* g) const {
    std::vector<T> fq((size_t)N * L);
    kernel(k, Q, fq.data());
    compat_grad_accum(*pw[k], N, L, b, fq.data(), g);
  }

  // ---- unrolled mean field and its adjoint ---------------------------------
  // The fork has no inference loop; this is the classic mean-field update
  // built from its pieces: Q0 = expAndNormalize(-U) (src/densecrf.cpp:321-336),
  // then T times  a = -U - sum_k PairwisePotential::apply_k(Q),  Q = softmax(a).
  static void softmax_row(const T* a, T* q, int L) {
    T mx = a[0];
    for (int l = 1; l < L; ++l) mx = std::max(mx, a[l]);
    T s = 0;
    for (int l = 0; l < L; ++l) {
      q[l] = std::exp(a[l] - mx);
      s += q[l];
    }
    for (int l = 0; l < L; ++l) q[l] = q[l] / s;
  }
  // Qs receives Q_0..Q_T, Ms the filtered fields K_k Q_t (t < T, k-major
  // within t); both are what the adjoint reads
  std::vector<T> meanfield(int T_, std::vector<std::vector<T>>* Qs, std::vector<std::vector<T>>* Ms) const {
    const size_t NL = (size_t)N * L;
    std::vector<T> a(NL), Q(NL), m(NL), c(NL);
    for (size_t i = 0; i < NL; ++i) a[i] = -U[i];
    for (int i = 0; i < N; ++i) softmax_row(&a[(size_t)i * L], &Q[(size_t)i * L], L);
    if (Qs) Qs->push_back(Q);
    for (int t = 0; t < T_; ++t) {
      for (size_t i = 0; i < NL; ++i) a[i] = -U[i];
      for (size_t k = 0; k < pw.size(); ++k) {
        kernel((int)k, Q.data(), m.data());
        if (Ms) Ms->push_back(m);
        for (int i = 0; i < N; ++i) compat_apply(*pw[k], L, &m[(size_t)i * L], &c[(size_t)i * L]);
        for (size_t i = 0; i < NL; ++i) a[i] -= c[i];
      }
      for (int i = 0; i < N; ++i) softmax_row(&a[(size_t)i * L], &Q[(size_t)i * L], L);
      if (Qs) Qs->push_back(Q);
    }
    return Q;
  }
  // Reverse pass of meanfield(): given dL/dQ_T, dU = dL/dU and, per
  // potential k, dL/d(parameters of k) in the layout of
  // LabelCompatibility::parameters (g holds them back to back).
  //   softmax:  ga = Q o (gQ - <Q, gQ>)   per pixel
  //   a = -U - sum_k C_k(K_k Q):  dU -= ga;  b_k = -ga;
  //   dparams_k += C_k.gradient(b_k, K_k Q)   (the reference's gradient);
  //   gQ_prev = sum_k applyTranspose_k(b_k)    (src/pairwise.cpp:209-213)
  void meanfield_backward(int T_, const std::vector<std::vector<T>>& Qs, const std::vector<std::vector<T>>& Ms,
                          const T* dQ, T* dU, double* g) const {
    const size_t NL = (size_t)N * L;
    const int K = (int)pw.size();
    std::vector<T> gQ(dQ, dQ + NL), ga(NL), b(NL), tmp(NL), gprev(NL);
    std::vector<int> goff(K + 1, 0);
    for (int k = 0; k < K; ++k) goff[k + 1] = goff[k] + compat_nparams(*pw[k], L);
    std::fill(dU, dU + NL, (T)0);
    auto sm_back = [&](const std::vector<T>& Q) {
      for (int i = 0; i < N; ++i) {
        T s = 0;
        for (int l = 0; l < L; ++l) s += Q[(size_t)i * L + l] * gQ[(size_t)i * L + l];
        for (int l = 0; l < L; ++l) ga[(size_t)i * L + l] = Q[(size_t)i * L + l] * (gQ[(size_t)i * L + l] - s);
      }
    };
    for (int t = T_; t >= 1; --t) {
      sm_back(Qs[t]);
      for (size_t i = 0; i < NL; ++i) {
        dU[i] -= ga[i];
        b[i] = -ga[i];
      }
      std::fill(gprev.begin(), gprev.end(), (T)0);
      for (int k = 0; k < K; ++k) {
        compat_grad_accum(*pw[k], N, L, b.data(), Ms[(size_t)(t - 1) * K + k].data(), g + goff[k]);
        apply(k, b.data(), tmp.data(), false, true);
        for (size_t i = 0; i < NL; ++i) gprev[i] += tmp[i];
      }
      gQ = gprev;
    }
    sm_back(Qs[0]);
    for (size_t i = 0; i < NL; ++i) dU[i] -= ga[i];
  }

  // ---- superpixel helpers, bug-compatible (src/densecrf.cpp:58-102) --------
  // pixels -> regions: res(l,r) = sum_{pixel < |r|} p(l,pixel)  (loop index!)
  std::vector<T> sp_to_regions(const std::vector<T>& p) const {
    std::vector<T> res((size_t)L * R, (T)0);
    for (int r = 0; r < R; ++r)
      for (int l = 0; l < L; ++l)
        for (size_t pix = 0; pix < members[r].size(); ++pix) {
          const size_t px = g_mut == 7 ? (size_t)members[r][pix] : pix;
          res[(size_t)r * L + l] = res[(size_t)r * L + l] + p[px * L + l];
        }
    return res;
  }
  // regions -> pixels: res(l,pixel) += p(l,r) for pixel < |r|
  std::vector<T> sp_to_pixels(const std::vector<T>& p) const {
    std::vector<T> res((size_t)L * N, (T)0);
    for (int r = 0; r < R; ++r)
      for (int l = 0; l < L; ++l)
        for (size_t pix = 0; pix < members[r].size(); ++pix)
          res[pix * L + l] = res[pix * L + l] + p[(size_t)r * L + l];
    return res;
  }
  // scalar form (:58-72), accumulated in T
  T sp_scalar(const std::vector<T>& p1, const std::vector<T>& p2) const {
    T res = 0;
    for (int r = 0; r < R; ++r)
      for (size_t pix = 0; pix < members[r].size(); ++pix)
        for (int l = 0; l < L; ++l) res = res + p1[(size_t)r * L + l] * p2[pix * L + l];
    return res;
  }

  // ---- QP (src/densecrf.cpp:340-442) ---------------------------------------
  struct QPStats {
    std::vector<double> energy, num, denom, step;
    int iters = 0;
  };
  std::vector<T> qp(const std::vector<T>& init, int max_iter, bool ref_stop, QPStats* st) {
    const size_t NL = (size_t)N * L, LR = (size_t)L * R;
    std::vector<T> Q = init, grad_y(NL), tmp(NL), s(NL), sx(NL), psisx(NL);
    std::vector<T> constant(LR), grad_z(LR), cgz(LR, 0), sx_z(LR), z(LR, (T)1);
    for (int r = 0; r < R; ++r)
      for (int l = 0; l < L; ++l) constant[(size_t)r * L + l] = e_r[r];
    auto Qm1 = [&]() {
      std::vector<T> v(NL);
      for (size_t i = 0; i < NL; ++i) v[i] = Q[i] - (T)1;
      return v;
    };
    auto descent_z = [&]() {  // src/qp.cpp:17-25
      for (size_t i = 0; i < LR; ++i) cgz[i] = grad_z[i] >= 0 ? 0 : 1;
    };
    {  // :373-375
      std::vector<T> nr = sp_to_regions(Qm1());
      for (size_t i = 0; i < LR; ++i) grad_z[i] = ((T)1 + nr[i]) * constant[i];
      descent_z();
    }
    grad_y = U;  // :377-382
    for (size_t k = 0; k < pw.size(); ++k) {
      apply((int)k, Q.data(), tmp.data());
      for (size_t i = 0; i < NL; ++i) grad_y[i] = grad_y[i] + (T)2 * tmp[i];
    }
    auto ho_bcast = [&]() {  // multiplySuperPixels((z - 1) .* constant)
      std::vector<T> t(LR);
      for (size_t i = 0; i < LR; ++i) t[i] = (z[i] - (T)1) * constant[i];
      return sp_to_pixels(t);
    };
    auto ho_energy = [&]() -> double {  // :387 / :435
      std::vector<T> zc(LR), omz(LR);
      for (size_t i = 0; i < LR; ++i) {
        zc[i] = z[i] * constant[i];
        omz[i] = ((T)1 - z[i]) * constant[i];
      }
      T zs = vec_sum(zc.data(), (long)LR);
      T sc = sp_scalar(omz, Qm1());
      return (double)(T)(zs + sc);
    };
    {
      std::vector<T> hb = ho_bcast();
      for (size_t i = 0; i < NL; ++i) grad_y[i] = grad_y[i] + hb[i];
    }
    double old_energy = std::numeric_limits<double>::max();
    double energy = 0;  // compute_LR_QP_value(Q, 0) (:940-957)
    {
      energy += dot(U.data(), Q.data(), NL);
      energy -= 0.0;
      for (size_t k = 0; k < pw.size(); ++k) {
        apply((int)k, Q.data(), tmp.data());
        energy += dot(Q.data(), tmp.data(), NL);
      }
      energy += 0.0;
      energy += ho_energy();
    }
    if (st) st->energy.push_back(energy);
    int it = 0;
    while (ref_stop ? (old_energy - energy) > 100.0 : it < max_iter) {
      if (ref_stop && max_iter > 0 && it >= max_iter) break;
      old_energy = energy;
      ++it;
      for (int i = 0; i < N; ++i) {  // descent_direction (src/qp.cpp:5-15)
        int m = 0;
        for (int l = 1; l < L; ++l)
          if (g_mut == 6 ? grad_y[(size_t)i * L + l] <= grad_y[(size_t)i * L + m]
                         : grad_y[(size_t)i * L + l] < grad_y[(size_t)i * L + m])
            m = l;
        for (int l = 0; l < L; ++l) s[(size_t)i * L + l] = l == m ? 1 : 0;
      }
      descent_z();
      for (size_t i = 0; i < NL; ++i) sx[i] = s[i] - Q[i];
      for (size_t i = 0; i < LR; ++i) sx_z[i] = cgz[i] - z[i];
      std::fill(psisx.begin(), psisx.end(), (T)0);
      for (size_t k = 0; k < pw.size(); ++k) {
        apply((int)k, sx.data(), tmp.data());
        for (size_t i = 0; i < NL; ++i) psisx[i] = psisx[i] + tmp[i];
      }
      std::vector<T> szc(LR), cgm1(LR);
      for (size_t i = 0; i < LR; ++i) {
        szc[i] = sx_z[i] * constant[i];
        cgm1[i] = (cgz[i] - (T)1) * constant[i];
      }
      double a = (double)sp_scalar(szc, Qm1());
      double b = (double)sp_scalar(cgm1, sx);
      double num = dot(U.data(), sx.data(), NL) + 2 * dot(Q.data(), psisx.data(), NL) + a + b +
                   (double)vec_sum(sx_z.data(), (long)LR);
      double denom = dot(sx.data(), psisx.data(), NL) + (double)sp_scalar(szc, sx);
      double step = g_mut == 5 ? -num / denom : -num / (2 * denom);
      if (st) {
        st->num.push_back(num);
        st->denom.push_back(denom);
      }
      if (denom == 0 || num == 0) {
        if (st) st->energy.push_back(energy);  // no update: the energy is unchanged
        break;
      }
      if (denom < 0) step = 1;
      if (st) st->step.push_back(step);
      const T stp = (T)step, stp2 = (T)(2 * step);
      for (size_t i = 0; i < NL; ++i) Q[i] = Q[i] + stp * sx[i];
      for (size_t i = 0; i < LR; ++i) z[i] = z[i] + stp * sx_z[i];
      {
        std::vector<T> hb = ho_bcast();
        for (size_t i = 0; i < NL; ++i) grad_y[i] = grad_y[i] + (stp2 * psisx[i] + stp * hb[i]);
      }
      {
        std::vector<T> nr = sp_to_regions(Qm1());
        for (size_t i = 0; i < LR; ++i) grad_z[i] = constant[i] + nr[i] * constant[i];
      }
      std::vector<T> gu(NL);
      for (size_t i = 0; i < NL; ++i) gu[i] = grad_y[i] + U[i];
      energy = dot(Q.data(), gu.data(), NL);
      energy += ho_energy();
      if (st) st->energy.push_back(energy);
    }
    if (st) st->iters = it;
    z_labels = z;
    return Q;
  }

  // ---- energies ------------------------------------------------------------
  // currentMap (src/densecrf.cpp:929-938): first maximum per pixel
  std::vector<int> current_map(const std::vector<T>& Q) const {
    std::vector<int> lab(N);
    for (int i = 0; i < N; ++i) {
      int m = 0;
      for (int l = 1; l < L; ++l)
        if (Q[(size_t)i * L + l] > Q[(size_t)i * L + m]) m = l;
      lab[i] = m;
    }
    return lab;
  }
  // unaryEnergy + pairwise_energy_true summed per pixel (:873-927)
  std::vector<T> pixel_energy(const std::vector<int>& lab) const {
    const size_t NL = (size_t)N * L;
    std::vector<T> r(N, (T)0), q(NL), o(NL);
    for (size_t k = 0; k < pw.size(); ++k) {
      for (int i = 0; i < N; ++i)
        for (int l = 0; l < L; ++l) q[(size_t)i * L + l] = lab[i] == l ? 0 : 1;
      apply((int)k, q.data(), o.data(), /*no_norm=*/true);
      for (int i = 0; i < N; ++i) {
        T v = (lab[i] >= 0 && lab[i] < L) ? -o[(size_t)i * L + lab[i]] : (T)0;
        r[i] = r[i] + v;
      }
    }
    std::vector<T> tot(N);
    for (int i = 0; i < N; ++i) {
      T u = (lab[i] >= 0 && lab[i] < L) ? U[(size_t)i * L + lab[i]] : (T)0;
      tot[i] = u + r[i];
    }
    return tot;
  }
  double assignment_energy_true(const std::vector<int>& lab) const {
    std::vector<T> t = pixel_energy(lab);
    double e = 0;
    for (int i = 0; i < N; ++i) e += (double)t[i];
    return e;
  }
  // assignment_energy_higher_order (:844-870), bug-compatible: compares the
  // label at loop index `pix` with the label of the region's first member and
  // always adds (R-1)*e_r per region.
  double assignment_energy_ho(const std::vector<int>& lab) const {
    double e = assignment_energy_true(lab);
    double ho = 0.0;
    for (int r = 0; r < R; ++r) {
      int prev = (int16_t)lab[members[r][0]];
      for (size_t pix = 0; pix < members[r].size(); ++pix)
        if (prev != (int16_t)lab[pix]) {
          e += (double)e_r[r];
          break;
        }
      ho += (double)(T)((T)(R - 1) * e_r[r]);  // int*float product (:866)
    }
    return e + ho;
  }
  // rescale (src/eigen_utils.cpp:120-140): global min/max to [0,1]
  static std::vector<T> rescale(const std::vector<T>& Q) {
    std::vector<T> out = Q;
    T mn = *std::min_element(out.begin(), out.end());
    for (auto& x : out) x = x - mn;
    T mx = *std::max_element(out.begin(), out.end());
    if (mx < 0) throw std::runtime_error("Maxval < 0, skipping");
    if (mx > 0)
      for (auto& x : out) x = x / mx;
    return out;
  }
  // compute_energy_LP_higher_order (:982-1037)
  double lp_energy(const std::vector<T>& Q) const {
    const size_t NL = (size_t)N * L;
    double e = dot(U.data(), Q.data(), NL);
    std::vector<T> rq = rescale(Q);
    std::vector<float> rqf(rq.begin(), rq.end());
    std::vector<T> tmp(NL);
    for (size_t k = 0; k < pw.size(); ++k) {
      apply_ord((int)k, rqf.data(), tmp.data());
      e -= dot(Q.data(), tmp.data(), NL);
    }
    for (int r = 0; r < R; ++r) {
      T mdl = -std::numeric_limits<T>::infinity();
      for (int l = 0; l < L; ++l) {
        T mx = -std::numeric_limits<T>::infinity(), mn = std::numeric_limits<T>::infinity();
        for (int p : members[r]) {
          mx = std::max(mx, Q[(size_t)p * L + l]);
          mn = std::min(mn, Q[(size_t)p * L + l]);
        }
        mdl = std::max(mdl, (T)(mx - mn));
      }
      e += (double)(T)(e_r[r] * mdl);
    }
    return e;
  }
  // computeUCondGrad (:803-830): `min`/`max` are ints, so Q is truncated.
  void ucondgrad(std::vector<T>& Us, const std::vector<T>& Q) const {
    std::fill(Us.begin(), Us.end(), (T)0);
    int min_ind = 0, max_ind = 0, min_lab = 0, max_lab = 0;
    for (int r = 0; r < R; ++r) {
      int mn = std::numeric_limits<int>::max(), mx = std::numeric_limits<int>::min();
      for (int l = 0; l < L; ++l)
        for (int p : members[r]) {
          T q = Q[(size_t)p * L + l];
          if ((T)mn > q) {
            min_ind = p;
            min_lab = l;
            mn = (int)q;
            if (g_mut == 8) mn = (int)std::floor(q * 1e6);
          }
          if ((T)mx < q) {
            max_ind = p;
            max_lab = l;
            mx = (int)q;
          }
        }
      Us[(size_t)min_ind * L + min_lab] = c_r[r] * e_r[r];
      Us[(size_t)max_ind * L + max_lab] = -c_r[r] * e_r[r];
    }
  }

  // ---- LP (src/densecrf.cpp:526-799) ----------------------------------------
  struct LPParams {
    float lambda = 1e-2f, dual_gap_tol = 1e3f, prox_energy_tol = 1e3f;
    int prox_max_iter = 10, fw_max_iter = 10, qp_max_iter = 1000;
    float qp_tol = 1e3f, qp_const = 1e-16f;
    bool best_int = true;
    float less_confident_percent = 0.f, confidence_tol = 0.95f;
  };
  struct LPStats {
    std::vector<int> qp_sweeps;
    std::vector<double> gap, delta, int_energy;
    int outer = 0;
    double init_lp_energy = 0, init_int_energy = 0;
  };
  static void renormalize(std::vector<T>& Q, int L, int N) {  // :444-455
    const double uniform = 1.0 / L;
    for (int i = 0; i < N; ++i) {
      T* c = &Q[(size_t)i * L];
      double sum = (double)ColSum<T>::run(c, L, i);
      if (sum == 0 || sum != sum)
        for (int l = 0; l < L; ++l) c[l] = (T)uniform;
      else {
        T s = (T)sum;
        for (int l = 0; l < L; ++l) c[l] = c[l] / s;
      }
    }
  }
  std::vector<T> max_rounding(const std::vector<T>& Q) const {  // :832-840
    std::vector<T> r((size_t)N * L, (T)0);
    std::vector<int> m = current_map(Q);
    for (int i = 0; i < N; ++i) r[(size_t)i * L + m[i]] = 1;
    return r;
  }
  // feasible_Q (:470-492) with sortCols (src/eigen_utils.cpp:96-113)
  void feasible(std::vector<T>& out, const std::vector<T>& Q) const {
    std::vector<T> col(L);
    for (int i = 0; i < N; ++i) {
      const T* c = &Q[(size_t)i * L];
      for (int l = 0; l < L; ++l) col[l] = c[l];
      std::sort(col.begin(), col.end(), [](T a, T b) { return a > b; });
      double sum = (double)(T)(ColSum<T>::run(c, L, i) - (T)1);
      int K = -1;
      for (int k = L; k > 0; --k) {
        double uk = (double)col[k - 1];
        if (sum / k < uk) {
          K = k;
          break;
        }
        sum -= uk;
      }
      for (int l = 0; l < L; ++l)
        out[(size_t)i * L + l] = (T)std::min(std::max((double)c[l] - sum / K, 0.0), 1.0);
    }
  }

  std::vector<T> lp(const std::vector<T>& init, const LPParams& P, LPStats* st) {
    const size_t NL = (size_t)N * L;
    const int M = L;
    std::vector<T> Q = init;
    renormalize(Q, L, N);
    std::vector<T> best_Q = Q;
    double int_energy = assignment_energy_ho(current_map(Q));
    double energy = lp_energy(Q);
    if (st) {
      st->init_lp_energy = energy;
      st->init_int_energy = int_energy;
    }
    if (energy > int_energy) {
      Q = max_rounding(Q);
      energy = lp_energy(Q);
    }
    double best_energy = energy, best_int_energy = int_energy;
    const T lambda = (T)P.lambda;
    std::vector<T> alpha(NL), u(NL), sa(NL), su(NL), beta_mat(NL), gamma(NL), cur_Q(NL), tmp(NL), Y(NL), posH(NL),
        negH(NL), qpv(N);
    std::vector<T> vg(L), vs(L), vt(L), vt2(L), vy(L);
    const T c_m1_over_M = (T)(-1.0 / (double)(float)M);
    const T c_2l_over_M = (T)((T)(2 * lambda) / (T)(float)M);
    const T c_1_over_M = (T)(1.0 / (double)(float)M);
    const T c_1m2_over_M = (T)(1 - 2.0 / (double)(float)M);
    const T qp_delta = (T)P.qp_const;
    int it = 0, count = 0;
    do {
      ++it;
      std::fill(beta_mat.begin(), beta_mat.end(), (T)0);
      std::fill(gamma.begin(), gamma.end(), (T)1);
      cur_Q = Q;
      std::fill(alpha.begin(), alpha.end(), (T)0);
      std::fill(u.begin(), u.end(), (T)0);
      int pit = 0;
      do {
        ++pit;
        std::fill(sa.begin(), sa.end(), (T)0);
        std::fill(su.begin(), su.end(), (T)0);
        for (size_t i = 0; i < NL; ++i) tmp[i] = (alpha[i] + u[i]) - U[i];  // :627
        for (int i = 0; i < N; ++i) {  // qp_gamma_multiplyC (:517-523)
          for (int l = 0; l < L; ++l) vt2[l] = tmp[(size_t)i * L + l];
          T sum = ColSum<T>::run_aligned(vt2.data(), L);
          for (int l = 0; l < L; ++l) Y[(size_t)i * L + l] = (c_m1_over_M * sum + vt2[l]) * lambda;
        }
        for (size_t i = 0; i < NL; ++i) {
          Y[i] = Y[i] + cur_Q[i];
          posH[i] = std::max(-Y[i], (T)0);
          negH[i] = std::max(Y[i], (T)0);
        }
        int qpit = 0;
        T qp_value = std::numeric_limits<T>::max();
        if (std::is_same<T, double>::value) qp_value = (T)std::numeric_limits<float>::max();
        do {  // :650-675
          for (int i = 0; i < N; ++i) {
            T* g = &gamma[(size_t)i * L];
            const T* ph = &posH[(size_t)i * L];
            const T* nh = &negH[(size_t)i * L];
            for (int l = 0; l < L; ++l) vg[l] = g[l];
            T sum = ColSum<T>::run_aligned(vg.data(), L);  // qp_gamma_step (:495-514)
            for (int l = 0; l < L; ++l) {
              T t = (sum - vg[l]) * c_2l_over_M;
              T stp = (t + ph[l]) + qp_delta;
              T t2 = (c_1_over_M * sum + c_1m2_over_M * vg[l]) * lambda;
              t2 = (t2 + nh[l]) + qp_delta;
              vs[l] = stp / t2;
              vg[l] = vg[l] * vs[l];
            }
            for (int l = 0; l < L; ++l) g[l] = vg[l];
            T s2 = ColSum<T>::run_aligned(vg.data(), L);
            for (int l = 0; l < L; ++l) {
              vt[l] = (c_m1_over_M * s2 + vg[l]) * lambda;
              vy[l] = Y[(size_t)i * L + l];
            }
            qpv[i] = (T)(0.5 * (double)ColSum<T>::dot_aligned(vg.data(), vt.data(), L) +
                         (double)ColSum<T>::dot_aligned(vg.data(), vy.data(), L));
          }
          T qp_value1 = vec_sum(qpv.data(), N);
          ++qpit;
          if (std::abs(qp_value - qp_value1) < (T)P.qp_tol) break;
          qp_value = qp_value1;
        } while (qpit < P.qp_max_iter);
        if (st) st->qp_sweeps.push_back(qpit);
        for (size_t i = 0; i < NL; ++i) beta_mat[i] = ((alpha[i] + u[i]) + gamma[i]) - U[i];  // :684-691
        for (int i = 0; i < N; ++i) {
          T b = -ColSum<T>::run(&beta_mat[(size_t)i * L], L, i);
          b = b / (T)M;
          for (int l = 0; l < L; ++l) beta_mat[(size_t)i * L + l] = b;
        }
        for (size_t i = 0; i < NL; ++i)  // :697
          Q[i] = lambda * ((((alpha[i] + beta_mat[i]) + u[i]) + gamma[i]) - U[i]) + cur_Q[i];
        std::vector<T> rq = rescale(Q);
        std::vector<float> rqf(rq.begin(), rq.end());
        for (size_t k = 0; k < pw.size(); ++k) {  // :708-717
          apply_ord((int)k, rqf.data(), tmp.data());
          for (size_t i = 0; i < NL; ++i) sa[i] = sa[i] + tmp[i];
        }
        ucondgrad(su, Q);  // :720
        for (size_t i = 0; i < NL; ++i) tmp[i] = (alpha[i] + u[i]) - (sa[i] + su[i]);  // :725-738
        double gap = dot(tmp.data(), Q.data(), NL);
        double tt = dot(tmp.data(), tmp.data(), NL);
        float delta = (float)(gap / ((double)(float)lambda * tt));
        delta = std::min(std::max(delta, 0.0f), 1.0f);
        if (st) {
          st->gap.push_back(gap);
          st->delta.push_back(delta);
        }
        const T dl = (T)delta;
        for (size_t i = 0; i < NL; ++i) {
          alpha[i] = alpha[i] + dl * (sa[i] - alpha[i]);
          u[i] = u[i] + dl * (su[i] - u[i]);
        }
      } while (pit < P.fw_max_iter);
      feasible(tmp, Q);  // :746-748
      Q = tmp;
      renormalize(Q, L, N);
      double prev_int = int_energy;
      int_energy = assignment_energy_ho(current_map(Q));
      if (st) st->int_energy.push_back(int_energy);
      if (P.best_int) {
        if (std::abs(int_energy - prev_int) < P.prox_energy_tol)
          ++count;
        else
          count = 0;
        if (count >= 5) break;
        if (int_energy < best_int_energy) {
          best_Q = Q;
          best_int_energy = int_energy;
        }
      } else {
        double prev = energy;
        energy = lp_energy(Q);
        if (std::abs(energy - prev) < P.prox_energy_tol)
          ++count;
        else
          count = 0;
        if (count >= 5) break;
        if (energy < best_energy) {
          best_Q = Q;
          best_energy = energy;
        }
      }
      if (P.less_confident_percent > 0) {  // :784-794 with less_confident_pixels (:457-462)
        int n_lc = 0;
        for (int i = 0; i < N; ++i) {
          const T* col = &best_Q[(size_t)i * L];
          T mx = col[0];
          for (int l = 1; l < L; ++l) mx = std::max(mx, col[l]);
          if ((float)mx <= P.confidence_tol) ++n_lc;
        }
        if (double(n_lc) / double(N) * 100 > (double)P.less_confident_percent) break;
      }
    } while (it < P.prox_max_iter);
    if (st) st->outer = it;
    return best_Q;
  }
};

template <> float CRF<float>::nrm(const Pairwise& p, int i) { return p.norm_f[i]; }
template <> double CRF<double>::nrm(const Pairwise& p, int i) { return p.norm_d[i]; }
template <> void CRF<float>::filt(const Lattice& lat, const float* in, float* out, int V, bool rev) {
  lat.filter_f(in, out, V, rev);
}
template <> void CRF<double>::filt(const Lattice& lat, const double* in, double* out, int V, bool rev) {
  lat.filter_d(in, out, V, rev);
}

// DenseKernel::initLattice (src/pairwise.cpp:40-57)
static void pairwise_init(Pairwise& p) {
  p.lat.init(p.feat.data(), (int)(p.feat.size() / p.d), p.d);
  const int N = p.lat.N;
  p.norm_f.assign(N, 1.f);
  p.norm_d.assign(N, 1.0);
  if (p.ntype != 0) {
    std::vector<float> ones(N, 1.f), kn(N);
    p.lat.filter_f(ones.data(), kn.data(), 1, false);
    std::vector<double> onesd(N, 1.0), knd(N);
    p.lat.filter_d(onesd.data(), knd.data(), 1, false);
    for (int i = 0; i < N; ++i) {
      if (p.ntype == 3) {
        p.norm_f[i] = (float)(1.0 / std::sqrt((double)kn[i] + 1e-20));
        p.norm_d[i] = 1.0 / std::sqrt(knd[i] + 1e-20);
      } else {
        p.norm_f[i] = (float)(1.0 / ((double)kn[i] + 1e-20));
        p.norm_d[i] = 1.0 / (knd[i] + 1e-20);
      }
    }
  }
}

}  // namespace orc

// ===========================================================================
// C ABI for ctypes (tests, fixture checks and the CPU baseline).
// ===========================================================================
using namespace orc;

struct OrcCRF {
  int N, L;
  std::vector<Pairwise*> pw;
  CRF<float>* f;
  CRF<double>* d;
};

static thread_local std::string g_err;
static std::map<void*, orc::Restricted> g_restricted;

extern "C" {

const char* orc_last_error() { return g_err.c_str(); }
void orc_set_mutation(int m) { orc::g_mut = m; }
void orc_set_reference_path(int scalar) { orc::g_scalar = scalar ? 1 : 0; }

void* orc_lattice_new(const float* feat, int N, int d) {
  Lattice* L = new Lattice();
  L->init(feat, N, d);
  return L;
}
void orc_lattice_free(void* h) {
  g_restricted.erase(h);
  delete (Lattice*)h;
}
int orc_lattice_M(void* h) { return ((Lattice*)h)->M; }
int orc_lattice_Npad(void* h) { return ((Lattice*)h)->Npad; }
void orc_lattice_get(void* h, int* offset, int* rank, float* bary, short* keys, int* nbr) {
  Lattice* L = (Lattice*)h;
  if (offset) std::copy(L->offset.begin(), L->offset.end(), offset);
  if (rank) std::copy(L->rank.begin(), L->rank.end(), rank);
  if (bary) std::copy(L->bary.begin(), L->bary.end(), bary);
  if (keys) std::copy(L->keys.begin(), L->keys.end(), keys);
  if (nbr) std::copy(L->nbr.begin(), L->nbr.end(), nbr);
}
void orc_filter_f(void* h, const float* in, float* out, int V, int reverse) {
  ((Lattice*)h)->filter_f(in, out, V, reverse != 0);
}
void orc_filter_d(void* h, const double* in, double* out, int V, int reverse) {
  ((Lattice*)h)->filter_d(in, out, V, reverse != 0);
}
void orc_ordered_f(void* h, const float* q, float* out, int V) { ((Lattice*)h)->ordered<float>(q, out, V); }
void orc_ordered_d(void* h, const float* q, double* out, int V) { ((Lattice*)h)->ordered<double>(q, out, V); }
void orc_lattice_gradient(void* h, const float* a, const float* b, int V, float* df) {
  ((Lattice*)h)->gradient(a, b, V, df);
}
void orc_lattice_dc(void* h, float* out, int low, int ml, int mh, int high) {
  orc::DcLattice dl(*(Lattice*)h);
  dl.dc(out, low, ml, mh, high);
}
void orc_uml_dc(const float* f, int N, int d, const int* ind, int L, float* out) { orc::uml_dc(f, N, d, ind, L, out); }
void orc_bf_apply(const float* f, int N, int d, const float* Q, int L, float* out) { orc::bf_apply(f, N, d, Q, L, out); }
void orc_bf_uml(const float* f, int N, int d, const int* ind, const float* Q, int L, float* out) {
  orc::bf_uml(f, N, d, ind, Q, L, out);
}
// restricted ordered filter; its stored state lives beside the lattice handle
int orc_ordered_restricted(void* h, const float* rin, const int* pI, int nI, const float* ext, int V, int store,
                           float* out) {
  try {
    g_restricted[h].run(*(Lattice*)h, rin, pI, nI, ext, V, store != 0, out);
  } catch (std::exception& e) {
    g_err = e.what();
    return -1;
  }
  return 0;
}

void* orc_crf_new(int N, int L) {
  OrcCRF* c = new OrcCRF{N, L, {}, new CRF<float>(N, L), new CRF<double>(N, L)};
  return c;
}
void orc_crf_free(void* h) {
  OrcCRF* c = (OrcCRF*)h;
  for (auto* p : c->pw) delete p;
  delete c->f;
  delete c->d;
  delete c;
}
void orc_crf_set_unary(void* h, const float* U) {
  OrcCRF* c = (OrcCRF*)h;
  size_t n = (size_t)c->N * c->L;
  for (size_t i = 0; i < n; ++i) {
    c->f->U[i] = U[i];
    c->d->U[i] = U[i];
  }
}
// features pixel-major (feat[i*d+k]); ntype as NormalizationType
void orc_crf_add_pairwise(void* h, const float* feat, int d, float w, int ntype) {
  OrcCRF* c = (OrcCRF*)h;
  Pairwise* p = new Pairwise();
  p->d = d;
  p->w = w;
  p->nn_w = w;
  p->ntype = ntype;
  p->feat.assign(feat, feat + (size_t)c->N * d);
  pairwise_init(*p);
  c->pw.push_back(p);
  c->f->pw.push_back(p);
  c->d->pw.push_back(p);
}
// as orc_crf_add_pairwise with a Diagonal (ctype 1, params v[L]) or Matrix
// (ctype 2, params m[L*L] column-major, stored as (m + m^T)/2) compatibility
void orc_crf_add_pairwise_compat(void* h, const float* feat, int d, int ctype, const float* params, int ntype) {
  OrcCRF* c = (OrcCRF*)h;
  const int L = c->L;
  orc_crf_add_pairwise(h, feat, d, ctype == 0 ? params[0] : 1.f, ntype);
  Pairwise* p = c->pw.back();
  p->ctype = ctype;
  if (ctype == 1) p->cp.assign(params, params + L);
  if (ctype == 2) {
    p->cp.assign((size_t)L * L, 0.f);
    for (int j = 0; j < L; ++j)
      for (int i = 0; i < L; ++i)
        p->cp[(size_t)j * L + i] = 0.5f * (params[(size_t)j * L + i] + params[(size_t)i * L + j]);
  }
  if (ctype != 0) p->nn_w = p->cp[0];
}
int orc_crf_nparams(void* h, int k) {
  OrcCRF* c = (OrcCRF*)h;
  return compat_nparams(*c->pw[k], c->L);
}
void orc_crf_pairwise_gradient_f(void* h, int k, const float* b, const float* Q, double* g) {
  OrcCRF* c = (OrcCRF*)h;
  std::fill(g, g + compat_nparams(*c->pw[k], c->L), 0.0);
  c->f->pairwise_gradient(k, b, Q, g);
}
void orc_crf_pairwise_gradient_d(void* h, int k, const double* b, const double* Q, double* g) {
  OrcCRF* c = (OrcCRF*)h;
  std::fill(g, g + compat_nparams(*c->pw[k], c->L), 0.0);
  c->d->pairwise_gradient(k, b, Q, g);
}
void orc_crf_meanfield_f(void* h, int T, float* Q) {
  OrcCRF* c = (OrcCRF*)h;
  std::vector<float> q = c->f->meanfield(T, nullptr, nullptr);
  std::copy(q.begin(), q.end(), Q);
}
// FP64 forward (Q_T into Q, may be null) and adjoint of dL/dQ_T = dQ:
// dU [N*L] and the parameter gradients of every potential back to back in g
void orc_crf_meanfield_grad_d(void* h, int T, const double* dQ, double* Q, double* dU, double* g) {
  OrcCRF* c = (OrcCRF*)h;
  std::vector<std::vector<double>> Qs, Ms;
  std::vector<double> q = c->d->meanfield(T, &Qs, &Ms);
  if (Q) std::copy(q.begin(), q.end(), Q);
  int np = 0;
  for (auto* p : c->pw) np += compat_nparams(*p, c->L);
  std::fill(g, g + np, 0.0);
  if (dQ) c->d->meanfield_backward(T, Qs, Ms, dQ, dU, g);
}
void orc_crf_meanfield_d(void* h, int T, double* Q) {
  OrcCRF* c = (OrcCRF*)h;
  std::vector<double> q = c->d->meanfield(T, nullptr, nullptr);
  std::copy(q.begin(), q.end(), Q);
}
int orc_crf_lattice_M(void* h, int k) { return ((OrcCRF*)h)->pw[k]->lat.M; }
void orc_crf_get_norm(void* h, int k, float* nf, double* nd) {
  Pairwise* p = ((OrcCRF*)h)->pw[k];
  if (nf) std::copy(p->norm_f.begin(), p->norm_f.end(), nf);
  if (nd) std::copy(p->norm_d.begin(), p->norm_d.end(), nd);
}
// one superpixel layer: region id per pixel in [0,R); e/c per region
void orc_crf_add_layer(void* h, const int* region, int R, const float* e, const float* cst) {
  OrcCRF* c = (OrcCRF*)h;
  std::vector<std::vector<int>> mem(R);
  for (int i = 0; i < c->N; ++i) mem[region[i]].push_back(i);
  int kept = 0;
  for (int r = 0; r < R; ++r) {
    if (mem[r].empty()) continue;  // regions without pixels are dropped
    ++kept;
    c->f->members.push_back(mem[r]);
    c->d->members.push_back(mem[r]);
    c->f->e_r.push_back(e[r]);
    c->d->e_r.push_back(e[r]);
    c->f->c_r.push_back(cst[r]);
    c->d->c_r.push_back(cst[r]);
  }
  c->f->R += kept;
  c->d->R += kept;
}
int orc_crf_apply_f(void* h, int k, const float* in, float* out, int no_norm, int transpose) {
  ((OrcCRF*)h)->f->apply(k, in, out, no_norm != 0, transpose != 0);
  return 0;
}
int orc_crf_apply_d(void* h, int k, const double* in, double* out, int no_norm, int transpose) {
  ((OrcCRF*)h)->d->apply(k, in, out, no_norm != 0, transpose != 0);
  return 0;
}
int orc_crf_apply_ord_f(void* h, int k, const float* q, float* out) {
  try {
    ((OrcCRF*)h)->f->apply_ord(k, q, out);
  } catch (std::exception& e) {
    g_err = e.what();
    return -1;
  }
  return 0;
}
int orc_crf_apply_ord_d(void* h, int k, const float* q, double* out) {
  try {
    ((OrcCRF*)h)->d->apply_ord(k, q, out);
  } catch (std::exception& e) {
    g_err = e.what();
    return -1;
  }
  return 0;
}

// QP.  dbl selects the FP64 oracle; init/out are float (converted).  stats:
// energies[max_iter+1], num/denom/step[max_iter]; returns iteration count.
int orc_crf_qp(void* h, int dbl, const float* init, float* out, int max_iter, int ref_stop, double* energies,
               double* num, double* denom, double* step, int cap) {
  OrcCRF* c = (OrcCRF*)h;
  size_t NL = (size_t)c->N * c->L;
  auto dump = [&](auto& st) {
    for (int i = 0; i < cap + 1; ++i)
      if (energies) energies[i] = i < (int)st.energy.size() ? st.energy[i] : NAN;
    for (size_t i = 0; i < st.num.size() && (int)i < cap; ++i) {
      if (num) num[i] = st.num[i];
      if (denom) denom[i] = st.denom[i];
    }
    for (size_t i = 0; i < st.step.size() && (int)i < cap; ++i)
      if (step) step[i] = st.step[i];
    return st.iters;
  };
  if (dbl) {
    std::vector<double> q0(init, init + NL);
    CRF<double>::QPStats st;
    std::vector<double> Q = c->d->qp(q0, max_iter, ref_stop != 0, &st);
    for (size_t i = 0; i < NL; ++i) out[i] = (float)Q[i];
    return dump(st);
  }
  std::vector<float> q0(init, init + NL);
  CRF<float>::QPStats st;
  std::vector<float> Q = c->f->qp(q0, max_iter, ref_stop != 0, &st);
  std::copy(Q.begin(), Q.end(), out);
  return dump(st);
}
void orc_crf_zlabels(void* h, float* z) {
  OrcCRF* c = (OrcCRF*)h;
  std::copy(c->f->z_labels.begin(), c->f->z_labels.end(), z);
}

// LP.  params: lambda, dual_gap_tol, prox_energy_tol, qp_tol, qp_const,
// less_confident_percent, confidence_tol as floats; ints prox_max_iter, fw_max_iter, qp_max_iter, best_int.  stats
// arrays are filled up to `cap` entries; returns outer iteration count or -1.
int orc_crf_lp(void* h, int dbl, const float* init, float* out, const float* fp, const int* ip, int* qp_sweeps,
               double* gap, double* delta, double* int_energy, int cap, double* init_energies) {
  OrcCRF* c = (OrcCRF*)h;
  size_t NL = (size_t)c->N * c->L;
  auto fill = [&](auto& P) {
    P.lambda = fp[0];
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
  };
  auto dump = [&](auto& st) {
    for (size_t i = 0; i < st.qp_sweeps.size() && (int)i < cap; ++i) {
      if (qp_sweeps) qp_sweeps[i] = st.qp_sweeps[i];
      if (gap) gap[i] = st.gap[i];
      if (delta) delta[i] = st.delta[i];
    }
    for (size_t i = 0; i < st.int_energy.size() && (int)i < cap; ++i)
      if (int_energy) int_energy[i] = st.int_energy[i];
    if (init_energies) {
      init_energies[0] = st.init_lp_energy;
      init_energies[1] = st.init_int_energy;
    }
    return st.outer;
  };
  try {
    if (dbl) {
      CRF<double>::LPParams P;
      fill(P);
      std::vector<double> q0(init, init + NL);
      CRF<double>::LPStats st;
      std::vector<double> Q = c->d->lp(q0, P, &st);
      for (size_t i = 0; i < NL; ++i) out[i] = (float)Q[i];
      return dump(st);
    }
    CRF<float>::LPParams P;
    fill(P);
    std::vector<float> q0(init, init + NL);
    CRF<float>::LPStats st;
    std::vector<float> Q = c->f->lp(q0, P, &st);
    std::copy(Q.begin(), Q.end(), out);
    return dump(st);
  } catch (std::exception& e) {
    g_err = e.what();
    return -1;
  }
}

// energies of a labelling / of a relaxed Q (float instantiation unless dbl)
double orc_crf_assignment_energy_true(void* h, int dbl, const int* lab) {
  OrcCRF* c = (OrcCRF*)h;
  std::vector<int> l(lab, lab + c->N);
  return dbl ? c->d->assignment_energy_true(l) : c->f->assignment_energy_true(l);
}
double orc_crf_assignment_energy_ho(void* h, int dbl, const int* lab) {
  OrcCRF* c = (OrcCRF*)h;
  std::vector<int> l(lab, lab + c->N);
  return dbl ? c->d->assignment_energy_ho(l) : c->f->assignment_energy_ho(l);
}
double orc_crf_lp_energy(void* h, int dbl, const float* Q) {
  OrcCRF* c = (OrcCRF*)h;
  size_t NL = (size_t)c->N * c->L;
  if (dbl) return c->d->lp_energy(std::vector<double>(Q, Q + NL));
  return c->f->lp_energy(std::vector<float>(Q, Q + NL));
}
void orc_crf_ucondgrad(void* h, const float* Q, float* Us) {
  OrcCRF* c = (OrcCRF*)h;
  size_t NL = (size_t)c->N * c->L;
  std::vector<float> q(Q, Q + NL), u(NL);
  c->f->ucondgrad(u, q);
  std::copy(u.begin(), u.end(), Us);
}
void orc_crf_current_map(void* h, const float* Q, int* lab) {
  OrcCRF* c = (OrcCRF*)h;
  size_t NL = (size_t)c->N * c->L;
  std::vector<int> m = c->f->current_map(std::vector<float>(Q, Q + NL));
  std::copy(m.begin(), m.end(), lab);
}
// expAndNormalize(-U) (src/densecrf.cpp:321-336)
void orc_unary_init(const float* U, float* Q, int N, int L) {
  std::vector<float> b(L);
  for (int i = 0; i < N; ++i) {
    float mx = -U[(size_t)i * L];
    for (int l = 0; l < L; ++l) mx = std::max(mx, -U[(size_t)i * L + l]);
    for (int l = 0; l < L; ++l) b[l] = std::exp(-U[(size_t)i * L + l] - mx);
    float s = ColSum<float>::run_aligned(b.data(), L);
    for (int l = 0; l < L; ++l) Q[(size_t)i * L + l] = b[l] / s;
  }
}
void orc_feasible(const float* Q, float* out, int N, int L) {
  CRF<float> c(N, L);
  std::vector<float> q(Q, Q + (size_t)N * L), o((size_t)N * L);
  c.feasible(o, q);
  std::copy(o.begin(), o.end(), out);
}
void orc_renormalize(float* Q, int N, int L) {
  std::vector<float> q(Q, Q + (size_t)N * L);
  CRF<float>::renormalize(q, L, N);
  std::copy(q.begin(), q.end(), Q);
}

}  // extern "C"
