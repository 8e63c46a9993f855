// Divide-and-conquer "upper minus lower" filter of one label:
//   * Permutohedral::compute_upper_minus_lower_dc(out, low, ml, mh, high)
//     (src/permutohedral.cpp:881-997): the pixels [mh, high) splatted with
//     weight 1, blurred over the lattice points of [low, ml) and [mh, high)
//     only, sliced into [low, ml) (added); then [low, ml) splatted, blurred
//     over the same points and sliced into [mh, high) (subtracted);
//   * PairwisePotential::apply_upper_minus_lower_dc(out, ind)
//     (src/pairwise.cpp:241-293): per label a lattice on the features in the
//     order ind(l, .), and apply_upper_minus_lower_sorted_slice's recursion:
//     blocks of at most 30 pixels by brute force (alpha 1/0.6), larger ones
//     as their two halves (sharing the middle pixel when the size is odd)
//     and then the lattice step between them.  Row l of the result stays in
//     the sorted order.
//
// The recursion runs on the host in the reference's order (left half, right
// half, then the node), one launch per node, so every output element gets
// its terms in the reference's order.  A node is one workgroup: it lists
// its active points (an atomic flag per point), and the blur axes are
// separated by barriers.  The point values live in two device arrays that
// are zero outside a node; each node clears what it touched.  Points outside
// the active list read as zero, as in the reference, whose blur updates only
// the listed points.  The work of a node is linear in its pixels and
// points; the reference's is linear in the whole lattice per node.
#include "lattice.hpp"

namespace dcrf {
namespace {

constexpr int kDcBlock = 1024;
constexpr int kDcLeaf = 30;  // SMALLEST_BLOCK, include/pairwise.h:31

struct DcTables {
  int M, d;
  const int* offset;
  const float* bary;
  const int* nbr;
  const int* seg;
  const int* sp_pix;
  const float* sp_w;
};

// splat weight 1 of the pixels [p0, p1) into A at the listed points, in pixel order
__device__ void dc_splat(const DcTables& t, const int* list, int nl, int p0, int p1, float* A) {
  for (int a = threadIdx.x; a < nl; a += blockDim.x) {
    const int p = list[a];
    float s = 0.f;
    for (int e = t.seg[p]; e < t.seg[p + 1]; ++e) {
      const int px = t.sp_pix[e];
      if (px >= p0 && px < p1) s = s + t.sp_w[e];
    }
    A[p + 1] = s;
  }
}

// the d+1 blur axes over the listed points (sum in double, rounded once);
// returns the buffer holding the result
__device__ float* dc_blur(const DcTables& t, const int* list, int nl, float* A, float* B) {
  float* src = A;
  float* dst = B;
  for (int j = 0; j <= t.d; ++j) {
    __syncthreads();
    const int* nb = t.nbr + (size_t)j * t.M * 2;
    for (int a = threadIdx.x; a < nl; a += blockDim.x) {
      const int p = list[a];
      const int n1 = nb[2 * p] + 1, n2 = nb[2 * p + 1] + 1;
      dst[p + 1] = (float)((double)src[p + 1] + 0.5 * (double)(src[n1] + src[n2]));
    }
    float* tmp = src;
    src = dst;
    dst = tmp;
  }
  __syncthreads();
  return src;
}

__global__ void __launch_bounds__(kDcBlock) k_dc_node(DcTables t, int low, int ml, int mh, int high, float alpha,
                                                      float* __restrict__ A, float* __restrict__ B,
                                                      int* __restrict__ flag, int* __restrict__ list,
                                                      float* __restrict__ out) {
  __shared__ int nl;
  const int d1 = t.d + 1;
  if (threadIdx.x == 0) nl = 0;
  __syncthreads();
  // active points: those of [mh, high) first, then of [low, ml) (the order
  // of the list does not change any value)
  const long nu = (long)(high - mh) * d1, nlo = (long)(ml - low) * d1;
  for (long q = threadIdx.x; q < nu + nlo; q += blockDim.x) {
    const long pix = q < nu ? mh + q / d1 : low + (q - nu) / d1;
    const int j = (int)(q < nu ? q % d1 : (q - nu) % d1);
    const int o = t.offset[pix * d1 + j];
    if (atomicExch(&flag[o], 1) == 0) list[atomicAdd(&nl, 1)] = o;
  }
  __syncthreads();
  const int n = nl;
  // upper: [mh, high) splatted, sliced into [low, ml)
  dc_splat(t, list, n, mh, high, A);
  float* v = dc_blur(t, list, n, A, B);
  for (int i = low + threadIdx.x; i < ml; i += blockDim.x) {
    float o = out[i];
    for (int j = 0; j < d1; ++j) o += t.bary[(size_t)i * d1 + j] * v[t.offset[(size_t)i * d1 + j] + 1] * alpha;
    out[i] = o;
  }
  __syncthreads();
  for (int a = threadIdx.x; a < n; a += blockDim.x) A[list[a] + 1] = B[list[a] + 1] = 0.f;
  __syncthreads();
  // lower: [low, ml) splatted, sliced into [mh, high)
  dc_splat(t, list, n, low, ml, A);
  v = dc_blur(t, list, n, A, B);
  for (int i = mh + threadIdx.x; i < high; i += blockDim.x) {
    float o = out[i];
    for (int j = 0; j < d1; ++j) o -= t.bary[(size_t)i * d1 + j] * v[t.offset[(size_t)i * d1 + j] + 1] * alpha;
    out[i] = o;
  }
  __syncthreads();
  for (int a = threadIdx.x; a < n; a += blockDim.x) {
    const int p = list[a];
    A[p + 1] = B[p + 1] = 0.f;
    flag[p] = 0;
  }
}

// a block of at most 30 sorted pixels by brute force: lower then upper
__global__ void __launch_bounds__(64) k_dc_leaf(int d, const float* __restrict__ fs, int mn, int mx,
                                                float* __restrict__ out) {
  const int c = mn + threadIdx.x;
  if (c >= mx) return;
  const double alpha = 1.0 / 0.6;
  float o = out[c];
  for (int b = mn; b < c; ++b) o = (float)((double)o - bf_kernel(&fs[(size_t)c * d], &fs[(size_t)b * d], d) * alpha);
  for (int b = c + 1; b < mx; ++b)
    o = (float)((double)o + bf_kernel(&fs[(size_t)c * d], &fs[(size_t)b * d], d) * alpha);
  out[c] = o;
}

__global__ void __launch_bounds__(kBlock) k_dc_gather(int N, int d, int L, int l, const int* __restrict__ ind,
                                                      const float* __restrict__ f, float* __restrict__ fs) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)N * d) return;
  const long j = t / d;
  fs[t] = f[(size_t)ind[j * L + l] * d + t % d];
}
__global__ void __launch_bounds__(kBlock) k_dc_scatter(int N, int L, int l, const float* __restrict__ col,
                                                       float* __restrict__ out) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j < N) out[(size_t)j * L + l] = col[j];
}

}  // namespace

struct DcScratch {
  DBuf<float> A, B;
  DBuf<int> flag, list;
  void ensure(int M, hipStream_t s) {
    if (flag.n >= (size_t)M + 2) return;
    A.resize(M + 2);
    B.resize(M + 2);
    flag.resize(M + 2);
    list.resize(M + 2);
    DCRF_HIP(hipMemsetAsync(A.get(), 0, (M + 2) * sizeof(float), s));
    DCRF_HIP(hipMemsetAsync(B.get(), 0, (M + 2) * sizeof(float), s));
    DCRF_HIP(hipMemsetAsync(flag.get(), 0, (M + 2) * sizeof(int), s));
  }
};

void Lattice::dc_step(float* out, int low, int ml, int mh, int high, DcScratch& sc) {
  sc.ensure(M_, stream_);
  const DcTables t{M_, d_, offset_.get(), bary_.get(), nbr_.get(), seg_.get(), sp_pix_.get(), sp_w_.get()};
  k_dc_node<<<1, kDcBlock, 0, stream_>>>(t, low, ml, mh, high, lattice_alpha(d_), sc.A.get(), sc.B.get(),
                                        sc.flag.get(), sc.list.get(), out);
  check_launch("k_dc_node");
}

void Lattice::compute_upper_minus_lower_dc(float* out, int low, int ml, int mh, int high) {
  DcScratch sc;
  dc_step(out, low, ml, mh, high, sc);
  DCRF_HIP(hipStreamSynchronize(stream_));
}

// apply_upper_minus_lower_sorted_slice (src/pairwise.cpp:262-292)
static void dc_sorted_slice(Lattice& lat, DcScratch& sc, const float* fs, int d, float* col, int mn, int mx) {
  const int size = mx - mn;
  DCRF_REQUIRE(size > 0, "apply_upper_minus_lower_dc: empty block");
  if (size <= kDcLeaf) {
    k_dc_leaf<<<1, 64, 0, lat.stream()>>>(d, fs, mn, mx, col);
    check_launch("k_dc_leaf");
    return;
  }
  const int ml = size % 2 == 0 ? mn + size / 2 : (int)std::floor(mn + size / 2.0);
  const int mh = size % 2 == 0 ? ml : ml + 1;
  dc_sorted_slice(lat, sc, fs, d, col, mn, mh);
  dc_sorted_slice(lat, sc, fs, d, col, ml, mx);
  lat.dc_step(col, mn, ml, mh, mx, sc);
}

// largest N * L of the divide-and-conquer filter (about 70,000 node launches)
constexpr long long kDcMaxPixelLabels = 1LL << 20;

void potential_apply_uml_dc(Potential& p, const int* ind, float* out, int L) {
  DCRF_REQUIRE(p.feat.get() != nullptr, "apply_upper_minus_lower_dc: the potential keeps no features");
  DCRF_REQUIRE(p.d <= 16, "apply_upper_minus_lower_dc: at most 16 features");
  hipStream_t s = p.lat->stream();
  const int N = p.lat->N(), d = p.d;
  // the recursion runs the reference's node order, one launch per node: about
  // N / 15 launches per label, so the form is bounded to small inputs
  DCRF_REQUIRE((long long)N * L <= kDcMaxPixelLabels,
               "apply_upper_minus_lower_dc: N * L above 2^20 (one launch per 15 pixels and label; use the ordered "
               "filter for large inputs)");
  float* tgt = out;
  if (p.ctype != kPotts) {
    p.ctmp.resize((size_t)N * L);
    tgt = p.ctmp.get();
  }
  DBuf<float> fs((size_t)N * d), col(N);
  DcScratch sc;
  for (int l = 0; l < L; ++l) {
    k_dc_gather<<<grid_for((size_t)N * d), kBlock, 0, s>>>(N, d, L, l, ind, p.feat.get(), fs.get());
    check_launch("k_dc_gather");
    Lattice lat(s);  // the reference builds a PairwisePotential on the sorted features per label
    lat.build(fs.get(), N, d);
    k_fill<<<grid_for(N), kBlock, 0, s>>>(col.get(), (size_t)N, 0.f);
    sc.A.release();  // sized per lattice
    sc.flag.release();
    dc_sorted_slice(lat, sc, fs.get(), d, col.get(), 0, N);
    k_dc_scatter<<<grid_for(N), kBlock, 0, s>>>(N, L, l, col.get(), tgt);
    check_launch("k_dc_scatter");
    DCRF_HIP(hipStreamSynchronize(s));
  }
  bf_compat(p, tgt, out, N, L, s);
  DCRF_HIP(hipStreamSynchronize(s));
}

}  // namespace dcrf
