// less_confident_pixels without device allocation (include/densecrf_ho_amd.h:
// dcrf_less_confident_pixels_noalloc).  The pixels are cut into tiles of
// kLcTile; one launch counts each tile's pixels whose largest label value is
// <= tol, one block scans the tile counts into tile offsets and the total, and
// a third launch writes each tile's pixel indices in ascending order from its
// offset.  The only memory beyond the caller's Q, indices and count is the
// file-scope device array `scratch` below: a tile count and a tile offset per
// tile and the total.  Q is read twice (once per pass over the tiles) instead
// of storing a flag per pixel, so the array's size does not grow with N.
//
// Every call on one device shares that device's copy of `scratch`; calls that
// overlap on one device race.  The count pass writes every tile count the
// scan reads, so nothing in it has to start at zero.

namespace dcrf {

constexpr int kLcTile = 1024;            // pixels per tile (4 per thread of a 256-thread block)
constexpr int kLcMaxTiles = 16384;       // largest tile count: N <= 16,777,216 pixels
constexpr int kLcScanThreads = 1024;     // one block scans the tile counts
constexpr long long kLcMaxPixels = (long long)kLcTile * kLcMaxTiles;
// scratch layout (ints): [0, kLcMaxTiles) tile counts, [kLcMaxTiles, 2 kLcMaxTiles)
// tile offsets, [2 kLcMaxTiles] the total
constexpr int kLcScratchInts = 2 * kLcMaxTiles + 64;

}  // namespace dcrf

static __device__ __align__(256) int scratch[dcrf::kLcScratchInts];

namespace dcrf {
namespace {

static_assert(kLcTile == 4 * kBlock, "four pixels per thread");
static_assert(kLcMaxTiles % kLcScanThreads == 0, "whole scan chunks");

// flags of the four consecutive pixels 4t..4t+3 of tile `tile`
__device__ inline int lc_flags4(const float* __restrict__ Q, int N, int L, float tol, int tile, int t, int (&f)[4]) {
  const long long base = (long long)tile * kLcTile + 4 * t;
  int n = 0;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const long long i = base + j;
    f[j] = 0;
    if (i < N) {
      const float* col = Q + i * L;
      float mx = col[0];
      for (int l = 1; l < L; ++l) mx = col[l] > mx ? col[l] : mx;
      f[j] = mx <= tol ? 1 : 0;
    }
    n += f[j];
  }
  return n;
}

__global__ void __launch_bounds__(kBlock) k_lc_tile_count(const float* __restrict__ Q, int N, int L, float tol,
                                                          int* __restrict__ cnt) {
  __shared__ int part[kBlock / 64];
  int f[4];
  int n = lc_flags4(Q, N, L, tol, blockIdx.x, threadIdx.x, f);
  for (int o = 32; o > 0; o >>= 1) n += __shfl_down(n, o, 64);
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = n;
  __syncthreads();
  if (threadIdx.x == 0) {
    int s = 0;
    for (int w = 0; w < kBlock / 64; ++w) s += part[w];
    cnt[blockIdx.x] = s;
  }
}

// exclusive scan of ntiles (<= kLcMaxTiles) tile counts by one block: each
// thread sums a contiguous chunk, the chunk sums are scanned in LDS, then each
// chunk is rescanned from its carry
__global__ void __launch_bounds__(kLcScanThreads) k_lc_tile_scan(const int* __restrict__ cnt, int ntiles,
                                                                 int* __restrict__ off, int* __restrict__ total) {
  constexpr int kPer = kLcMaxTiles / kLcScanThreads;
  __shared__ int sums[kLcScanThreads];
  const int t = threadIdx.x;
  const int b0 = t * kPer;
  int s = 0;
#pragma unroll
  for (int k = 0; k < kPer; ++k) s += (b0 + k < ntiles) ? cnt[b0 + k] : 0;
  sums[t] = s;
  __syncthreads();
  // Hillis-Steele inclusive scan of the chunk sums
  for (int o = 1; o < kLcScanThreads; o <<= 1) {
    const int v = t >= o ? sums[t - o] : 0;
    __syncthreads();
    sums[t] += v;
    __syncthreads();
  }
  int carry = t > 0 ? sums[t - 1] : 0;
#pragma unroll
  for (int k = 0; k < kPer; ++k) {
    if (b0 + k < ntiles) {
      off[b0 + k] = carry;
      carry += cnt[b0 + k];
    }
  }
  if (t == kLcScanThreads - 1) *total = sums[t];
}

__global__ void __launch_bounds__(kBlock) k_lc_tile_write(const float* __restrict__ Q, int N, int L, float tol,
                                                          const int* __restrict__ off, int* __restrict__ idx) {
  __shared__ int wsum[kBlock / 64];
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  int f[4];
  const int n = lc_flags4(Q, N, L, tol, blockIdx.x, t, f);
  // inclusive prefix over the wave, then the earlier waves' totals
  int inc = n;
  for (int o = 1; o < 64; o <<= 1) {
    const int v = __shfl_up(inc, o, 64);
    if (lane >= o) inc += v;
  }
  if (lane == 63) wsum[w] = inc;
  __syncthreads();
  int pos = off[blockIdx.x] + inc - n;
  for (int k = 0; k < w; ++k) pos += wsum[k];
  const int base = blockIdx.x * kLcTile + 4 * t;
#pragma unroll
  for (int j = 0; j < 4; ++j)
    if (f[j]) idx[pos++] = base + j;
}

}  // namespace

// the device address of this device's `scratch`, looked up once per device
// with that device current (hipGetSymbolAddress; the array's host-side name
// is not a device address)
static int* lc_scratch_ptr() {
  constexpr int kMaxDevices = 64;
  static std::mutex mu;
  static int* ptrs[kMaxDevices] = {};
  int dev = 0;
  DCRF_HIP(hipGetDevice(&dev));
  DCRF_REQUIRE(dev >= 0 && dev < kMaxDevices, "less_confident_pixels (no allocation): device id out of range");
  std::lock_guard<std::mutex> g(mu);
  if (!ptrs[dev]) {
    void* p = nullptr;
    DCRF_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    ptrs[dev] = (int*)p;
  }
  return ptrs[dev];
}

// Q (L x N) and idx (N) are device pointers; returns the count.  Launches on
// stream s and waits for it.
int less_confident_pixels_noalloc_dev(const float* Q, int N, int L, float tol, int* idx, hipStream_t s) {
  DCRF_REQUIRE((long long)N <= kLcMaxPixels,
               "dcrf_less_confident_pixels_noalloc: N exceeds the 16,777,216 pixels its fixed scratch holds");
  if (N == 0) return 0;
  int* sc = lc_scratch_ptr();
  int* cnt = sc;
  int* off = sc + kLcMaxTiles;
  int* total = sc + 2 * kLcMaxTiles;
  const int ntiles = (N + kLcTile - 1) / kLcTile;
  k_lc_tile_count<<<ntiles, kBlock, 0, s>>>(Q, N, L, tol, cnt);
  k_lc_tile_scan<<<1, kLcScanThreads, 0, s>>>(cnt, ntiles, off, total);
  k_lc_tile_write<<<ntiles, kBlock, 0, s>>>(Q, N, L, tol, off, idx);
  check_launch("less_confident_pixels (no allocation)");
  int n = 0;
  DCRF_HIP(hipMemcpyAsync(&n, total, sizeof(int), hipMemcpyDeviceToHost, s));
  DCRF_HIP(hipStreamSynchronize(s));
  return n;
}

}  // namespace dcrf
