// Shared plumbing of the dense-CRF engine: error reporting, device buffers,
// launch helpers and deterministic FP64 reductions.
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdint>
#include <cstdio>
#include <map>
#include <mutex>
#include <stdexcept>
#include <cstring>
#include <string>
#include <vector>

namespace dcrf {

// ---------------------------------------------------------------------------
// errors: every C entry point catches dcrf::Error (and std::exception) and
// turns it into a status code plus a thread-local message.
// ---------------------------------------------------------------------------
struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

enum Status : int {
  kOk = 0,
  kInvalidArgument = 1,
  kOutOfBounds = 2,     // "Value out of bounds" of the ordered filter
  kMaxvalNegative = 3,  // "Maxval < 0, skipping" of rescale
  kHipError = 4,
  kInternal = 5,
};

void set_last_error(const std::string& msg);

#define DCRF_HIP(call)                                                                     \
  do {                                                                                     \
    hipError_t e_ = (call);                                                                \
    if (e_ != hipSuccess)                                                                  \
      throw ::dcrf::Error(::dcrf::kHipError, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

#define DCRF_REQUIRE(cond, msg)                                         \
  do {                                                                  \
    if (!(cond)) throw ::dcrf::Error(::dcrf::kInvalidArgument, (msg)); \
  } while (0)

inline void check_launch(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) throw Error(kHipError, std::string("launch of ") + what + ": " + hipGetErrorString(e));
}

// ---------------------------------------------------------------------------
// kernel timing: HIP events recorded around launches on their own stream,
// resolved when the report is read.  Off unless enabled.
// ---------------------------------------------------------------------------
struct Profiler {
  struct Rec {
    std::string name;
    hipEvent_t a, b;
  };
  std::atomic<bool> on{false};
  std::string prefix;  // when not empty, only launches whose timer name starts with it are timed
  std::mutex mu;
  std::vector<Rec> recs;
  std::map<std::string, std::pair<long, double>> done;  // name -> (count, ms)
  static Profiler& get() {
    static Profiler p;
    return p;
  }
  void add(const std::string& n, hipEvent_t a, hipEvent_t b) {
    std::lock_guard<std::mutex> g(mu);
    recs.push_back({n, a, b});
    if (recs.size() > 4096) resolve_locked();
  }
  void resolve_locked() {
    for (auto& r : recs) {
      float ms = 0.f;
      if (hipEventSynchronize(r.b) == hipSuccess && hipEventElapsedTime(&ms, r.a, r.b) == hipSuccess) {
        auto& d = done[r.name];
        d.first += 1;
        d.second += ms;
      }
      (void)hipEventDestroy(r.a);
      (void)hipEventDestroy(r.b);
    }
    recs.clear();
  }
  // a counted quantity reported like a kernel: (calls, total) in place of
  // (launches, ms)
  void count(const std::string& n, double v) {
    std::lock_guard<std::mutex> g(mu);
    auto& d = done[n];
    d.first += 1;
    d.second += v;
  }
  void clear() {
    std::lock_guard<std::mutex> g(mu);
    resolve_locked();
    done.clear();
  }
  std::string report() {
    std::lock_guard<std::mutex> g(mu);
    resolve_locked();
    std::string out;
    char line[256];
    for (auto& kv : done) {
      snprintf(line, sizeof(line), "%s %ld %.6f\n", kv.first.c_str(), kv.second.first, kv.second.second);
      out += line;
    }
    return out;
  }
};

inline bool prof_on() { return Profiler::get().on.load(std::memory_order_relaxed); }

// times the launches issued during its lifetime on stream s
struct KTimer {
  hipEvent_t a = nullptr;
  hipStream_t s;
  std::string name;
  KTimer(const char* n, hipStream_t st) : s(st) {
    if (!Profiler::get().on.load(std::memory_order_relaxed)) return;
    const std::string& pf = Profiler::get().prefix;
    if (!pf.empty() && std::strncmp(n, pf.c_str(), pf.size()) != 0) return;
    name = n;
    if (hipEventCreate(&a) != hipSuccess || hipEventRecord(a, s) != hipSuccess) a = nullptr;
  }
  KTimer(const std::string& n, hipStream_t st) : KTimer(n.c_str(), st) {}
  ~KTimer() {
    if (!a) return;
    hipEvent_t b;
    if (hipEventCreate(&b) == hipSuccess && hipEventRecord(b, s) == hipSuccess) Profiler::get().add(name, a, b);
  }
};

// ---------------------------------------------------------------------------
// device buffer (owning, typed, resizable; never shrinks)
// ---------------------------------------------------------------------------
template <class T>
struct DBuf {
  T* p = nullptr;
  size_t n = 0;
  DBuf() = default;
  explicit DBuf(size_t count) { resize(count); }
  DBuf(const DBuf&) = delete;
  DBuf& operator=(const DBuf&) = delete;
  DBuf(DBuf&& o) noexcept : p(o.p), n(o.n) {
    o.p = nullptr;
    o.n = 0;
  }
  DBuf& operator=(DBuf&& o) noexcept {
    if (this != &o) {
      release();
      p = o.p;
      n = o.n;
      o.p = nullptr;
      o.n = 0;
    }
    return *this;
  }
  ~DBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  void resize(size_t count) {
    if (count <= n && p) return;
    release();
    if (count == 0) return;
    DCRF_HIP(hipMalloc(&p, count * sizeof(T)));
    n = count;
  }
  T* get() const { return p; }
  size_t bytes() const { return n * sizeof(T); }
};

// Pinned, coherent host memory that kernels write directly (small status
// words the host reads after a stream synchronisation, with no copy launch).
template <class T>
struct HBuf {
  T* p = nullptr;
  size_t n = 0;
  HBuf() = default;
  HBuf(const HBuf&) = delete;
  HBuf& operator=(const HBuf&) = delete;
  ~HBuf() {
    if (p) (void)hipHostFree(p);
  }
  void resize(size_t count) {
    if (count <= n && p) return;
    if (p) (void)hipHostFree(p);
    p = nullptr;
    n = 0;
    DCRF_HIP(hipHostMalloc((void**)&p, std::max<size_t>(count, 1) * sizeof(T), hipHostMallocCoherent));
    n = count;
  }
  T* get() const { return p; }
};

// ---------------------------------------------------------------------------
// launch helpers
// ---------------------------------------------------------------------------
// Threads per block of most kernels.  Kernels that are always launched with
// kBlock threads index with the constant: blockDim.x is a 2-byte read of the
// dispatch packet, which gfx950 serves through the vector memory path, so it
// would put one dependent load ahead of every address the kernel forms.
constexpr int kBlock = 256;
inline unsigned grid_for(size_t work, int block = kBlock, size_t cap = 1u << 30) {
  size_t g = (work + block - 1) / block;
  if (g == 0) g = 1;
  if (g > cap) g = cap;
  return (unsigned)g;
}

// Division by a runtime constant with one high multiply (valid for n < 2^31).
// d = 2^s uses a shift; otherwise m = floor(2^(32+s)/d) + 1 with
// s = floor(log2 d), and q = umulhi(n, m) >> s.
struct FastDiv {
  uint32_t d = 1, m = 0, s = 0;
  FastDiv() = default;
  explicit FastDiv(uint32_t dd) : d(dd) {
    s = 31 - __builtin_clz(dd);
    m = ((1u << s) == dd) ? 0u : (uint32_t)(((uint64_t)1 << (32 + s)) / dd + 1);
  }
  __device__ __host__ inline uint32_t div(uint32_t n) const {
#ifdef __HIP_DEVICE_COMPILE__
    return m == 0 ? (n >> s) : (__umulhi(n, m) >> s);
#else
    return m == 0 ? (n >> s) : (uint32_t)(((uint64_t)n * m) >> 32) >> s;
#endif
  }
};

// XCD-aware block order.  Workgroups are dispatched round-robin over the 8
// XCDs (each with its own L2); this maps them so that XCD x works on one
// contiguous range of logical blocks, keeping neighbouring rows (lattice
// points that share pixels, pixels that share vertices) in one L2.
constexpr unsigned kNumXcd = 8;
__device__ inline unsigned xcd_block() {
  const unsigned g = gridDim.x, b = blockIdx.x;
  if (g < kNumXcd) return b;
  const unsigned per = g / kNumXcd, rem = g % kNumXcd;
  const unsigned x = b % kNumXcd, i = b / kNumXcd;
  return x * per + min(x, rem) + i;
}

// Lanes of one wave exchanging values through LDS: the wave's earlier LDS
// writes are visible to its later reads, with no barrier across the block
// (other waves go on with their own rows).
__device__ inline void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Number of blocks used by grid-stride reduction kernels: a fixed count so the
// combination order of partial sums (and therefore the result) is
// deterministic for a given problem size.
constexpr int kRedBlocks = 1024;

// ---------------------------------------------------------------------------
// block-level FP64 sum (wave shuffles then LDS); result valid in thread 0
// ---------------------------------------------------------------------------
__device__ inline double wave_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  return v;
}

template <int NV>
__device__ inline void block_sum(double (&v)[NV]) {
  __shared__ double red[NV][kBlock / 64];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int i = 0; i < NV; ++i) {
    double s = wave_sum(v[i]);
    if (lane == 0) red[i][wid] = s;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int i = 0; i < NV; ++i) {
      double s = 0;
      for (int w = 0; w < kBlock / 64; ++w) s += red[i][w];
      v[i] = s;
    }
  }
  __syncthreads();
}

// sum `nparts` per-block partials (stride `stride` between sums of the same
// quantity) in index order; one block, result written to out[0..NV)
template <int NV>
__global__ void finalize_sums(const double* __restrict__ parts, int nparts, double* __restrict__ out) {
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

}  // namespace dcrf
