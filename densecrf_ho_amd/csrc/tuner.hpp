// Launch-shape autotuner.
//
// Kernels with grid-stride loops and no cross-block reduction compute the same
// bits for any grid size, but their speed depends on it (occupancy, L2 reuse
// per XCD, tail effects).  For such a kernel the first launch at a given
// problem size times every candidate grid on the live data with HIP events
// (each candidate once to warm up, once timed) and keeps the fastest; later
// launches of that (kernel, size class) reuse the choice.  The launch callback
// must be idempotent: running it again gives the same output.
//
// Off when DCRF_AUTOTUNE=0 (the first candidate is used) and while a stream is
// being captured into a graph.  report() lists every decision.
#pragma once

#include <functional>
#include <map>
#include <mutex>
#include <sstream>
#include <string>
#include <vector>

#include "common.hpp"

namespace dcrf {

class Tuner {
 public:
  static Tuner& get() {
    static Tuner t;
    return t;
  }

  // grid to launch `kernel` with for `work` units (size class: power of two)
  unsigned pick(const char* kernel, long work, const std::vector<unsigned>& cands, hipStream_t s,
                const std::function<void(unsigned)>& launch) {
    if (cands.empty()) return 1;
    if (!enabled_ || cands.size() == 1) return cands[0];
    int cls = 0;
    while ((1L << cls) < work && cls < 62) ++cls;
    const std::string key = std::string(kernel) + "@" + std::to_string(cls);
    {
      std::lock_guard<std::mutex> g(mu_);
      auto it = best_.find(key);
      if (it != best_.end()) return it->second.grid;
    }
    // no timing inside a graph capture: a launch there records, it does not run
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(s, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone) return cands[0];
    hipEvent_t a = nullptr, b = nullptr;
    DCRF_HIP(hipEventCreate(&a));
    DCRF_HIP(hipEventCreate(&b));
    unsigned best = cands[0];
    float best_ms = 1e30f;
    std::ostringstream trail;
    for (unsigned c : cands) {
      launch(c);  // warm-up
      DCRF_HIP(hipEventRecord(a, s));
      launch(c);
      DCRF_HIP(hipEventRecord(b, s));
      DCRF_HIP(hipEventSynchronize(b));
      float ms = 0.f;
      DCRF_HIP(hipEventElapsedTime(&ms, a, b));
      trail << " " << c << ":" << ms;
      if (ms < best_ms) {
        best_ms = ms;
        best = c;
      }
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    std::lock_guard<std::mutex> g(mu_);
    best_[key] = Choice{best, best_ms, trail.str()};
    return best;
  }

  // the tuned grid if this (kernel, size class) was tuned, else `fallback`
  unsigned cached(const char* kernel, long work, unsigned fallback) {
    if (!enabled_) return fallback;
    int cls = 0;
    while ((1L << cls) < work && cls < 62) ++cls;
    std::lock_guard<std::mutex> g(mu_);
    auto it = best_.find(std::string(kernel) + "@" + std::to_string(cls));
    return it == best_.end() ? fallback : it->second.grid;
  }

  // "kernel@size_class grid ms | candidate:ms ..." per line
  std::string report() {
    std::lock_guard<std::mutex> g(mu_);
    std::ostringstream o;
    for (auto& kv : best_) o << kv.first << " " << kv.second.grid << " " << kv.second.ms << " |" << kv.second.trail << "\n";
    return o.str();
  }
  void clear() {
    std::lock_guard<std::mutex> g(mu_);
    best_.clear();
  }
  void set_enabled(bool on) { enabled_ = on; }

 private:
  Tuner() {
    const char* e = std::getenv("DCRF_AUTOTUNE");
    enabled_ = !(e && e[0] == '0');
  }
  struct Choice {
    unsigned grid;
    float ms;
    std::string trail;
  };
  std::mutex mu_;
  std::map<std::string, Choice> best_;
  bool enabled_ = true;
};

// candidate grids for a grid-stride launch whose full grid is `full`: the full
// grid and caps at 16384, 8192, 4096 and 2048 blocks (those below it)
inline std::vector<unsigned> grid_candidates(unsigned full) {
  std::vector<unsigned> c{full};
  for (unsigned cap : {16384u, 8192u, 4096u, 2048u})
    if (cap < full) c.push_back(cap);
  return c;
}

}  // namespace dcrf
