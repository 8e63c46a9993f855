// C ABI (include/densecrf_ho_amd.h): status codes, host/device staging and
// the thread-local error message.
#include <cstring>
#include <string>

#include "../../include/densecrf_ho_amd.h"
#include "../../include/densecrf_ho_amd_io.hpp"
#include "batch.hpp"
#include "tuner.hpp"
#include "crf.hpp"
#include "crf_kernels.hpp"

struct dcrf_lattice_s {
  hipStream_t s = nullptr;
  std::unique_ptr<dcrf::Lattice> lat;
  dcrf::DBuf<float> minmax;
  ~dcrf_lattice_s() {
    if (s) {
      (void)hipStreamSynchronize(s);
      lat.reset();
      (void)hipStreamDestroy(s);
    }
  }
};
struct dcrf_crf_s {
  std::unique_ptr<dcrf::CRF> crf;
};
struct dcrf_batch_s {
  std::unique_ptr<dcrf::Batch> b;
  std::vector<int> W, H;  // image sizes when given (2-D kernels)
};
struct dcrf_pairwise_s {
  hipStream_t s = nullptr;
  dcrf::Potential p;
  dcrf::DBuf<float> mm0;
  ~dcrf_pairwise_s() {
    if (s) {
      (void)hipStreamSynchronize(s);
      p.lat.reset();
      p.norm.release();
      mm0.release();
      (void)hipStreamDestroy(s);
    }
  }
};

namespace dcrf {

static thread_local std::string g_last_error;
void set_last_error(const std::string& msg) { g_last_error = msg; }

template <class F>
static int guarded(F&& f) {
  try {
    f();
    return kOk;
  } catch (const Error& e) {
    set_last_error(e.what());
    return e.code;
  } catch (const std::bad_alloc&) {
    set_last_error("out of host memory");
    return kInternal;
  } catch (const std::exception& e) {
    set_last_error(e.what());
    return kInternal;
  }
}

// device view of a caller buffer: either the caller's device pointer or a
// staged copy of a host buffer
template <class T>
struct Staged {
  T* d = nullptr;
  DBuf<T> own;
  T* host = nullptr;
  size_t n = 0;
  hipStream_t s;
  Staged(const T* p, size_t count, bool on_device, hipStream_t st, bool copy_in) : n(count), s(st) {
    if (on_device) {
      d = const_cast<T*>(p);
      return;
    }
    own.resize(count);
    d = own.get();
    host = const_cast<T*>(p);
    if (copy_in && count > 0) DCRF_HIP(hipMemcpyAsync(d, p, count * sizeof(T), hipMemcpyHostToDevice, s));
  }
  void copy_out() {
    if (host && n > 0) DCRF_HIP(hipMemcpyAsync(host, d, n * sizeof(T), hipMemcpyDeviceToHost, s));
  }
};

namespace {
__global__ void __launch_bounds__(kBlock) k_features_2d(float* __restrict__ f, int W, int H, float sx, float sy, const unsigned char* im,
                              float sr, float sg, float sb) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= W * H) return;
  const int x = i % W, y = i / W;
  const int d = im ? 5 : 2;
  f[(size_t)i * d + 0] = (float)x / sx;
  f[(size_t)i * d + 1] = (float)y / sy;
  if (im) {
    f[(size_t)i * d + 2] = (float)im[(size_t)i * 3 + 0] / sr;
    f[(size_t)i * d + 3] = (float)im[(size_t)i * 3 + 1] / sg;
    f[(size_t)i * d + 4] = (float)im[(size_t)i * 3 + 2] / sb;
  }
}

__global__ void __launch_bounds__(kBlock) k_minmax(const float* __restrict__ a, size_t n, float* __restrict__ parts) {
  __shared__ float smn[kBlock], smx[kBlock];
  float mn = INFINITY, mx = -INFINITY;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    const float v = a[i];
    mn = fminf(mn, v);
    mx = fmaxf(mx, v);
    if (v != v) mn = mx = v;  // NaN propagates
  }
  smn[threadIdx.x] = mn;
  smx[threadIdx.x] = mx;
  __syncthreads();
  for (int o = kBlock / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) {
      smn[threadIdx.x] = fminf(smn[threadIdx.x], smn[threadIdx.x + o]);
      smx[threadIdx.x] = fmaxf(smx[threadIdx.x], smx[threadIdx.x + o]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    parts[2 * blockIdx.x] = smn[0];
    parts[2 * blockIdx.x + 1] = smx[0];
  }
}
}  // namespace

// host-side [min, max] of a device array (bounds checks of the ordered filter)
static void minmax_host(const float* d, size_t n, hipStream_t s, float* mn, float* mx) {
  const int nb = 256;
  DBuf<float> parts(2 * nb);
  k_minmax<<<nb, kBlock, 0, s>>>(d, n, parts.get());
  check_launch("k_minmax");
  std::vector<float> h(2 * nb);
  DCRF_HIP(hipMemcpyAsync(h.data(), parts.get(), h.size() * sizeof(float), hipMemcpyDeviceToHost, s));
  DCRF_HIP(hipStreamSynchronize(s));
  float a = INFINITY, b = -INFINITY;
  bool nan = false;
  for (int i = 0; i < nb; ++i) {
    if (h[2 * i] != h[2 * i] || h[2 * i + 1] != h[2 * i + 1]) nan = true;
    a = std::min(a, h[2 * i]);
    b = std::max(b, h[2 * i + 1]);
  }
  *mn = nan ? NAN : a;
  *mx = nan ? NAN : b;
}

// PairwisePotential::apply_upper_minus_lower_ord throws when Q leaves [0,1]
// (src/pairwise.cpp:218-219).  NaN passes the reference's test and is
// rejected here as out of bounds too.
void check_unit_interval(const float* d, size_t n, hipStream_t s) {
  float mn, mx;
  minmax_host(d, n, s, &mn, &mx);
  if (!(mx <= 1.f) || !(mn >= 0.f)) throw Error(kOutOfBounds, "Value out of bounds");
}

// pI of the restricted ordered filter: strictly ascending pixel indices in
// [0, N) (the reference walks the pixels and pI together and needs that order)
void check_pixel_list(const int* d_pI, int nI, int N, bool store, hipStream_t s) {
  DCRF_REQUIRE(nI >= 0 && nI <= N, "restricted ordered filter: bad pixel count");
  // the reference reads pI[0] when it stores (src/permutohedral.cpp:724-730): an
  // empty list has no defined result there, so it is refused
  DCRF_REQUIRE(nI > 0 || !store, "restricted ordered filter: a call with store needs at least one pixel");
  if (nI == 0) return;
  std::vector<int> h(nI);
  DCRF_HIP(hipMemcpyAsync(h.data(), d_pI, nI * sizeof(int), hipMemcpyDeviceToHost, s));
  DCRF_HIP(hipStreamSynchronize(s));
  for (int i = 0; i < nI; ++i)
    DCRF_REQUIRE(h[i] >= 0 && h[i] < N && (i == 0 || h[i] > h[i - 1]),
                 "restricted ordered filter: pixel indices must be strictly ascending and in [0, N)");
}

// ind of the brute-force sorted filters: every entry a pixel index in [0, N)
void check_index_matrix(const int* d_ind, size_t n, int N, hipStream_t s) {
  std::vector<int> h(n);
  DCRF_HIP(hipMemcpyAsync(h.data(), d_ind, n * sizeof(int), hipMemcpyDeviceToHost, s));
  DCRF_HIP(hipStreamSynchronize(s));
  for (size_t i = 0; i < n; ++i) DCRF_REQUIRE(h[i] >= 0 && h[i] < N, "brute-force filter: ind entry out of range");
}

}  // namespace dcrf

using namespace dcrf;

extern "C" {

const char* dcrf_last_error(void) { return g_last_error.c_str(); }
const char* dcrf_version(void) { return "densecrf_ho_amd 0.1 (gfx950)"; }

int dcrf_set_reference_path(int portable) {
  return guarded([&] {
    DCRF_REQUIRE(portable == 0 || portable == 1, "dcrf_set_reference_path: 0 (SSE2 build) or 1 (portable build)");
    reference_path().store(portable);
  });
}
int dcrf_reference_path(void) { return reference_path().load(); }

// ---- lattice ------------------------------------------------------------------
int dcrf_lattice_create(const float* features, int N, int d, int on_device, dcrf_lattice_t* out) {
  return guarded([&] {
    DCRF_REQUIRE(features && out && N > 0 && d > 0, "dcrf_lattice_create: bad arguments");
    auto h = std::make_unique<dcrf_lattice_s>();
    DCRF_HIP(hipStreamCreateWithFlags(&h->s, hipStreamNonBlocking));
    h->lat = std::make_unique<Lattice>(h->s);
    Staged<float> f(features, (size_t)N * d, on_device != 0, h->s, true);
    h->lat->build(f.d, N, d);
    h->minmax.resize(2);
    DCRF_HIP(hipMemsetAsync(h->minmax.get(), 0, 2 * sizeof(float), h->s));
    DCRF_HIP(hipStreamSynchronize(h->s));
    *out = h.release();
  });
}
int dcrf_lattice_destroy(dcrf_lattice_t lat) {
  return guarded([&] { delete lat; });
}
int dcrf_lattice_info(dcrf_lattice_t lat, int* M, int* Npad) {
  return guarded([&] {
    DCRF_REQUIRE(lat, "dcrf_lattice_info: null handle");
    if (M) *M = lat->lat->M();
    if (Npad) *Npad = lat->lat->Npad();
  });
}
int dcrf_lattice_export(dcrf_lattice_t lat, int* offset, int* rank, float* barycentric, short* keys, int* neighbors) {
  return guarded([&] {
    DCRF_REQUIRE(lat, "dcrf_lattice_export: null handle");
    lat->lat->export_tables(offset, rank, barycentric, keys, neighbors);
  });
}
int dcrf_lattice_compute(dcrf_lattice_t lat, const float* in, float* out, int V, int reverse, int on_device) {
  return guarded([&] {
    DCRF_REQUIRE(lat && in && out && V > 0, "dcrf_lattice_compute: bad arguments");
    const size_t n = (size_t)lat->lat->N() * V;
    hipStream_t s = lat->s;
    Staged<float> a(in, n, on_device != 0, s, true);
    Staged<float> b(out, n, on_device != 0, s, false);
    SplatArgs sa;
    sa.kind = SplatIn::kPlain;
    sa.in = a.d;
    SliceArgs so;
    so.out = b.d;
    lat->lat->filter(sa, so, V, reverse != 0);
    b.copy_out();
    DCRF_HIP(hipStreamSynchronize(s));
  });
}
int dcrf_lattice_compute_upper_minus_lower_ord(dcrf_lattice_t lat, const float* Q, float* out, int V, int on_device) {
  return guarded([&] {
    DCRF_REQUIRE(lat && Q && out && V > 0, "dcrf_lattice_compute_upper_minus_lower_ord: bad arguments");
    const size_t n = (size_t)lat->lat->N() * V;
    hipStream_t s = lat->s;
    Staged<float> a(Q, n, on_device != 0, s, true);
    Staged<float> b(out, n, on_device != 0, s, false);
    check_unit_interval(a.d, n, s);
    SliceArgs so;
    so.out = b.d;
    lat->lat->ordered(a.d, lat->minmax.get(), V, so);
    b.copy_out();
    DCRF_HIP(hipStreamSynchronize(s));
  });
}

int dcrf_lattice_compute_upper_minus_lower_ord_restricted(dcrf_lattice_t lat, float* rout, const float* rQ,
                                                           const int* pI, int nI, const float* Q, int V, int store,
                                                           int on_device) {
  return guarded([&] {
    DCRF_REQUIRE(lat && V > 0 && nI >= 0 && (nI == 0 || (rout && rQ && pI)) && (!store || Q),
                 "dcrf_lattice_compute_upper_minus_lower_ord_restricted: bad arguments");
    const int N = lat->lat->N();
    hipStream_t s = lat->s;
    const size_t nr = (size_t)nI * V;
    Staged<int> ix(pI, nI, on_device != 0, s, true);
    Staged<float> a(rQ, nr, on_device != 0, s, true);
    Staged<float> o(rout, nr, on_device != 0, s, true);  // added to
    Staged<float> e(store ? Q : nullptr, store ? (size_t)N * V : 0, on_device != 0, s, store != 0);
    check_pixel_list(ix.d, nI, N, store != 0, s);
    if (nr) check_unit_interval(a.d, nr, s);
    if (store) check_unit_interval(e.d, (size_t)N * V, s);
    lat->lat->ordered_restricted(a.d, ix.d, nI, e.d, V, store != 0, o.d);
    o.copy_out();
    DCRF_HIP(hipStreamSynchronize(s));
  });
}

int dcrf_lattice_compute_upper_minus_lower_dc(dcrf_lattice_t lat, float* out, int low, int middle_low,
                                              int middle_high, int high, int on_device) {
  return guarded([&] {
    DCRF_REQUIRE(lat && out, "dcrf_lattice_compute_upper_minus_lower_dc: bad arguments");
    const int N = lat->lat->N();
    DCRF_REQUIRE(0 <= low && low <= middle_low && middle_low <= middle_high && middle_high <= high && high <= N,
                 "dcrf_lattice_compute_upper_minus_lower_dc: ranges must satisfy 0 <= low <= middle_low <= "
                 "middle_high <= high <= N");
    Staged<float> o(out, N, on_device != 0, lat->s, true);  // added to
    lat->lat->compute_upper_minus_lower_dc(o.d, low, middle_low, middle_high, high);
    o.copy_out();
    DCRF_HIP(hipStreamSynchronize(lat->s));
  });
}
int dcrf_lattice_gradient(dcrf_lattice_t lat, float* df, const float* a, const float* b, int V, int on_device) {
  return guarded([&] {
    DCRF_REQUIRE(lat && df && a && b && V > 0, "dcrf_lattice_gradient: bad arguments");
    const int N = lat->lat->N(), d = lat->lat->d();
    hipStream_t s = lat->s;
    Staged<float> x(a, (size_t)N * V, on_device != 0, s, true);
    Staged<float> y(b, (size_t)N * V, on_device != 0, s, true);
    Staged<float> o(df, (size_t)N * d, on_device != 0, s, false);
    lat->lat->gradient(x.d, y.d, V, o.d);
    o.copy_out();
    DCRF_HIP(hipStreamSynchronize(s));
  });
}

// ---- pairwise potential ---------------------------------------------------------
int dcrf_pairwise_create(const float* features, int N, int d, float potts_w, int ntype, int on_device,
                         dcrf_pairwise_t* out) {
  return guarded([&] {
    DCRF_REQUIRE(features && out && N > 0 && d > 0, "dcrf_pairwise_create: bad arguments");
    auto h = std::make_unique<dcrf_pairwise_s>();
    DCRF_HIP(hipStreamCreateWithFlags(&h->s, hipStreamNonBlocking));
    Staged<float> f(features, (size_t)N * d, on_device != 0, h->s, true);
    potential_init(h->p, f.d, N, d, potts_w, ntype, h->s);
    h->p.ktype = kConstKernel;  // PairwisePotential's default (include/pairwise.h:71)
    h->mm0.resize(2);
    DCRF_HIP(hipMemsetAsync(h->mm0.get(), 0, 2 * sizeof(float), h->s));
    DCRF_HIP(hipStreamSynchronize(h->s));
    *out = h.release();
  });
}
int dcrf_pairwise_destroy(dcrf_pairwise_t pw) {
  return guarded([&] { delete pw; });
}
int dcrf_pairwise_apply(dcrf_pairwise_t pw, const float* in, float* out, int L, int transpose, int on_device) {
  return guarded([&] {
    DCRF_REQUIRE(pw && in && out && L > 0, "dcrf_pairwise_apply: bad arguments");
    const size_t n = (size_t)pw->p.lat->N() * L;
    Staged<float> a(in, n, on_device != 0, pw->s, true);
    Staged<float> b(out, n, on_device != 0, pw->s, false);
    potential_apply(pw->p, a.d, b.d, L, transpose != 0, false);
    b.copy_out();
    DCRF_HIP(hipStreamSynchronize(pw->s));
  });
}
int dcrf_pairwise_apply_upper_minus_lower_ord(dcrf_pairwise_t pw, const float* Q, float* out, int L, int on_device) {
  return guarded([&] {
    DCRF_REQUIRE(pw && Q && out && L > 0, "dcrf_pairwise_apply_upper_minus_lower_ord: bad arguments");
    const size_t n = (size_t)pw->p.lat->N() * L;
    Staged<float> a(Q, n, on_device != 0, pw->s, true);
    Staged<float> b(out, n, on_device != 0, pw->s, false);
    check_unit_interval(a.d, n, pw->s);
    potential_apply_ord(pw->p, a.d, pw->mm0.get(), b.d, L);
    b.copy_out();
    DCRF_HIP(hipStreamSynchronize(pw->s));
  });
}

static void apply_ord_restricted(Potential& p, hipStream_t s, float* rout, const float* rQ, const int* pI, int nI,
                                 const float* Q, int L, int store, int on_device) {
  const int N = p.lat->N();
  const size_t nr = (size_t)nI * L;
  Staged<int> ix(pI, nI, on_device != 0, s, true);
  Staged<float> a(rQ, nr, on_device != 0, s, true);
  Staged<float> o(rout, nr, on_device != 0, s, false);
  Staged<float> e(store ? Q : nullptr, store ? (size_t)N * L : 0, on_device != 0, s, store != 0);
  check_pixel_list(ix.d, nI, N, store != 0, s);
  if (nr) check_unit_interval(a.d, nr, s);
  if (store) check_unit_interval(e.d, (size_t)N * L, s);
  potential_apply_ord_restricted(p, o.d, a.d, ix.d, nI, e.d, L, store != 0);
  o.copy_out();
  DCRF_HIP(hipStreamSynchronize(s));
}
int dcrf_pairwise_apply_upper_minus_lower_ord_restricted(dcrf_pairwise_t pw, float* rout, const float* rQ,
                                                         const int* pI, int nI, const float* Q, int L, int store,
                                                         int on_device) {
  return guarded([&] {
    DCRF_REQUIRE(pw && L > 0 && nI >= 0 && (nI == 0 || (rout && rQ && pI)) && (!store || Q),
                 "dcrf_pairwise_apply_upper_minus_lower_ord_restricted: bad arguments");
    apply_ord_restricted(pw->p, pw->s, rout, rQ, pI, nI, Q, L, store, on_device);
  });
}

// kind: 0 apply_bf(Q), 1 upper_minus_lower_bf(ind), 2 _bf_ord(ind, Q), 3 _dc(ind)
static void apply_bf_common(Potential& p, hipStream_t s, int kind, const int* ind, const float* Q, float* out, int L,
                            int on_device) {
  const int N = p.lat->N();
  const size_t n = (size_t)N * L;
  const bool uses_q = kind == 0 || kind == 2;
  Staged<int> ix(kind ? ind : nullptr, kind ? n : 0, on_device != 0, s, kind != 0);
  Staged<float> q(uses_q ? Q : nullptr, uses_q ? n : 0, on_device != 0, s, uses_q);
  Staged<float> o(out, n, on_device != 0, s, false);
  if (kind) check_index_matrix(ix.d, n, N, s);
  if (kind == 0)
    potential_apply_bf(p, q.d, o.d, L);
  else if (kind == 3)
    potential_apply_uml_dc(p, ix.d, o.d, L);
  else
    potential_apply_uml_bf(p, ix.d, q.d, o.d, L, kind == 2);
  o.copy_out();
  DCRF_HIP(hipStreamSynchronize(s));
}
int dcrf_pairwise_apply_bf(dcrf_pairwise_t pw, const float* Q, float* out, int L, int on_device) {
  return guarded([&] {
    DCRF_REQUIRE(pw && Q && out && L > 0, "dcrf_pairwise_apply_bf: bad arguments");
    apply_bf_common(pw->p, pw->s, 0, nullptr, Q, out, L, on_device);
  });
}
int dcrf_pairwise_apply_upper_minus_lower_bf(dcrf_pairwise_t pw, const int* ind, float* out, int L, int on_device) {
  return guarded([&] {
    DCRF_REQUIRE(pw && ind && out && L > 0, "dcrf_pairwise_apply_upper_minus_lower_bf: bad arguments");
    apply_bf_common(pw->p, pw->s, 1, ind, nullptr, out, L, on_device);
  });
}
int dcrf_pairwise_apply_upper_minus_lower_dc(dcrf_pairwise_t pw, const int* ind, float* out, int L, int on_device) {
  return guarded([&] {
    DCRF_REQUIRE(pw && ind && out && L > 0, "dcrf_pairwise_apply_upper_minus_lower_dc: bad arguments");
    apply_bf_common(pw->p, pw->s, 3, ind, nullptr, out, L, on_device);
  });
}
int dcrf_pairwise_apply_upper_minus_lower_bf_ord(dcrf_pairwise_t pw, const int* ind, const float* Q, float* out,
                                                 int L, int on_device) {
  return guarded([&] {
    DCRF_REQUIRE(pw && ind && Q && out && L > 0, "dcrf_pairwise_apply_upper_minus_lower_bf_ord: bad arguments");
    apply_bf_common(pw->p, pw->s, 2, ind, Q, out, L, on_device);
  });
}

int dcrf_pairwise_set_compatibility(dcrf_pairwise_t pw, int compat_type, const float* params, int L) {
  return guarded([&] {
    DCRF_REQUIRE(pw && params && compat_type >= DCRF_COMPAT_POTTS && compat_type <= DCRF_COMPAT_MATRIX &&
                     (compat_type == DCRF_COMPAT_POTTS || L > 0),
                 "dcrf_pairwise_set_compatibility: bad arguments");
    potential_set_compat(pw->p, compat_type, params, L, pw->s);
  });
}
int dcrf_pairwise_set_kernel_type(dcrf_pairwise_t pw, int ktype) {
  return guarded([&] {
    DCRF_REQUIRE(pw && ktype >= 0 && ktype <= 2, "dcrf_pairwise_set_kernel_type: bad arguments");
    DCRF_REQUIRE(pw->p.kparams.empty(), "dcrf_pairwise_set_kernel_type: set before any kernel parameters");
    pw->p.ktype = ktype;
  });
}
int dcrf_pairwise_kernel_nparams(dcrf_pairwise_t pw) { return pw ? potential_kernel_nparams(pw->p) : -1; }
int dcrf_pairwise_kernel_parameters(dcrf_pairwise_t pw, float* out) {
  return guarded([&] {
    DCRF_REQUIRE(pw, "dcrf_pairwise_kernel_parameters: bad arguments");
    const std::vector<float> v = potential_kernel_parameters(pw->p);
    DCRF_REQUIRE(out || v.empty(), "dcrf_pairwise_kernel_parameters: null output");
    std::copy(v.begin(), v.end(), out);
  });
}
int dcrf_pairwise_set_kernel_parameters(dcrf_pairwise_t pw, const float* v, int n) {
  return guarded([&] {
    DCRF_REQUIRE(pw && (v || n == 0), "dcrf_pairwise_set_kernel_parameters: bad arguments");
    potential_set_kernel_parameters(pw->p, v, n, pw->s);
  });
}
static void kernel_gradient_common(Potential& p, hipStream_t s, const float* b, const float* Q, int L, float* out,
                                   int on_device) {
  const size_t n = (size_t)p.lat->N() * L;
  Staged<float> x(b, n, on_device != 0, s, true);
  Staged<float> q(Q, n, on_device != 0, s, true);
  potential_pairwise_kernel_gradient(p, x.d, q.d, L, out);
}
int dcrf_pairwise_kernel_gradient(dcrf_pairwise_t pw, const float* b, const float* Q, int L, float* out,
                                  int on_device) {
  return guarded([&] {
    DCRF_REQUIRE(pw && b && Q && L > 0 && (out || potential_kernel_nparams(pw->p) == 0),
                 "dcrf_pairwise_kernel_gradient: bad arguments");
    kernel_gradient_common(pw->p, pw->s, b, Q, L, out, on_device);
  });
}

// ---- CRF ------------------------------------------------------------------------
int dcrf_crf_create(int N, int L, dcrf_crf_t* out) {
  return guarded([&] {
    DCRF_REQUIRE(out, "dcrf_crf_create: null output");
    auto h = std::make_unique<dcrf_crf_s>();
    h->crf = std::make_unique<CRF>(N, L);
    *out = h.release();
  });
}
int dcrf_crf_destroy(dcrf_crf_t crf) {
  return guarded([&] { delete crf; });
}
int dcrf_crf_set_unary(dcrf_crf_t crf, const float* unary, int on_device) {
  return guarded([&] {
    DCRF_REQUIRE(crf && unary, "dcrf_crf_set_unary: bad arguments");
    CRF& c = *crf->crf;
    Staged<float> u(unary, (size_t)c.N() * c.L(), on_device != 0, c.stream(), true);
    c.set_unary(u.d);
    DCRF_HIP(hipStreamSynchronize(c.stream()));
  });
}
int dcrf_crf_add_pairwise(dcrf_crf_t crf, const float* features, int d, float potts_w, int ntype, int on_device) {
  return guarded([&] {
    DCRF_REQUIRE(crf && features && d > 0, "dcrf_crf_add_pairwise: bad arguments");
    CRF& c = *crf->crf;
    Staged<float> f(features, (size_t)c.N() * d, on_device != 0, c.stream(), true);
    c.add_pairwise(f.d, d, potts_w, ntype);
  });
}
int dcrf_crf_add_pairwise_compat(dcrf_crf_t crf, const float* features, int d, int compat_type,
                                 const float* params, int ntype, int on_device) {
  return guarded([&] {
    DCRF_REQUIRE(crf && features && params && d > 0, "dcrf_crf_add_pairwise_compat: bad arguments");
    CRF& c = *crf->crf;
    DCRF_REQUIRE(compat_type >= DCRF_COMPAT_POTTS && compat_type <= DCRF_COMPAT_MATRIX,
                 "dcrf_crf_add_pairwise_compat: unknown compatibility type");
    Staged<float> f(features, (size_t)c.N() * d, on_device != 0, c.stream(), true);
    c.add_pairwise(f.d, d, compat_type == DCRF_COMPAT_POTTS ? params[0] : 1.f, ntype);
    potential_set_compat(c.pairwise(c.num_pairwise() - 1), compat_type, params, c.L(), c.stream());
  });
}
int dcrf_crf_pairwise_nparams(dcrf_crf_t crf, int k, int* n) {
  return guarded([&] {
    DCRF_REQUIRE(crf && n && k >= 0 && k < crf->crf->num_pairwise(), "dcrf_crf_pairwise_nparams: bad arguments");
    *n = potential_nparams(crf->crf->pairwise(k), crf->crf->L());
  });
}
int dcrf_crf_pairwise_parameters(dcrf_crf_t crf, int k, float* params) {
  return guarded([&] {
    DCRF_REQUIRE(crf && params && k >= 0 && k < crf->crf->num_pairwise(),
                 "dcrf_crf_pairwise_parameters: bad arguments");
    potential_parameters(crf->crf->pairwise(k), crf->crf->L(), params);
  });
}
int dcrf_crf_pairwise_set_parameters(dcrf_crf_t crf, int k, const float* params) {
  return guarded([&] {
    DCRF_REQUIRE(crf && params && k >= 0 && k < crf->crf->num_pairwise(),
                 "dcrf_crf_pairwise_set_parameters: bad arguments");
    potential_set_parameters(crf->crf->pairwise(k), crf->crf->L(), params, crf->crf->stream());
    crf->crf->touch_model();
  });
}
int dcrf_crf_pairwise_gradient(dcrf_crf_t crf, int k, const float* b, const float* Q, float* grad, int on_device) {
  return guarded([&] {
    DCRF_REQUIRE(crf && b && Q && grad && k >= 0 && k < crf->crf->num_pairwise(),
                 "dcrf_crf_pairwise_gradient: bad arguments");
    CRF& c = *crf->crf;
    const size_t n = (size_t)c.N() * c.L();
    Staged<float> sb(b, n, on_device != 0, c.stream(), true);
    Staged<float> sq(Q, n, on_device != 0, c.stream(), true);
    c.pairwise_gradient(k, sb.d, sq.d, grad);
  });
}
int dcrf_crf_meanfield(dcrf_crf_t crf, int iterations, float* Q, int on_device) {
  return guarded([&] {
    DCRF_REQUIRE(crf && Q, "dcrf_crf_meanfield: bad arguments");
    CRF& c = *crf->crf;
    const size_t n = (size_t)c.N() * c.L();
    Staged<float> q(Q, n, on_device != 0, c.stream(), false);
    c.meanfield(iterations, q.d);
    q.copy_out();
    DCRF_HIP(hipStreamSynchronize(c.stream()));
  });
}
int dcrf_crf_meanfield_backward(dcrf_crf_t crf, const float* dQ, float* dU, double* dparams, int on_device) {
  return guarded([&] {
    DCRF_REQUIRE(crf && dQ && dU, "dcrf_crf_meanfield_backward: bad arguments");
    CRF& c = *crf->crf;
    const size_t n = (size_t)c.N() * c.L();
    Staged<float> g(dQ, n, on_device != 0, c.stream(), true);
    Staged<float> u(dU, n, on_device != 0, c.stream(), false);
    c.meanfield_backward(g.d, u.d, dparams);
    u.copy_out();
    DCRF_HIP(hipStreamSynchronize(c.stream()));
  });
}
int dcrf_crf_meanfield_id(dcrf_crf_t crf, long long* id) {
  return guarded([&] {
    DCRF_REQUIRE(crf && id, "dcrf_crf_meanfield_id: bad arguments");
    *id = crf->crf->meanfield_id();
  });
}
int dcrf_crf_cross_entropy(dcrf_crf_t crf, const float* Q, const int* labels, float* dQ, double* loss,
                           int on_device) {
  return guarded([&] {
    DCRF_REQUIRE(crf && Q && labels && dQ && loss, "dcrf_crf_cross_entropy: bad arguments");
    CRF& c = *crf->crf;
    const size_t n = (size_t)c.N() * c.L();
    Staged<float> q(Q, n, on_device != 0, c.stream(), true);
    Staged<int> l(labels, c.N(), on_device != 0, c.stream(), true);
    Staged<float> g(dQ, n, on_device != 0, c.stream(), false);
    *loss = c.cross_entropy(q.d, l.d, g.d);
    g.copy_out();
    DCRF_HIP(hipStreamSynchronize(c.stream()));
  });
}
int dcrf_crf_add_pairwise_gaussian(dcrf_crf_t crf, int W, int H, float sx, float sy, float potts_w) {
  return guarded([&] {
    DCRF_REQUIRE(crf, "dcrf_crf_add_pairwise_gaussian: null handle");
    CRF& c = *crf->crf;
    DCRF_REQUIRE(W > 0 && H > 0 && (long)W * H == c.N(), "dcrf_crf_add_pairwise_gaussian: W*H must equal N");
    DBuf<float> f((size_t)c.N() * 2);
    k_features_2d<<<grid_for(c.N()), kBlock, 0, c.stream()>>>(f.get(), W, H, sx, sy, nullptr, 1.f, 1.f, 1.f);
    check_launch("k_features_2d");
    c.add_pairwise(f.get(), 2, potts_w, kNormSymmetric);
  });
}
int dcrf_crf_add_pairwise_bilateral(dcrf_crf_t crf, int W, int H, float sx, float sy, float sr, float sg, float sb,
                                    const unsigned char* image, float potts_w) {
  return guarded([&] {
    DCRF_REQUIRE(crf && image, "dcrf_crf_add_pairwise_bilateral: bad arguments");
    CRF& c = *crf->crf;
    DCRF_REQUIRE(W > 0 && H > 0 && (long)W * H == c.N(), "dcrf_crf_add_pairwise_bilateral: W*H must equal N");
    DBuf<unsigned char> im((size_t)c.N() * 3);
    DCRF_HIP(hipMemcpyAsync(im.get(), image, (size_t)c.N() * 3, hipMemcpyHostToDevice, c.stream()));
    DBuf<float> f((size_t)c.N() * 5);
    k_features_2d<<<grid_for(c.N()), kBlock, 0, c.stream()>>>(f.get(), W, H, sx, sy, im.get(), sr, sg, sb);
    check_launch("k_features_2d");
    c.add_pairwise(f.get(), 5, potts_w, kNormSymmetric);
  });
}
int dcrf_crf_add_superpixel_layer(dcrf_crf_t crf, const int* region, int R, const float* e, const float* cst) {
  return guarded([&] {
    DCRF_REQUIRE(crf && region && e && cst && R > 0, "dcrf_crf_add_superpixel_layer: bad arguments");
    crf->crf->add_layer(region, R, e, cst);
  });
}
int dcrf_crf_apply(dcrf_crf_t crf, int k, const float* in, float* out, int transpose, int no_norm, int on_device) {
  return guarded([&] {
    DCRF_REQUIRE(crf && in && out, "dcrf_crf_apply: bad arguments");
    CRF& c = *crf->crf;
    DCRF_REQUIRE(k >= 0 && k < c.num_pairwise(), "dcrf_crf_apply: no such potential");
    const size_t n = (size_t)c.N() * c.L();
    Staged<float> a(in, n, on_device != 0, c.stream(), true);
    Staged<float> b(out, n, on_device != 0, c.stream(), false);
    c.apply(k, a.d, b.d, transpose != 0, no_norm != 0);
    b.copy_out();
    DCRF_HIP(hipStreamSynchronize(c.stream()));
  });
}
int dcrf_crf_apply_upper_minus_lower_ord(dcrf_crf_t crf, int k, const float* Q, float* out, int on_device) {
  return guarded([&] {
    DCRF_REQUIRE(crf && Q && out, "dcrf_crf_apply_upper_minus_lower_ord: bad arguments");
    CRF& c = *crf->crf;
    DCRF_REQUIRE(k >= 0 && k < c.num_pairwise(), "dcrf_crf_apply_upper_minus_lower_ord: no such potential");
    const size_t n = (size_t)c.N() * c.L();
    Staged<float> a(Q, n, on_device != 0, c.stream(), true);
    Staged<float> b(out, n, on_device != 0, c.stream(), false);
    c.apply_ord(k, a.d, b.d, true);
    b.copy_out();
    DCRF_HIP(hipStreamSynchronize(c.stream()));
  });
}
int dcrf_crf_apply_upper_minus_lower_ord_restricted(dcrf_crf_t crf, int k, float* rout, const float* rQ,
                                                    const int* pI, int nI, const float* Q, int store, int on_device) {
  return guarded([&] {
    DCRF_REQUIRE(crf && nI >= 0 && (nI == 0 || (rout && rQ && pI)) && (!store || Q),
                 "dcrf_crf_apply_upper_minus_lower_ord_restricted: bad arguments");
    CRF& c = *crf->crf;
    DCRF_REQUIRE(k >= 0 && k < c.num_pairwise(), "dcrf_crf_apply_upper_minus_lower_ord_restricted: no such potential");
    apply_ord_restricted(c.pairwise(k), c.stream(), rout, rQ, pI, nI, Q, c.L(), store, on_device);
  });
}
int dcrf_crf_apply_bf(dcrf_crf_t crf, int k, const float* Q, float* out, int on_device) {
  return guarded([&] {
    DCRF_REQUIRE(crf && Q && out, "dcrf_crf_apply_bf: bad arguments");
    CRF& c = *crf->crf;
    DCRF_REQUIRE(k >= 0 && k < c.num_pairwise(), "dcrf_crf_apply_bf: no such potential");
    apply_bf_common(c.pairwise(k), c.stream(), 0, nullptr, Q, out, c.L(), on_device);
  });
}
int dcrf_crf_apply_upper_minus_lower_bf(dcrf_crf_t crf, int k, const int* ind, float* out, int on_device) {
  return guarded([&] {
    DCRF_REQUIRE(crf && ind && out, "dcrf_crf_apply_upper_minus_lower_bf: bad arguments");
    CRF& c = *crf->crf;
    DCRF_REQUIRE(k >= 0 && k < c.num_pairwise(), "dcrf_crf_apply_upper_minus_lower_bf: no such potential");
    apply_bf_common(c.pairwise(k), c.stream(), 1, ind, nullptr, out, c.L(), on_device);
  });
}
int dcrf_crf_apply_upper_minus_lower_dc(dcrf_crf_t crf, int k, const int* ind, float* out, int on_device) {
  return guarded([&] {
    DCRF_REQUIRE(crf && ind && out, "dcrf_crf_apply_upper_minus_lower_dc: bad arguments");
    CRF& c = *crf->crf;
    DCRF_REQUIRE(k >= 0 && k < c.num_pairwise(), "dcrf_crf_apply_upper_minus_lower_dc: no such potential");
    apply_bf_common(c.pairwise(k), c.stream(), 3, ind, nullptr, out, c.L(), on_device);
  });
}
int dcrf_crf_apply_upper_minus_lower_bf_ord(dcrf_crf_t crf, int k, const int* ind, const float* Q, float* out,
                                            int on_device) {
  return guarded([&] {
    DCRF_REQUIRE(crf && ind && Q && out, "dcrf_crf_apply_upper_minus_lower_bf_ord: bad arguments");
    CRF& c = *crf->crf;
    DCRF_REQUIRE(k >= 0 && k < c.num_pairwise(), "dcrf_crf_apply_upper_minus_lower_bf_ord: no such potential");
    apply_bf_common(c.pairwise(k), c.stream(), 2, ind, Q, out, c.L(), on_device);
  });
}
int dcrf_crf_pairwise_kernel_gradient(dcrf_crf_t crf, int k, const float* b, const float* Q, float* out,
                                      int on_device) {
  return guarded([&] {
    DCRF_REQUIRE(crf && b && Q && out, "dcrf_crf_pairwise_kernel_gradient: bad arguments");
    CRF& c = *crf->crf;
    DCRF_REQUIRE(k >= 0 && k < c.num_pairwise(), "dcrf_crf_pairwise_kernel_gradient: no such potential");
    kernel_gradient_common(c.pairwise(k), c.stream(), b, Q, c.L(), out, on_device);
  });
}
int dcrf_crf_add_pairwise_potential(dcrf_crf_t crf, dcrf_pairwise_t pw) {
  return guarded([&] {
    DCRF_REQUIRE(crf && pw && pw->p.lat, "dcrf_crf_add_pairwise_potential: bad arguments");
    CRF& c = *crf->crf;
    Potential& src = pw->p;
    DCRF_REQUIRE(src.lat->N() == c.N(), "dcrf_crf_add_pairwise_potential: the potential has another pixel count");
    DCRF_REQUIRE(src.ctype == kPotts || src.cL == c.L(),
                 "dcrf_crf_add_pairwise_potential: the compatibility has another label count");
    DCRF_REQUIRE(src.feat.get() != nullptr, "dcrf_crf_add_pairwise_potential: the potential keeps no features");
    DCRF_HIP(hipStreamSynchronize(pw->s));
    // the same potential over the CRF's stream: features, normalisation,
    // compatibility, kernel type and parameters; the no-norm copy takes
    // parameters()(0) now, as DenseCRF::addPairwiseEnergy does (src/densecrf.cpp:161-168)
    c.add_pairwise(src.feat.get(), src.d, src.w, src.ntype);
    Potential& dst = c.pairwise(c.num_pairwise() - 1);
    if (src.ctype != kPotts) potential_set_compat(dst, src.ctype, src.cparam.data(), src.cL, c.stream());
    dst.ktype = src.ktype;
    if (!src.kparams.empty())
      potential_set_kernel_parameters(dst, src.kparams.data(), (int)src.kparams.size(), c.stream());
    DCRF_HIP(hipStreamSynchronize(c.stream()));
  });
}
int dcrf_crf_pairwise_set_kernel_type(dcrf_crf_t crf, int k, int ktype) {
  return guarded([&] {
    DCRF_REQUIRE(crf && ktype >= 0 && ktype <= 2, "dcrf_crf_pairwise_set_kernel_type: bad arguments");
    CRF& c = *crf->crf;
    DCRF_REQUIRE(k >= 0 && k < c.num_pairwise(), "dcrf_crf_pairwise_set_kernel_type: no such potential");
    DCRF_REQUIRE(c.pairwise(k).kparams.empty(), "dcrf_crf_pairwise_set_kernel_type: set before any kernel parameters");
    c.pairwise(k).ktype = ktype;
  });
}
int dcrf_crf_pairwise_kernel_nparams(dcrf_crf_t crf, int k, int* n) {
  return guarded([&] {
    DCRF_REQUIRE(crf && n, "dcrf_crf_pairwise_kernel_nparams: bad arguments");
    CRF& c = *crf->crf;
    DCRF_REQUIRE(k >= 0 && k < c.num_pairwise(), "dcrf_crf_pairwise_kernel_nparams: no such potential");
    *n = potential_kernel_nparams(c.pairwise(k));
  });
}
int dcrf_crf_unary_init(dcrf_crf_t crf, float* Q, int on_device) {
  return guarded([&] {
    DCRF_REQUIRE(crf && Q, "dcrf_crf_unary_init: bad arguments");
    CRF& c = *crf->crf;
    Staged<float> b(Q, (size_t)c.N() * c.L(), on_device != 0, c.stream(), false);
    c.unary_init(b.d);
    b.copy_out();
    DCRF_HIP(hipStreamSynchronize(c.stream()));
  });
}
int dcrf_crf_qp(dcrf_crf_t crf, const float* init, float* out, int max_iter, int ref_stop, int on_device,
                dcrf_qp_stats* stats) {
  return guarded([&] {
    DCRF_REQUIRE(crf && init && out, "dcrf_crf_qp: bad arguments");
    CRF& c = *crf->crf;
    const size_t n = (size_t)c.N() * c.L();
    Staged<float> a(init, n, on_device != 0, c.stream(), true);
    Staged<float> b(out, n, on_device != 0, c.stream(), false);
    const int cap = stats ? stats->capacity : 0;
    QPResult r = c.qp(a.d, b.d, max_iter, ref_stop != 0, cap);
    b.copy_out();
    DCRF_HIP(hipStreamSynchronize(c.stream()));
    if (stats) {
      stats->iterations = r.iterations;
      stats->updates = r.updates;
      for (size_t i = 0; i < r.energy.size() && stats->energy; ++i) stats->energy[i] = r.energy[i];
      for (size_t i = 0; i < r.num.size(); ++i) {
        if (stats->num) stats->num[i] = r.num[i];
        if (stats->denom) stats->denom[i] = r.denom[i];
      }
      for (size_t i = 0; i < r.step.size() && stats->step; ++i) stats->step[i] = r.step[i];
    }
  });
}
void dcrf_lp_params_default(dcrf_lp_params* p) {
  if (!p) return;
  p->prox_reg_const = 1e-2f;
  p->dual_gap_tol = 1e3f;
  p->prox_energy_tol = 1e3f;
  p->prox_max_iter = 10;
  p->fw_max_iter = 10;
  p->qp_max_iter = 1000;
  p->qp_tol = 1e3f;
  p->qp_const = 1e-16f;
  p->best_int = 1;
  p->accel_prox = 1;
  p->work_set_size = 10;
  p->approx_fw_iter = 10;
  p->less_confident_percent = 0.f;
  p->confidence_tol = 0.95f;
}
int dcrf_crf_lp(dcrf_crf_t crf, const float* init, float* out, const dcrf_lp_params* params, int on_device,
                dcrf_lp_stats* stats) {
  return guarded([&] {
    DCRF_REQUIRE(crf && init && out, "dcrf_crf_lp: bad arguments");
    CRF& c = *crf->crf;
    dcrf_lp_params d;
    dcrf_lp_params_default(&d);
    const dcrf_lp_params& P = params ? *params : d;
    LPParams lp;
    lp.prox_reg_const = P.prox_reg_const;
    lp.dual_gap_tol = P.dual_gap_tol;
    lp.prox_energy_tol = P.prox_energy_tol;
    lp.prox_max_iter = P.prox_max_iter;
    lp.fw_max_iter = P.fw_max_iter;
    lp.qp_max_iter = P.qp_max_iter;
    lp.qp_tol = P.qp_tol;
    lp.qp_const = P.qp_const;
    lp.best_int = P.best_int != 0;
    lp.less_confident_percent = P.less_confident_percent;
    lp.confidence_tol = P.confidence_tol;
    const size_t n = (size_t)c.N() * c.L();
    Staged<float> a(init, n, on_device != 0, c.stream(), true);
    Staged<float> b(out, n, on_device != 0, c.stream(), false);
    LPResult r = c.lp(a.d, b.d, lp, stats ? stats->capacity : 0);
    b.copy_out();
    DCRF_HIP(hipStreamSynchronize(c.stream()));
    if (stats) {
      stats->outer = r.outer;
      stats->inner = (int)r.qp_sweeps.size();
      for (size_t i = 0; i < r.qp_sweeps.size() && (int)i < stats->capacity; ++i) {
        if (stats->qp_sweeps) stats->qp_sweeps[i] = r.qp_sweeps[i];
        if (stats->dual_gap) stats->dual_gap[i] = r.gap[i];
        if (stats->delta) stats->delta[i] = r.delta[i];
      }
      for (size_t i = 0; i < r.int_energy.size() && (int)i < stats->capacity; ++i)
        if (stats->int_energy) stats->int_energy[i] = r.int_energy[i];
      stats->init_lp_energy = r.init_lp_energy;
      stats->init_int_energy = r.init_int_energy;
    }
  });
}
int dcrf_crf_current_map(dcrf_crf_t crf, const float* Q, int* labels, int on_device) {
  return guarded([&] {
    DCRF_REQUIRE(crf && Q && labels, "dcrf_crf_current_map: bad arguments");
    CRF& c = *crf->crf;
    Staged<float> a(Q, (size_t)c.N() * c.L(), on_device != 0, c.stream(), true);
    Staged<int> b(labels, (size_t)c.N(), on_device != 0, c.stream(), false);
    c.current_map(a.d, b.d);
    b.copy_out();
    DCRF_HIP(hipStreamSynchronize(c.stream()));
  });
}
int dcrf_crf_assignment_energy_true(dcrf_crf_t crf, const int* labels, int on_device, double* energy) {
  return guarded([&] {
    DCRF_REQUIRE(crf && labels && energy, "dcrf_crf_assignment_energy_true: bad arguments");
    CRF& c = *crf->crf;
    Staged<int> a(labels, (size_t)c.N(), on_device != 0, c.stream(), true);
    *energy = c.assignment_energy_true(a.d);
  });
}
int dcrf_crf_unary_energy(dcrf_crf_t crf, const int* labels, float* out, int on_device) {
  return guarded([&] {
    DCRF_REQUIRE(crf && labels && out, "dcrf_crf_unary_energy: bad arguments");
    CRF& c = *crf->crf;
    Staged<int> a(labels, (size_t)c.N(), on_device != 0, c.stream(), true);
    Staged<float> o(out, (size_t)c.N(), on_device != 0, c.stream(), false);
    c.unary_energy(a.d, o.d);
    o.copy_out();
    DCRF_HIP(hipStreamSynchronize(c.stream()));
  });
}
int dcrf_crf_pairwise_energy_true(dcrf_crf_t crf, const int* labels, int term, float* out, int on_device) {
  return guarded([&] {
    DCRF_REQUIRE(crf && labels && out, "dcrf_crf_pairwise_energy_true: bad arguments");
    CRF& c = *crf->crf;
    Staged<int> a(labels, (size_t)c.N(), on_device != 0, c.stream(), true);
    Staged<float> o(out, (size_t)c.N(), on_device != 0, c.stream(), false);
    c.pairwise_energy_true(a.d, term, o.d);
    o.copy_out();
    DCRF_HIP(hipStreamSynchronize(c.stream()));
  });
}
int dcrf_crf_assignment_energy_higher_order(dcrf_crf_t crf, const int* labels, int on_device, double* energy) {
  return guarded([&] {
    DCRF_REQUIRE(crf && labels && energy, "dcrf_crf_assignment_energy_higher_order: bad arguments");
    CRF& c = *crf->crf;
    Staged<int> a(labels, (size_t)c.N(), on_device != 0, c.stream(), true);
    *energy = c.assignment_energy_ho(a.d);
  });
}
int dcrf_crf_energy_higher_order(dcrf_crf_t crf, const float* Q, int on_device, double* energy) {
  return guarded([&] {
    DCRF_REQUIRE(crf && Q && energy, "dcrf_crf_energy_higher_order: bad arguments");
    CRF& c = *crf->crf;
    Staged<float> a(Q, (size_t)c.N() * c.L(), on_device != 0, c.stream(), true);
    *energy = c.energy_higher_order(a.d);
  });
}
int dcrf_crf_lp_energy(dcrf_crf_t crf, const float* Q, int on_device, double* energy) {
  return guarded([&] {
    DCRF_REQUIRE(crf && Q && energy, "dcrf_crf_lp_energy: bad arguments");
    CRF& c = *crf->crf;
    Staged<float> a(Q, (size_t)c.N() * c.L(), on_device != 0, c.stream(), true);
    *energy = c.lp_energy(a.d);
  });
}
int dcrf_crf_ucondgrad(dcrf_crf_t crf, const float* Q, float* Us, int on_device) {
  return guarded([&] {
    DCRF_REQUIRE(crf && Q && Us, "dcrf_crf_ucondgrad: bad arguments");
    CRF& c = *crf->crf;
    const size_t n = (size_t)c.N() * c.L();
    Staged<float> a(Q, n, on_device != 0, c.stream(), true);
    Staged<float> b(Us, n, on_device != 0, c.stream(), false);
    c.ucondgrad(a.d, b.d);
    b.copy_out();
    DCRF_HIP(hipStreamSynchronize(c.stream()));
  });
}
int dcrf_crf_qp_value(dcrf_crf_t crf, const float* Q, int on_device, double* energy) {
  return guarded([&] {
    DCRF_REQUIRE(crf && Q && energy, "dcrf_crf_qp_value: bad arguments");
    CRF& c = *crf->crf;
    Staged<float> a(Q, (size_t)c.N() * c.L(), on_device != 0, c.stream(), true);
    *energy = c.qp_value(a.d);
  });
}

// ---- device memory and timing ------------------------------------------------------
int dcrf_malloc(void** ptr, size_t bytes) {
  return guarded([&] {
    DCRF_REQUIRE(ptr, "dcrf_malloc: null output");
    *ptr = nullptr;
    if (bytes) DCRF_HIP(hipMalloc(ptr, bytes));
  });
}
int dcrf_free(void* ptr) {
  return guarded([&] {
    if (ptr) DCRF_HIP(hipFree(ptr));
  });
}
int dcrf_memcpy(void* dst, const void* src, size_t bytes, int kind) {
  return guarded([&] {
    DCRF_REQUIRE(kind >= 0 && kind <= 2, "dcrf_memcpy: kind must be 0, 1 or 2");
    const hipMemcpyKind k[3] = {hipMemcpyHostToDevice, hipMemcpyDeviceToHost, hipMemcpyDeviceToDevice};
    if (bytes) DCRF_HIP(hipMemcpy(dst, src, bytes, k[kind]));
  });
}
int dcrf_device_synchronize(void) {
  return guarded([&] { DCRF_HIP(hipDeviceSynchronize()); });
}
int dcrf_profile_enable(int on) {
  return guarded([&] {
    Profiler::get().clear();
    Profiler::get().on.store(on != 0);
  });
}
int dcrf_profile_filter(const char* prefix) {
  return guarded([&] { Profiler::get().prefix = prefix ? prefix : ""; });
}
// ---- batches ------------------------------------------------------------------
int dcrf_batch_create(int L, int B, const int* npix, dcrf_batch_t* out) {
  return guarded([&] {
    DCRF_REQUIRE(npix && out && B > 0, "dcrf_batch_create: bad arguments");
    auto h = std::make_unique<dcrf_batch_s>();
    h->b = std::make_unique<Batch>(L, std::vector<int>(npix, npix + B));
    *out = h.release();
  });
}
int dcrf_batch_create_2d(int L, int B, const int* W, const int* H, dcrf_batch_t* out) {
  return guarded([&] {
    DCRF_REQUIRE(W && H && out && B > 0, "dcrf_batch_create_2d: bad arguments");
    std::vector<int> n(B);
    for (int b = 0; b < B; ++b) {
      DCRF_REQUIRE(W[b] > 0 && H[b] > 0 && (long)W[b] * H[b] < (1L << 31), "dcrf_batch_create_2d: bad image size");
      n[b] = W[b] * H[b];
    }
    auto h = std::make_unique<dcrf_batch_s>();
    h->b = std::make_unique<Batch>(L, n);
    h->W.assign(W, W + B);
    h->H.assign(H, H + B);
    *out = h.release();
  });
}
int dcrf_batch_destroy(dcrf_batch_t b) {
  return guarded([&] { delete b; });
}
int dcrf_batch_pixel_offsets(dcrf_batch_t b, long long* offsets) {
  return guarded([&] {
    DCRF_REQUIRE(b && offsets, "dcrf_batch_pixel_offsets: bad arguments");
    const auto& o = b->b->pixel_offsets();
    for (size_t i = 0; i < o.size(); ++i) offsets[i] = o[i];
  });
}
int dcrf_batch_set_graphs(dcrf_batch_t b, int on) {
  return guarded([&] {
    DCRF_REQUIRE(b, "dcrf_batch_set_graphs: null handle");
    b->b->set_graphs(on != 0);
  });
}
int dcrf_batch_set_value_type(dcrf_batch_t b, int type) {
  return guarded([&] {
    DCRF_REQUIRE(b && (type == DCRF_VALUES_F32 || type == DCRF_VALUES_BF16), "dcrf_batch_set_value_type: bad arguments");
    b->b->set_bf16_rows(type == DCRF_VALUES_BF16);
  });
}
int dcrf_crf_set_value_type(dcrf_crf_t crf, int type) {
  return guarded([&] {
    DCRF_REQUIRE(crf && (type == DCRF_VALUES_F32 || type == DCRF_VALUES_BF16), "dcrf_crf_set_value_type: bad arguments");
    CRF& c = *crf->crf;
    for (int k = 0; k < c.num_pairwise(); ++k) c.pairwise(k).lat->set_bf16_rows(type == DCRF_VALUES_BF16);
    c.touch_model();
  });
}
int dcrf_batch_graph_captures(dcrf_batch_t b, long long* n) {
  return guarded([&] {
    DCRF_REQUIRE(b && n, "dcrf_batch_graph_captures: bad arguments");
    *n = b->b->graph_captures();
  });
}
int dcrf_batch_set_unary(dcrf_batch_t b, int img, const float* unary, int on_device) {
  return guarded([&] {
    DCRF_REQUIRE(b && unary && img >= 0 && img < b->b->B(), "dcrf_batch_set_unary: bad arguments");
    Batch& B = *b->b;
    const auto& o = B.pixel_offsets();
    Staged<float> u(unary, (size_t)(o[img + 1] - o[img]) * B.L(), on_device != 0, B.stream(), true);
    B.set_unary(img, u.d);
    DCRF_HIP(hipStreamSynchronize(B.stream()));
  });
}
int dcrf_batch_add_pairwise(dcrf_batch_t b, const float* const* features, int d, float potts_w, int ntype,
                            int on_device) {
  return guarded([&] {
    DCRF_REQUIRE(b && features && d > 0, "dcrf_batch_add_pairwise: bad arguments");
    Batch& B = *b->b;
    const auto& o = B.pixel_offsets();
    std::vector<std::unique_ptr<Staged<float>>> st;
    std::vector<const float*> f;
    for (int i = 0; i < B.B(); ++i) {
      DCRF_REQUIRE(features[i], "dcrf_batch_add_pairwise: null feature array");
      st.push_back(std::make_unique<Staged<float>>(features[i], (size_t)(o[i + 1] - o[i]) * d, on_device != 0,
                                                   B.stream(), true));
      f.push_back(st.back()->d);
    }
    B.add_pairwise(f, d, potts_w, ntype);
  });
}
static void batch_add_2d(dcrf_batch_t b, float sx, float sy, const unsigned char* const* images, float sr, float sg,
                         float sbl, float w) {
  DCRF_REQUIRE(b && !b->W.empty(), "batch: 2-D kernels need a batch made by dcrf_batch_create_2d");
  Batch& B = *b->b;
  const int d = images ? 5 : 2;
  std::vector<DBuf<float>> feats;
  std::vector<const float*> f;
  for (int i = 0; i < B.B(); ++i) {
    const int W = b->W[i], H = b->H[i];
    DBuf<unsigned char> im;
    if (images) {
      DCRF_REQUIRE(images[i], "batch: null image");
      im.resize((size_t)W * H * 3);
      DCRF_HIP(hipMemcpyAsync(im.get(), images[i], (size_t)W * H * 3, hipMemcpyHostToDevice, B.stream()));
    }
    feats.emplace_back((size_t)W * H * d);
    k_features_2d<<<grid_for((size_t)W * H), kBlock, 0, B.stream()>>>(feats.back().get(), W, H, sx, sy,
                                                                       images ? im.get() : nullptr, sr, sg, sbl);
    check_launch("k_features_2d");
    DCRF_HIP(hipStreamSynchronize(B.stream()));
    f.push_back(feats.back().get());
  }
  B.add_pairwise(f, d, w, kNormSymmetric);
}
int dcrf_batch_add_pairwise_gaussian(dcrf_batch_t b, float sx, float sy, float potts_w) {
  return guarded([&] { batch_add_2d(b, sx, sy, nullptr, 1.f, 1.f, 1.f, potts_w); });
}
int dcrf_batch_add_pairwise_bilateral(dcrf_batch_t b, float sx, float sy, float sr, float sg, float sb,
                                      const unsigned char* const* images, float potts_w) {
  return guarded([&] {
    DCRF_REQUIRE(images, "dcrf_batch_add_pairwise_bilateral: null images");
    batch_add_2d(b, sx, sy, images, sr, sg, sb, potts_w);
  });
}
int dcrf_batch_unary_init(dcrf_batch_t b, float* Q, int on_device) {
  return guarded([&] {
    DCRF_REQUIRE(b && Q, "dcrf_batch_unary_init: bad arguments");
    Batch& B = *b->b;
    Staged<float> q(Q, (size_t)B.N() * B.L(), on_device != 0, B.stream(), false);
    B.unary_init(q.d);
    q.copy_out();
    DCRF_HIP(hipStreamSynchronize(B.stream()));
  });
}
int dcrf_batch_qp(dcrf_batch_t b, const float* init, float* out, int max_iter, int ref_stop, int on_device,
                  int* iterations) {
  return guarded([&] {
    DCRF_REQUIRE(b && out, "dcrf_batch_qp: bad arguments");
    Batch& B = *b->b;
    const size_t n = (size_t)B.N() * B.L();
    std::unique_ptr<Staged<float>> a;
    if (init) a = std::make_unique<Staged<float>>(init, n, on_device != 0, B.stream(), true);
    Staged<float> o(out, n, on_device != 0, B.stream(), false);
    std::vector<int> it = B.qp(a ? a->d : nullptr, o.d, max_iter, ref_stop != 0);
    o.copy_out();
    DCRF_HIP(hipStreamSynchronize(B.stream()));
    if (iterations) std::copy(it.begin(), it.end(), iterations);
  });
}
int dcrf_batch_lp(dcrf_batch_t b, const float* init, float* out, const dcrf_lp_params* params, int on_device,
                  int* sweeps, int cap) {
  return guarded([&] {
    DCRF_REQUIRE(b && init && out, "dcrf_batch_lp: bad arguments");
    Batch& B = *b->b;
    dcrf_lp_params def;
    dcrf_lp_params_default(&def);
    const dcrf_lp_params& p = params ? *params : def;
    LPParams P;
    P.prox_reg_const = p.prox_reg_const;
    P.dual_gap_tol = p.dual_gap_tol;
    P.prox_energy_tol = p.prox_energy_tol;
    P.prox_max_iter = p.prox_max_iter;
    P.fw_max_iter = p.fw_max_iter;
    P.qp_max_iter = p.qp_max_iter;
    P.qp_tol = p.qp_tol;
    P.qp_const = p.qp_const;
    P.best_int = p.best_int != 0;
    P.less_confident_percent = p.less_confident_percent;
    P.confidence_tol = p.confidence_tol;
    const size_t n = (size_t)B.N() * B.L();
    Staged<float> a(init, n, on_device != 0, B.stream(), true);
    Staged<float> o(out, n, on_device != 0, B.stream(), false);
    std::vector<std::vector<int>> sw = B.lp(a.d, o.d, P);
    o.copy_out();
    DCRF_HIP(hipStreamSynchronize(B.stream()));
    if (sweeps && cap > 0)
      for (int i = 0; i < B.B(); ++i)
        for (int j = 0; j < cap; ++j) sweeps[(size_t)i * cap + j] = j < (int)sw[i].size() ? sw[i][j] : 0;
  });
}
int dcrf_batch_assignment_energy_true(dcrf_batch_t b, const int* labels, int on_device, double* energies) {
  return guarded([&] {
    DCRF_REQUIRE(b && labels && energies, "dcrf_batch_assignment_energy_true: bad arguments");
    Batch& B = *b->b;
    Staged<int> l(labels, (size_t)B.N(), on_device != 0, B.stream(), true);
    B.assignment_energy_true(l.d, energies);
  });
}

int dcrf_autotune_enable(int on) {
  return guarded([&] { Tuner::get().set_enabled(on != 0); });
}
int dcrf_autotune_report(char* buf, int len) {
  std::string r;
  const int rc = guarded([&] { r = Tuner::get().report(); });
  if (rc != kOk) return -rc;
  if (buf && len > 0) {
    const size_t n = std::min((size_t)len - 1, r.size());
    std::memcpy(buf, r.data(), n);
    buf[n] = 0;
  }
  return (int)r.size() + 1;
}
int dcrf_profile_report(char* buf, int len) {
  std::string r;
  const int rc = guarded([&] { r = Profiler::get().report(); });
  if (rc != kOk) return -rc;
  if (buf && len > 0) {
    const size_t n = std::min((size_t)len - 1, r.size());
    std::memcpy(buf, r.data(), n);
    buf[n] = 0;
  }
  return (int)r.size() + 1;
}

// ---- file formats (host only) ------------------------------------------------------
int dcrf_io_load_unary(const char* path, float* out, size_t cap, int* W, int* H, int* L) {
  return guarded([&] {
    DCRF_REQUIRE(path && W && H && L, "dcrf_io_load_unary: bad arguments");
    dcrf_ho::img_size sz{-1, -1};
    dcrf_ho::MatrixXf u = dcrf_ho::load_unary(path, sz);
    *W = sz.width;
    *H = sz.height;
    *L = (int)u.rows();
    if (out) {
      DCRF_REQUIRE(cap >= (size_t)u.size(), "dcrf_io_load_unary: output too small");
      std::memcpy(out, u.data(), (size_t)u.size() * sizeof(float));
    }
  });
}
int dcrf_io_load_image(const char* path, unsigned char* out, size_t cap, int* W, int* H) {
  return guarded([&] {
    DCRF_REQUIRE(path && W && H, "dcrf_io_load_image: bad arguments");
    dcrf_ho::img_size sz{-1, -1};
    std::unique_ptr<unsigned char[]> im(dcrf_ho::load_image(path, sz));
    *W = sz.width;
    *H = sz.height;
    if (out) {
      DCRF_REQUIRE(cap >= (size_t)sz.width * sz.height * 3, "dcrf_io_load_image: output too small");
      std::memcpy(out, im.get(), (size_t)sz.width * sz.height * 3);
    }
  });
}
int dcrf_io_save_map(const float* Q, int W, int H, int L, const char* path) {
  return guarded([&] {
    DCRF_REQUIRE(Q && path && W > 0 && H > 0 && L > 0, "dcrf_io_save_map: bad arguments");
    dcrf_ho::MatrixXf q(L, (long)W * H);
    std::memcpy(q.data(), Q, (size_t)L * W * H * sizeof(float));
    dcrf_ho::save_map(q, dcrf_ho::img_size{W, H}, path, "MSRC");
  });
}

int dcrf_superpixel_terms(const unsigned char* img, int N, const int* region, int R, float constant, float normaliser,
                          float* e, float* c) {
  return guarded([&] {
    DCRF_REQUIRE(img && region && e && c && N > 0 && R > 0, "dcrf_superpixel_terms: bad arguments");
    dcrf_ho::superpixel_terms(img, N, region, R, constant, normaliser, e, c);
  });
}

int dcrf_less_confident_pixels(const float* Q, int N, int L, float tol, int* indices, int* count, int on_device) {
  return guarded([&] {
    DCRF_REQUIRE(count && N >= 0 && L > 0 && (N == 0 || (Q && indices)), "dcrf_less_confident_pixels: bad arguments");
    *count = 0;
    if (N == 0) return;  // an empty field has no pixel (src/densecrf.cpp:457-462 returns an empty list)
    hipStream_t s = nullptr;
    Staged<float> q(Q, (size_t)N * L, on_device != 0, s, true);
    Staged<int> o(indices, (size_t)N, on_device != 0, s, false);
    const int n = less_confident_pixels_dev(q.d, N, L, tol, o.d, s);
    o.n = (size_t)n;  // only the first n entries are written
    o.copy_out();
    DCRF_HIP(hipStreamSynchronize(s));
    *count = n;
  });
}

// ---- unary energies (include/unary.h:31-57) -----------------------------------------
int dcrf_logistic_unary(const float* Lw, const float* f, int L, int F, int N, float* U, int on_device) {
  return guarded([&] {
    DCRF_REQUIRE(Lw && f && U && L > 0 && F > 0 && N > 0, "dcrf_logistic_unary: bad arguments");
    hipStream_t s = nullptr;
    Staged<float> w(Lw, (size_t)L * F, on_device != 0, s, true);
    Staged<float> x(f, (size_t)F * N, on_device != 0, s, true);
    Staged<float> u(U, (size_t)L * N, on_device != 0, s, false);
    logistic_unary_dev(w.d, x.d, L, F, N, u.d, s);
    u.copy_out();
    DCRF_HIP(hipStreamSynchronize(s));
  });
}
int dcrf_logistic_unary_gradient(const float* b, const float* f, int L, int F, int N, float* g, int on_device) {
  return guarded([&] {
    DCRF_REQUIRE(b && f && g && L > 0 && F > 0 && N > 0, "dcrf_logistic_unary_gradient: bad arguments");
    hipStream_t s = nullptr;
    Staged<float> x(b, (size_t)L * N, on_device != 0, s, true);
    Staged<float> y(f, (size_t)F * N, on_device != 0, s, true);
    Staged<float> o(g, (size_t)L * F, on_device != 0, s, false);
    DBuf<float> part(logistic_grad_parts(L, F, N));
    logistic_unary_gradient_dev(x.d, y.d, L, F, N, o.d, part.get(), s);
    o.copy_out();
    DCRF_HIP(hipStreamSynchronize(s));
  });
}
int dcrf_crf_set_unary_logistic(dcrf_crf_t crf, const float* Lw, const float* f, int F, int on_device) {
  return guarded([&] {
    DCRF_REQUIRE(crf && Lw && f && F > 0, "dcrf_crf_set_unary_logistic: bad arguments");
    CRF& c = *crf->crf;
    hipStream_t s = c.stream();
    Staged<float> w(Lw, (size_t)c.L() * F, on_device != 0, s, true);
    Staged<float> x(f, (size_t)F * c.N(), on_device != 0, s, true);
    DBuf<float> u((size_t)c.L() * c.N());
    logistic_unary_dev(w.d, x.d, c.L(), F, c.N(), u.get(), s);
    c.set_unary(u.get());
    DCRF_HIP(hipStreamSynchronize(s));
  });
}

// a device (or managed) allocation, not host memory: the no-allocation entry
// point cannot stage host buffers
static void require_device_pointer(const void* p, const char* what) {
  hipPointerAttribute_t a;
  const hipError_t e = hipPointerGetAttributes(&a, p);
  if (e != hipSuccess) (void)hipGetLastError();  // leave no error for the next launch check
  if (e != hipSuccess || (a.type != hipMemoryTypeDevice && a.type != hipMemoryTypeManaged))
    throw Error(kInvalidArgument, std::string("dcrf_less_confident_pixels_noalloc: ") + what +
                                      " must be a device pointer");
}

int dcrf_less_confident_pixels_noalloc(const float* Q, int N, int L, float tol, int* indices, int* count) {
  return guarded([&] {
    DCRF_REQUIRE(count && N >= 0 && L > 0 && (N == 0 || (Q && indices)),
                 "dcrf_less_confident_pixels_noalloc: bad arguments");
    *count = 0;
    DCRF_REQUIRE((long long)N <= kLcMaxPixels,
                 "dcrf_less_confident_pixels_noalloc: N exceeds the 16,777,216 pixels its fixed scratch holds");
    if (N == 0) return;
    require_device_pointer(Q, "Q");
    require_device_pointer(indices, "indices");
    *count = less_confident_pixels_noalloc_dev(Q, N, L, tol, indices, nullptr);
  });
}

}  // extern "C"
