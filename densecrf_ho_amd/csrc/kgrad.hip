// Kernel (feature) gradients of the pairwise terms:
//   * Permutohedral::gradient (src/permutohedral.cpp:1095-1178):
//       df = d/df a^T K b, N x d, from two splat + blur passes (a forward,
//       b with the axes reversed) and a per-pixel "slice" through the
//       simplex vertices ordered by rank, rotated and multiplied by the
//       elevation matrix;
//   * DenseKernel::featureGradient / gradient (src/pairwise.cpp:90-122,
//     :170-181) for the four normalisation types, and the parameter
//     gradient of a DIAG_KERNEL (d values) or FULL_KERNEL (d x d);
//   * PairwisePotential::kernelGradient (:402-407) = kernel gradient of b
//     against the compatibility applied to Q.
// The blur of these passes sums in double and rounds once, as the
// reference's gradient does; the slice follows its float order per pixel.
#include "crf.hpp"

namespace dcrf {
namespace {

constexpr int kKgMaxD = 16;

struct KgConst {
  float scale[kKgMaxD];  // scale_factor of :1104-1106
};

// One thread per pixel: for each value k in order, the rotated differences
// r_a[j] = alpha v[o0] - alpha v[o1] of the vertices ranked d - rank and the
// next one, then grad_j += x_k * scale_{j-1} (sm - j r_a[j]), sm += r_a[j]
// (:1137-1165).  df = grad (first pass) or df + grad (second pass).
__global__ void __launch_bounds__(kBlock) k_kgrad_slice(int N, int d, int V, const int* __restrict__ offset,
                                                        const int* __restrict__ rank, const float* __restrict__ vals,
                                                        const float* __restrict__ x, float alpha, KgConst kc,
                                                        int accumulate, float* __restrict__ df) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const int d1 = d + 1;
  int o0[kKgMaxD + 1], o1[kKgMaxD + 1];
  for (int j = 0; j <= d; ++j) {
    const int r0 = d - rank[(size_t)i * d1 + j];
    const int r1 = r0 + 1 > d ? 0 : r0 + 1;
    o0[j] = offset[(size_t)i * d1 + r0] + 1;
    o1[j] = offset[(size_t)i * d1 + r1] + 1;
  }
  float grad[kKgMaxD];
  for (int j = 0; j < d; ++j) grad[j] = 0.f;
  for (int k = 0; k < V; ++k) {
    const float xk = x[(size_t)i * V + k];
    float sm = alpha * vals[(size_t)o0[0] * V + k] - alpha * vals[(size_t)o1[0] * V + k];
    for (int j = 1; j <= d; ++j) {
      const float ra = alpha * vals[(size_t)o0[j] * V + k] - alpha * vals[(size_t)o1[j] * V + k];
      const float v = kc.scale[j - 1] * (sm - (float)j * ra);
      grad[j - 1] = grad[j - 1] + xk * v;
      sm = sm + ra;
    }
  }
  for (int j = 0; j < d; ++j) {
    float* o = df + (size_t)i * d + j;
    *o = accumulate ? *o + grad[j] : grad[j];
  }
}

// elementwise pieces of featureGradient
// out = ((a * fb + fa * b) * 0.5) * norm^3 (symmetric), (a * fb) * norm^2
// (after: fa null), (fa * b) * norm^2 (before: fb null)
__global__ void __launch_bounds__(kBlock) k_kg_mix(size_t n, int V, const float* __restrict__ a,
                                                   const float* __restrict__ b, const float* __restrict__ fa,
                                                   const float* __restrict__ fb, const float* __restrict__ norm,
                                                   float* __restrict__ out) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const float nm = norm[t / V];
  if (fa && fb) {
    out[t] = ((a[t] * fb[t] + fa[t] * b[t]) * 0.5f) * ((nm * nm) * nm);
  } else if (fb) {
    out[t] = (a[t] * fb[t]) * (nm * nm);
  } else {
    out[t] = (fa[t] * b[t]) * (nm * nm);
  }
}
__global__ void __launch_bounds__(kBlock) k_kg_scale(size_t n, int V, const float* __restrict__ a,
                                                     const float* __restrict__ norm, float* __restrict__ out) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) out[t] = a[t] * norm[t / V];
}
__global__ void __launch_bounds__(kBlock) k_kg_sub(size_t n, const float* __restrict__ r, float* __restrict__ g) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) g[t] = -r[t] + g[t];
}

// parameter gradient partials: DIAG entry e = sum_i f[i,e] fg[i,e]; FULL
// entry e = (r, c) column-major: sum_i fg[i,r] f[i,c]; one entry per
// blockIdx.y, FP64 partial per block
__global__ void __launch_bounds__(kBlock) k_kg_param_parts(int N, int d, int full, const float* __restrict__ f,
                                                           const float* __restrict__ fg, double* __restrict__ parts) {
  const int e = blockIdx.y;
  const int r = full ? e % d : e, c = full ? e / d : e;
  double v[1] = {0.0};
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += (long)gridDim.x * blockDim.x)
    v[0] += (double)fg[(size_t)i * d + r] * (double)f[(size_t)i * d + c];
  block_sum<1>(v);
  if (threadIdx.x == 0) parts[(size_t)e * gridDim.x + blockIdx.x] = v[0];
}
__global__ void __launch_bounds__(kBlock) k_kg_param_final(int nb, const double* __restrict__ parts,
                                                           float* __restrict__ out) {
  double v[1] = {0.0};
  for (int b = threadIdx.x; b < nb; b += blockDim.x) v[0] += parts[(size_t)blockIdx.x * nb + b];
  block_sum<1>(v);
  if (threadIdx.x == 0) out[blockIdx.x] = (float)v[0];
}

// features transformed by the kernel parameters: diag(v) f or V f
__global__ void __launch_bounds__(kBlock) k_kfeat(int N, int d, int full, const float* __restrict__ v,
                                                  const float* __restrict__ f, float* __restrict__ out) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)N * d) return;
  const long i = t / d;
  const int r = (int)(t % d);
  if (!full) {
    out[t] = v[r] * f[t];
  } else {
    float acc = 0.f;
    for (int c = 0; c < d; ++c) acc = acc + v[(size_t)c * d + r] * f[(size_t)i * d + c];
    out[t] = acc;
  }
}

}  // namespace

void Lattice::gradient(const float* a, const float* b, int V, float* df) {
  DCRF_REQUIRE(V > 0 && d_ <= kKgMaxD, "gradient: value size must be positive and d at most 16");
  ensure_values(V);
  KgConst kc;
  const float inv_std = (float)(std::sqrt(2.0 / 3.0) * (d_ + 1));
  for (int i = 0; i < d_; ++i) kc.scale[i] = (float)(1.0 / std::sqrt(double((i + 2) * (i + 1))) * inv_std);
  const float alpha = 1.0f / (1 + powf(2.f, (float)-d_)) / (float)(d_ + 1);
  for (int dir = 0; dir < 2; ++dir) {
    SplatArgs in;
    in.kind = SplatIn::kPlain;
    in.in = dir ? b : a;
    const float* vals = splat_blur(in, V, dir == 1, true, std::string());
    k_kgrad_slice<<<grid_for(N_), kBlock, 0, stream_>>>(N_, d_, V, offset_.get(), rank_.get(), vals, dir ? a : b,
                                                         alpha, kc, dir, df);
    check_launch("k_kgrad_slice");
  }
}

// DenseKernel::featureGradient (src/pairwise.cpp:95-122): N x d into fg
void potential_feature_gradient(Potential& p, const float* a, const float* b, int V, float* fg) {
  Lattice& lat = *p.lat;
  hipStream_t s = lat.stream();
  const int N = lat.N();
  const size_t n = (size_t)N * V;
  const unsigned g = grid_for(n);
  if (p.ntype == kNoNorm) {
    lat.gradient(a, b, V, fg);
    return;
  }
  DBuf<float> fa(p.ntype != kNormAfter ? n : 0), fb(p.ntype != kNormBefore ? n : 0), x(n), ones(n), t1(n), t2(n);
  DBuf<float> r((size_t)N * lat.d());
  k_fill<<<g, kBlock, 0, s>>>(ones.get(), n, 1.f);
  SliceArgs o;
  SplatArgs in;
  const float* ga = a;
  const float* gb = b;
  if (p.ntype == kNormSymmetric) {
    in.kind = SplatIn::kScaled;
    in.norm = p.norm.get();
    in.in = a;
    o.out = fa.get();
    lat.filter(in, o, V, true);  // K^T (a diag(norm))
    in.in = b;
    o.out = fb.get();
    lat.filter(in, o, V, false);  // K (b diag(norm))
    k_kg_mix<<<g, kBlock, 0, s>>>(n, V, a, b, fa.get(), fb.get(), p.norm.get(), x.get());
    k_kg_scale<<<g, kBlock, 0, s>>>(n, V, a, p.norm.get(), t1.get());
    k_kg_scale<<<g, kBlock, 0, s>>>(n, V, b, p.norm.get(), t2.get());
    ga = t1.get();
    gb = t2.get();
  } else if (p.ntype == kNormAfter) {
    in.kind = SplatIn::kPlain;
    in.in = b;
    o.out = fb.get();
    lat.filter(in, o, V, false);
    k_kg_mix<<<g, kBlock, 0, s>>>(n, V, a, b, nullptr, fb.get(), p.norm.get(), x.get());
    k_kg_scale<<<g, kBlock, 0, s>>>(n, V, a, p.norm.get(), t1.get());
    ga = t1.get();
  } else {  // before
    in.kind = SplatIn::kPlain;
    in.in = a;
    o.out = fa.get();
    lat.filter(in, o, V, true);
    k_kg_mix<<<g, kBlock, 0, s>>>(n, V, a, b, fa.get(), nullptr, p.norm.get(), x.get());
    k_kg_scale<<<g, kBlock, 0, s>>>(n, V, b, p.norm.get(), t2.get());
    gb = t2.get();
  }
  check_launch("k_kg_mix");
  lat.gradient(x.get(), ones.get(), V, r.get());
  lat.gradient(ga, gb, V, fg);
  const size_t nd = (size_t)N * lat.d();
  k_kg_sub<<<grid_for(nd), kBlock, 0, s>>>(nd, r.get(), fg);
  check_launch("k_kg_sub");
  DCRF_HIP(hipStreamSynchronize(s));  // the temporaries are freed on return
}

int potential_kernel_nparams(const Potential& p) {
  return p.ktype == kDiagKernel ? p.d : p.ktype == kFullKernel ? p.d * p.d : 0;
}

// DenseKernel::gradient (:170-181) of a^T K b w.r.t. the kernel parameters;
// host out (potential_kernel_nparams floats)
void potential_kernel_gradient(Potential& p, const float* a, const float* b, int V, float* host_out) {
  const int np = potential_kernel_nparams(p);
  if (np == 0) return;
  DCRF_REQUIRE(p.feat.get() != nullptr, "kernel gradient: the potential keeps no features");
  Lattice& lat = *p.lat;
  hipStream_t s = lat.stream();
  const int N = lat.N(), d = lat.d();
  DBuf<float> fg((size_t)N * d), out(np);
  potential_feature_gradient(p, a, b, V, fg.get());
  const int nb = 256;
  DBuf<double> parts((size_t)np * nb);
  k_kg_param_parts<<<dim3(nb, np), kBlock, 0, s>>>(N, d, p.ktype == kFullKernel ? 1 : 0, p.feat.get(), fg.get(),
                                                   parts.get());
  k_kg_param_final<<<np, kBlock, 0, s>>>(nb, parts.get(), out.get());
  check_launch("k_kg_param");
  DCRF_HIP(hipMemcpyAsync(host_out, out.get(), np * sizeof(float), hipMemcpyDeviceToHost, s));
  DCRF_HIP(hipStreamSynchronize(s));
}

void potential_set_kernel_parameters(Potential& p, const float* v, int n, hipStream_t s) {
  const int np = potential_kernel_nparams(p);
  DCRF_REQUIRE(n == np, "setKernelParameters: parameter count does not match the kernel type");
  if (np == 0) return;
  DCRF_REQUIRE(p.feat.get() != nullptr, "setKernelParameters: the potential keeps no features");
  const int N = p.lat->N(), d = p.d;
  DBuf<float> dv(np), f2((size_t)N * d);
  DCRF_HIP(hipMemcpyAsync(dv.get(), v, np * sizeof(float), hipMemcpyHostToDevice, s));
  k_kfeat<<<grid_for((size_t)N * d), kBlock, 0, s>>>(N, d, p.ktype == kFullKernel ? 1 : 0, dv.get(), p.feat.get(),
                                                     f2.get());
  check_launch("k_kfeat");
  DBuf<float> orig = std::move(p.feat);
  const float nn_w = p.nn_w;
  potential_init(p, f2.get(), N, d, p.w, p.ntype, s);
  p.feat = std::move(orig);  // the kernel keeps the features as given (f_)
  p.nn_w = nn_w;
  p.kparams.assign(v, v + np);
  DCRF_HIP(hipStreamSynchronize(s));
}

// DenseKernel::parameters (:136-146): the values set last, else ones(d)
// (DIAG) or the identity (FULL)
std::vector<float> potential_kernel_parameters(const Potential& p) {
  const int np = potential_kernel_nparams(p);
  if ((int)p.kparams.size() == np) return p.kparams;
  std::vector<float> v(np, 0.f);
  for (int i = 0; i < p.d; ++i) v[p.ktype == kFullKernel ? i * p.d + i : i] = 1.f;
  return v;
}

// PairwisePotential::kernelGradient(b, Q) (src/pairwise.cpp:402-407): the
// kernel gradient of b against compatibility(Q); b, Q N x L on the device
void potential_pairwise_kernel_gradient(Potential& p, const float* b, const float* Q, int L, float* host_out) {
  hipStream_t s = p.lat->stream();
  const size_t n = (size_t)p.lat->N() * L;
  DBuf<float> lq(n);
  if (p.ctype == kPotts) {
    DCRF_HIP(hipMemcpyAsync(lq.get(), Q, n * sizeof(float), hipMemcpyDeviceToDevice, s));
    k_rscale<<<grid_for(n), kBlock, 0, s>>>(lq.get(), n, -p.w);
    check_launch("k_rscale");
  } else {
    compat_apply_rows(p, Q, lq.get(), p.lat->N(), L, s);
  }
  potential_kernel_gradient(p, b, lq.get(), L, host_out);
}

}  // namespace dcrf
