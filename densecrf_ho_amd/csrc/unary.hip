// LogisticUnaryEnergy (src/unary.cpp:53-68, include/unary.h:49-57) on matrix
// cores: get() = L * f and gradient(b) = b * f^T, both with the fp32 MFMA
// v_mfma_f32_16x16x4_f32.  Layouts are Eigen's column-major ones:
//   Lw  L x F, element (l, k) at k*L + l
//   f   F x N, element (k, p) at p*F + k      (features contiguous per pixel)
//   U   L x N, element (l, p) at p*L + l      (labels contiguous per pixel)
//   b   L x N, as U
//   g   L x F, as Lw (the gradient is returned as g resized to L*F x 1)
// Both products read and write whole pixel ranges, which are contiguous runs of
// f, b and U; they are staged through LDS so global accesses coalesce.  The
// work is bound by HBM (about L F / (2 (L + F)) flops per byte, below the fp32
// MFMA / HBM ratio of MI355X): MFMA keeps the arithmetic off the vector ALUs.
//
// MFMA 16x16x4 f32 operand layout (wave64): lane t holds A[t % 16][t / 16] and
// B[t / 16][t % 16]; the result D (16 x 16) is 4 floats per lane, register r
// holding D[4 (t / 16) + r][t % 16].

namespace dcrf {
namespace {

typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kUTile = 64;  // pixels per workgroup tile (16 per wave)
constexpr int kUKc = 32;    // features per LDS chunk
constexpr int kUPad = kUKc + 1;

template <int MT>  // 16-label tiles: L <= 16 MT
__global__ void __launch_bounds__(256) k_logistic_get(const float* __restrict__ Lw, const float* __restrict__ f,
                                                      int L, int F, int N, float* __restrict__ U) {
  __shared__ float fs[kUTile][kUPad];
  __shared__ float ws[MT * 16][kUPad];
  __shared__ float os[kUTile * MT * 16];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const long long p0 = (long long)blockIdx.x * kUTile;
  const int np = (int)min((long long)kUTile, (long long)N - p0);
  f32x4 acc[MT];
#pragma unroll
  for (int m = 0; m < MT; ++m) acc[m] = f32x4{0.f, 0.f, 0.f, 0.f};
  for (int k0 = 0; k0 < F; k0 += kUKc) {
    const int kc = min(kUKc, F - k0);
    for (int i = tid; i < kUTile * kUKc; i += 256) {
      const int p = i / kUKc, k = i % kUKc;
      fs[p][k] = (p < np && k < kc) ? f[(p0 + p) * F + k0 + k] : 0.f;
    }
    for (int i = tid; i < MT * 16 * kUKc; i += 256) {
      const int k = i / (MT * 16), l = i % (MT * 16);
      ws[l][k] = (l < L && k < kc) ? Lw[(size_t)(k0 + k) * L + l] : 0.f;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < kUKc; kk += 4) {
      const float b = fs[w * 16 + (lane & 15)][kk + (lane >> 4)];
#pragma unroll
      for (int m = 0; m < MT; ++m) {
        const float a = ws[m * 16 + (lane & 15)][kk + (lane >> 4)];
        acc[m] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, acc[m], 0, 0, 0);
      }
    }
    __syncthreads();
  }
  // D[label][pixel] -> os[pixel * L + label], then one contiguous store of the tile
  const int px = w * 16 + (lane & 15);
#pragma unroll
  for (int m = 0; m < MT; ++m)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int l = m * 16 + 4 * (lane >> 4) + r;
      if (l < L) os[px * L + l] = acc[m][r];
    }
  __syncthreads();
  for (int i = tid; i < np * L; i += 256) U[p0 * L + i] = os[i];
}

constexpr int kGPix = 4096;  // pixels per workgroup of the gradient
constexpr int kGChunk = 64;  // pixels per LDS chunk
constexpr int kGFb = 64;     // features per workgroup (4 waves x 16)

// partial b f^T over pixels [blockIdx.x kGPix, +kGPix) and features
// [blockIdx.y kGFb, +kGFb): part[blockIdx.x][k][l] (L x F per pixel block)
template <int MT>
__global__ void __launch_bounds__(256) k_logistic_grad_part(const float* __restrict__ bq, const float* __restrict__ f,
                                                            int L, int F, int N, float* __restrict__ part) {
  __shared__ float bs[kGChunk][MT * 16 + 1];
  __shared__ float fs[kGChunk][kGFb + 1];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const long long pbeg = (long long)blockIdx.x * kGPix;
  const long long pend = min(pbeg + kGPix, (long long)N);
  const int fb = blockIdx.y * kGFb;
  const int nf = min(kGFb, F - fb);
  f32x4 acc[MT];
#pragma unroll
  for (int m = 0; m < MT; ++m) acc[m] = f32x4{0.f, 0.f, 0.f, 0.f};
  for (long long c0 = pbeg; c0 < pend; c0 += kGChunk) {
    const int nc = (int)min((long long)kGChunk, pend - c0);
    for (int i = tid; i < kGChunk * L; i += 256) {  // nc pixels x L labels, contiguous in b
      const int p = i / L, l = i % L;
      bs[p][l] = p < nc ? bq[(c0 + p) * L + l] : 0.f;
    }
    for (int i = tid; i < kGChunk * MT * 16; i += 256) {  // zero the padding labels
      const int p = i / (MT * 16), l = i % (MT * 16);
      if (l >= L) bs[p][l] = 0.f;
    }
    for (int i = tid; i < kGChunk * kGFb; i += 256) {
      const int p = i / kGFb, k = i % kGFb;
      fs[p][k] = (p < nc && k < nf) ? f[(c0 + p) * F + fb + k] : 0.f;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < kGChunk; kk += 4) {  // 4 pixels per MFMA step
      const float bb = fs[kk + (lane >> 4)][w * 16 + (lane & 15)];
#pragma unroll
      for (int m = 0; m < MT; ++m) {
        const float a = bs[kk + (lane >> 4)][m * 16 + (lane & 15)];
        acc[m] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, bb, acc[m], 0, 0, 0);
      }
    }
    __syncthreads();
  }
  const int k = fb + w * 16 + (lane & 15);
  if (k >= F) return;
  float* out = part + (size_t)blockIdx.x * L * F;
#pragma unroll
  for (int m = 0; m < MT; ++m)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int l = m * 16 + 4 * (lane >> 4) + r;
      if (l < L) out[(size_t)k * L + l] = acc[m][r];
    }
}

// g[i] = sum over pixel blocks of part[b][i], in block order (FP64)
__global__ void __launch_bounds__(kBlock) k_logistic_grad_sum(const float* __restrict__ part, int nblk, size_t n,
                                                              float* __restrict__ g) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double s = 0;
  for (int b = 0; b < nblk; ++b) s += (double)part[(size_t)b * n + i];
  g[i] = (float)s;
}

}  // namespace

// U (L x N) = Lw (L x F) f (F x N); device pointers, stream s
void logistic_unary_dev(const float* Lw, const float* f, int L, int F, int N, float* U, hipStream_t s) {
  DCRF_REQUIRE(L >= 1 && L <= kMaxLabels && F >= 1 && N >= 1, "LogisticUnaryEnergy: 1..64 labels, F >= 1, N >= 1");
  const unsigned grid = (unsigned)((N + kUTile - 1) / kUTile);
  KTimer tm("logistic_get", s);
  if (L <= 16)
    k_logistic_get<1><<<grid, 256, 0, s>>>(Lw, f, L, F, N, U);
  else if (L <= 32)
    k_logistic_get<2><<<grid, 256, 0, s>>>(Lw, f, L, F, N, U);
  else if (L <= 48)
    k_logistic_get<3><<<grid, 256, 0, s>>>(Lw, f, L, F, N, U);
  else
    k_logistic_get<4><<<grid, 256, 0, s>>>(Lw, f, L, F, N, U);
  check_launch("k_logistic_get");
}

// g (L x F) = b (L x N) f^T; device pointers; part holds logistic_grad_parts(L, F, N) floats
size_t logistic_grad_parts(int L, int F, int N) {
  return (size_t)((N + kGPix - 1) / kGPix) * L * F;
}
void logistic_unary_gradient_dev(const float* b, const float* f, int L, int F, int N, float* g, float* part,
                                 hipStream_t s) {
  DCRF_REQUIRE(L >= 1 && L <= kMaxLabels && F >= 1 && N >= 1, "LogisticUnaryEnergy: 1..64 labels, F >= 1, N >= 1");
  const int nblk = (N + kGPix - 1) / kGPix;
  const dim3 grid((unsigned)nblk, (unsigned)((F + kGFb - 1) / kGFb));
  KTimer tm("logistic_grad", s);
  if (L <= 16)
    k_logistic_grad_part<1><<<grid, 256, 0, s>>>(b, f, L, F, N, part);
  else if (L <= 32)
    k_logistic_grad_part<2><<<grid, 256, 0, s>>>(b, f, L, F, N, part);
  else if (L <= 48)
    k_logistic_grad_part<3><<<grid, 256, 0, s>>>(b, f, L, F, N, part);
  else
    k_logistic_grad_part<4><<<grid, 256, 0, s>>>(b, f, L, F, N, part);
  const size_t n = (size_t)L * F;
  k_logistic_grad_sum<<<grid_for(n), kBlock, 0, s>>>(part, nblk, n, g);
  check_launch("k_logistic_grad");
}

}  // namespace dcrf
