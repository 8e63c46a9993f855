// Coded ordered filter: the lower pass of seqCompute_upper_minus_lower_ord
// (src/permutohedral.cpp:590-701) with every (lattice point, label) of the
// ordered state held as a code, and its nine bins stored only when the code
// says so.
//
// LO[b] (b = 1..9) of label k at point p sums the weights of the entries whose
// bin_lo >= b.  When every entry of p with a nonzero weight has the same bin c
// for label k, LO[b] is the point's total weight Tb for b <= c and 0 above:
// the sum of all its weights in entry order, which is the float the total-
// weight pass computes (the same additions in the same order; entries of
// weight zero add nothing).  A blur axis keeps that form when the point and
// both neighbours carry the same c (o + 0.5 * (a + b) on Tb values is the
// total weight's own blur).  So per (point, label) the state is
//   code 0..9   uniform: LO[b] = b <= code ? Tb_s : 0 (Tb_s the total weight
//               after s blur axes)
//   kCodeZero   every weight zero (Tb_s = 0 as well): all bins 0; matches
//               any uniform code (absent neighbours, the zero row)
//   kCodeRow    the nine values are stored in the row
// and the results equal the dense rows' bit for bit.  In the LP's fields
// most labels of most points are uniform (a pixel's labels sit in one or two
// bins), so the splat, the d+1 blur axes and the slice move the rows of the
// mixed labels only.
//
// Rows are label-major (bin b of label k at k * 9 + b - 1, RS values per row),
// in the ordered row order (Morton order per part, row 0 the zero row);
// codes are one byte per (row, label) in rows of LC = L rounded up to 4.
// The upper pass, the bin-edge correction rows and the full upper rows of
// parts in full mode are those of the dense filter.  With six vertices per
// pixel (d = 5) the last blur axis writes dense bin-major rows and the dense
// slice reads them (most labels of a pixel read bin 1, one run per row); with
// three (d = 2) the slice reads the codes and the stored rows.

namespace dcrf {
namespace {

constexpr uint8_t kCodeZero = 14, kCodeRow = 15;

inline int ocode_stride(int L) { return (L + 3) & ~3; }

// the code of o + 0.5 * (a + b) from the codes of its three rows
__device__ inline int code_join(int c, int x) {
  if (c == kCodeRow || x == kCodeRow) return kCodeRow;
  if (x == kCodeZero) return c;
  if (c == kCodeZero) return x;
  return c == x ? c : kCodeRow;
}
__device__ inline int code3(int co, int ca, int cb) { return code_join(code_join(co, ca), cb); }

// slot r (bin r + 1) of a row with code c (kCodeZero reads Tb = 0)
__device__ inline float coded_slot(int c, int r, float tb) { return r + 1 <= c ? tb : 0.f; }

__device__ inline void ld9(const float* p, float (&v)[9]) {
  const Float4u a = *reinterpret_cast<const Float4u*>(p), b = *reinterpret_cast<const Float4u*>(p + 4);
  v[0] = a.x, v[1] = a.y, v[2] = a.z, v[3] = a.w, v[4] = b.x, v[5] = b.y, v[6] = b.z, v[7] = b.w;
  v[8] = p[8];
}
__device__ inline void st9(float* p, const float (&v)[9], bool nt) {
  if (nt) {
#pragma unroll
    for (int r = 0; r < 9; ++r) __builtin_nontemporal_store(v[r], p + r);
  } else {
    *reinterpret_cast<Float4u*>(p) = Float4u{v[0], v[1], v[2], v[3]};
    *reinterpret_cast<Float4u*>(p + 4) = Float4u{v[4], v[5], v[6], v[7]};
    p[8] = v[8];
  }
}

// Lower pass, one lane per (point chunk, quad of labels), as k_osplat: the
// chunk's weights by bin in the lane's LDS slots (entry order), the bins'
// suffix sums.  Per label a bit set of the bins that got a nonzero weight
// gives the code.  A one-chunk point writes its codes and its mixed labels'
// rows; a chunk of a point with several writes its codes, its mixed labels'
// partial rows and (quad 0) its total weight, for k_ccombine.
template <int TAG>  // the lattice dimension (2, 5, or 0 for others): kernel names in profiles
__global__ void __launch_bounds__(kBlock) k_csplat(const uint8_t* __restrict__ bins, int L, int LP, int RS, int LC,
                                                   long nlanes, const int* __restrict__ sp_pix,
                                                   const float* __restrict__ sp_w, const int* __restrict__ seg,
                                                   const int* __restrict__ chstart, const int* __restrict__ chpoint,
                                                   const int* __restrict__ pstart, const int* __restrict__ orow,
                                                   float* __restrict__ values, uint8_t* __restrict__ codes,
                                                   float* __restrict__ part, uint8_t* __restrict__ pcode,
                                                   float* __restrict__ ptb) {
  __shared__ float hist[4 * kOHist * kBlock];
  float* hl = hist + threadIdx.x;  // this lane's slots, stride kBlock
#pragma unroll
  for (int i = 0; i < 4 * kOHist; ++i) hl[i * kBlock] = 0.f;
  const int nq = (L + 3) >> 2;
  for (long g = (long)xcd_block() * kBlock + threadIdx.x; g < nlanes; g += (long)gridDim.x * kBlock) {
    const int c = (int)(g / nq), q = (int)(g - (long)c * nq);
    const int p = chpoint[c];
    const int c0 = chstart[p], nc = chstart[p + 1] - c0;
    const int b = seg[p] + (c - c0) * kChunk;
    const int e = min(seg[p + 1], b + kChunk);
    static_assert(kOSBatch == 8, "entries load as two 16-byte words");
    int px[kOSBatch];
    float w[kOSBatch];
    auto entries = [&](int t) {
      const Int4u p0 = *reinterpret_cast<const Int4u*>(sp_pix + t), p1 = *reinterpret_cast<const Int4u*>(sp_pix + t + 4);
      const Float4u w0 = *reinterpret_cast<const Float4u*>(sp_w + t), w1 = *reinterpret_cast<const Float4u*>(sp_w + t + 4);
      px[0] = p0.x, px[1] = p0.y, px[2] = p0.z, px[3] = p0.w, px[4] = p1.x, px[5] = p1.y, px[6] = p1.z, px[7] = p1.w;
      w[0] = w0.x, w[1] = w0.y, w[2] = w0.z, w[3] = w0.w, w[4] = w1.x, w[5] = w1.y, w[6] = w1.z, w[7] = w1.w;
    };
    uint32_t used[4] = {0u, 0u, 0u, 0u};  // bins with a nonzero weight, per label
    float tw = 0.f;                       // the chunk's total weight, in entry order
    entries(b);
    for (int t0 = b; t0 < e; t0 += kOSBatch) {
      uint32_t bw[kOSBatch];
      float wc[kOSBatch];
#pragma unroll
      for (int i = 0; i < kOSBatch; ++i) {
        bw[i] = *reinterpret_cast<const uint32_t*>(bins + (size_t)px[i] * LP + 4 * q);
        wc[i] = t0 + i < e ? w[i] : 0.f;
      }
      if (t0 + kOSBatch < e) entries(t0 + kOSBatch);
#pragma unroll
      for (int i = 0; i < kOSBatch; ++i) {
        const float wi = wc[i];
        const uint32_t nz = wi != 0.f ? 1u : 0u;
        tw = tw + wi;
#pragma unroll
        for (int kk = 0; kk < 4; ++kk) {
          const int cb = min((int)((bw[i] >> (8 * kk)) & 15u), kOHist - 1);
          float* slot = &hl[(kk * kOHist + cb) * kBlock];
          *slot = *slot + wi;
          used[kk] |= nz << cb;
        }
      }
    }
    float acc[4][kOBins];
    uint32_t cw = 0;
    int code[4];
#pragma unroll
    for (int kk = 0; kk < 4; ++kk) {
      float h[kOHist];
#pragma unroll
      for (int cc = 0; cc < kOHist; ++cc) {
        h[cc] = hl[(kk * kOHist + cc) * kBlock];
        hl[(kk * kOHist + cc) * kBlock] = 0.f;  // the lane's next chunk starts from zero
      }
      float sum = h[kOHist - 1];
      acc[kk][kOBins - 1] = sum;
#pragma unroll
      for (int r = kOBins - 2; r >= 0; --r) acc[kk][r] = sum = sum + h[r + 1];
      const uint32_t u = used[kk];
      code[kk] = u == 0u ? kCodeZero : (u & (u - 1u)) == 0u ? 31 - __clz(u) : kCodeRow;
      if (4 * q + kk >= L) code[kk] = kCodeZero;  // row padding
      cw |= (uint32_t)code[kk] << (8 * kk);
    }
    const bool one = nc == 1;
    const size_t slot = one ? (size_t)orow[p] : (size_t)(pstart[p] + (c - c0));
    *reinterpret_cast<uint32_t*>((one ? codes : pcode) + slot * LC + 4 * q) = cw;
    float* dst = (one ? values : part) + slot * RS;
#pragma unroll
    for (int kk = 0; kk < 4; ++kk)
      if (code[kk] == kCodeRow) st9(dst + (size_t)(4 * q + kk) * kOBins, acc[kk], false);
    if (!one && q == 0) ptb[slot] = tw;
  }
}

// points with several chunks, one lane per (point, label): the chunks' codes
// give the point's; a mixed label sums its chunk partials in chunk order (a
// uniform chunk's partial is its total weight up to its bin), as k_combine_if
__global__ void __launch_bounds__(kBlock) k_ccombine(int L, int RS, int LC, RowMap rm, const int* __restrict__ multi,
                                                     int nmulti, const int* __restrict__ chstart,
                                                     const int* __restrict__ pstart, const float* __restrict__ part,
                                                     const uint8_t* __restrict__ pcode,
                                                     const float* __restrict__ ptb, const int* __restrict__ orow,
                                                     float* __restrict__ values, uint8_t* __restrict__ codes) {
  long m;
  int k, ks;
  rm.at(L, m, k, ks);
  if (m < 0 || m >= nmulti) return;
  const int p = multi[m];
  const int s0 = pstart[p], s1 = s0 + chstart[p + 1] - chstart[p];
  int code = kCodeZero;
  for (int s = s0; s < s1 && code != kCodeRow; ++s) code = code_join(code, pcode[(size_t)s * LC + k]);
  const size_t r = (size_t)orow[p];
  codes[r * LC + k] = (uint8_t)code;
  if (code != kCodeRow) return;
  float acc[kOBins];
  auto part_row = [&](int s, float (&v)[kOBins]) {
    const int pc = pcode[(size_t)s * LC + k];
    if (pc == kCodeRow) {
      ld9(part + (size_t)s * RS + (size_t)k * kOBins, v);
    } else {
      const float t = ptb[s];
#pragma unroll
      for (int i = 0; i < kOBins; ++i) v[i] = coded_slot(pc, i, t);
    }
  };
  part_row(s0, acc);
  for (int s = s0 + 1; s < s1; ++s) {
    float v[kOBins];
    part_row(s, v);
#pragma unroll
    for (int i = 0; i < kOBins; ++i) acc[i] = acc[i] + v[i];
  }
  st9(values + r * RS + (size_t)k * kOBins, acc, false);
}

// One blur axis, one lane per (row, quad of labels): the output codes of four
// labels from one 4-byte code word of each of the three rows; the wave's
// mixed (row, label) pairs are then listed in LDS and worked one per lane (a
// wave covers floor(64 / quads) rows, and a pair's loads cost the same
// whichever lane works it, so the mixed pairs of ten rows share a few load
// instructions instead of each row's lanes issuing their own).  Then, as
// k_blur2, the correction rows and the upper rows of parts in full mode.
constexpr int kWcItems = 256;  // listed pairs per wave: at most four per lane
template <int TAG>  // the lattice dimension (2, 5, or 0 for others): kernel names in profiles
__global__ void __launch_bounds__(kBlock) k_cblur_wc(int L, int RS, int LC, RowMap rm, uint32_t M,
                                                     const int* __restrict__ nbr, const float* __restrict__ src,
                                                     float* __restrict__ dst, const uint8_t* __restrict__ csrc,
                                                     uint8_t* __restrict__ cdst, const float* __restrict__ tb_in,
                                                     const float* __restrict__ dsrc,
                                                     float* __restrict__ ddst, int R4, FastDiv divR4, uint32_t nq,
                                                     size_t half_q, EdgeMap edge, int nt) {
  __shared__ int4 items[kBlock / 64][kWcItems];
  const int wv = (int)(threadIdx.x >> 6), lane = (int)(threadIdx.x & 63);
  const int nqc = (L + 3) >> 2;
  long i;
  int q, qs;
  rm.at(nqc, i, q, qs);
  const bool on = i >= 0 && i < (long)M;
  const int r = (int)i + 1;
  int n1 = 0, n2 = 0;
  uint32_t co = 0, ca = 0, cb = 0;
  if (on) {
    n1 = nbr[2 * i] + 1;
    n2 = nbr[2 * i + 1] + 1;
    co = *reinterpret_cast<const uint32_t*>(csrc + (size_t)r * LC + 4 * q);
    ca = *reinterpret_cast<const uint32_t*>(csrc + (size_t)n1 * LC + 4 * q);
    cb = *reinterpret_cast<const uint32_t*>(csrc + (size_t)n2 * LC + 4 * q);
  }
  int code[4];
  uint32_t out = 0, mixed = 0;
#pragma unroll
  for (int kk = 0; kk < 4; ++kk) {
    code[kk] = code3((int)((co >> (8 * kk)) & 255u), (int)((ca >> (8 * kk)) & 255u), (int)((cb >> (8 * kk)) & 255u));
    if (4 * q + kk >= L) code[kk] = kCodeZero;
    out |= (uint32_t)code[kk] << (8 * kk);
    if (on && code[kk] == kCodeRow) mixed |= 1u << kk;
  }
  if (on) *reinterpret_cast<uint32_t*>(cdst + (size_t)r * LC + 4 * q) = out;
  // the wave's list: this lane's pairs at its exclusive prefix
  const int n = __popc(mixed);
  int inc = n;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int v = __shfl_up(inc, o, 64);
    if (lane >= o) inc += v;
  }
  const int total = __shfl(inc, 63);
  const int first = inc - n;
  {
    int at = first;
#pragma unroll
    for (int kk = 0; kk < 4; ++kk)
      if (mixed & (1u << kk)) {
        const uint32_t cc = ((co >> (8 * kk)) & 255u) | (((ca >> (8 * kk)) & 255u) << 8) |
                            (((cb >> (8 * kk)) & 255u) << 16);
        items[wv][at++] = make_int4(r, n1, n2, (int)((uint32_t)(4 * q + kk) | (cc << 8)));
      }
  }
  __syncthreads();  // the lists are written
  for (int base = 0; base < total; base += 64) {
    const int t = base + lane;
    if (t < total) {
      const int4 it = items[wv][t];
      const int k = it.w & 255;
      const int c0 = (it.w >> 8) & 255, c1 = (it.w >> 16) & 255, c2 = (int)((uint32_t)it.w >> 24);
      float o[kOBins], a[kOBins], bb[kOBins];
      auto row = [&](int rr, int cc, float (&v)[kOBins]) {
        if (cc == kCodeRow) {
          ld9(src + (size_t)rr * RS + (size_t)k * kOBins, v);
        } else {
          const float tt = tb_in[rr];
#pragma unroll
          for (int x = 0; x < kOBins; ++x) v[x] = coded_slot(cc, x, tt);
        }
      };
      row(it.x, c0, o);
      row(it.y, c1, a);
      row(it.z, c2, bb);
#pragma unroll
      for (int x = 0; x < kOBins; ++x) o[x] = o[x] + 0.5f * (a[x] + bb[x]);
      st9(dst + (size_t)it.x * RS + (size_t)k * kOBins, o, nt != 0);
    }
  }
  const uint32_t stride = gridDim.x * kBlock, t0 = xcd_block() * kBlock + threadIdx.x;
  if (edge.any_corr())  // correction rows: two quads per point
    for (uint32_t t = t0; t < 2 * M; t += stride) blur_quad(nbr, dsrc, ddst, 0, t >> 1, 2u, t & 1u, false);
  if (edge.any_full())  // upper rows of the points of parts in full mode
    for (uint32_t t = t0; t < nq; t += stride) {
      const uint32_t pt = divR4.div(t);
      if (edge.point_full(pt)) blur_quad(nbr, src, dst, half_q, pt, (uint32_t)R4, t - pt * (uint32_t)R4, nt);
    }
}

// The last blur axis with its output written as dense bin-major rows (bin b
// of label k at (b - 1) * L + k, every label: a uniform one from the total
// weight after the axis, tb_out), read by the dense bin-major slice
// (k_oslice), which loads only the bins a pixel's value selects.
__global__ void __launch_bounds__(kBlock) k_cblur_expand(int L, int RS, int LC, RowMap rm, uint32_t M,
                                                         const int* __restrict__ nbr, const float* __restrict__ src,
                                                         float* __restrict__ dst, const uint8_t* __restrict__ csrc,
                                                         const float* __restrict__ tb_in,
                                                         const float* __restrict__ tb_out,
                                                         const float* __restrict__ dsrc, float* __restrict__ ddst,
                                                         int R4, FastDiv divR4, uint32_t nq, size_t half_q,
                                                         EdgeMap edge) {
  {
    long i;
    int k, ks;
    rm.at(L, i, k, ks);
    if (i >= 0 && i < (long)M) {
      const int n1 = nbr[2 * i] + 1, n2 = nbr[2 * i + 1] + 1;
      const size_t r = (size_t)i + 1;
      const int co = csrc[r * LC + k], ca = csrc[(size_t)n1 * LC + k], cb = csrc[(size_t)n2 * LC + k];
      const int code = code3(co, ca, cb);
      float o[kOBins];
      if (code == kCodeRow) {
        float a[kOBins], bb[kOBins];
        auto row = [&](size_t rr, int cc, float (&v)[kOBins]) {
          if (cc == kCodeRow) {
            ld9(src + rr * RS + (size_t)k * kOBins, v);
          } else {
            const float t = tb_in[rr];
#pragma unroll
            for (int x = 0; x < kOBins; ++x) v[x] = coded_slot(cc, x, t);
          }
        };
        row(r, co, o);
        row((size_t)n1, ca, a);
        row((size_t)n2, cb, bb);
#pragma unroll
        for (int x = 0; x < kOBins; ++x) o[x] = o[x] + 0.5f * (a[x] + bb[x]);
      } else {
        const float t = tb_out[r];
#pragma unroll
        for (int x = 0; x < kOBins; ++x) o[x] = coded_slot(code, x, t);
      }
#pragma unroll
      for (int x = 0; x < kOBins; ++x) __builtin_nontemporal_store(o[x], dst + r * RS + (size_t)x * L + k);
    }
  }
  const uint32_t stride = gridDim.x * kBlock, t0 = xcd_block() * kBlock + threadIdx.x;
  if (edge.any_corr())  // correction rows: two quads per point
    for (uint32_t t = t0; t < 2 * M; t += stride) blur_quad(nbr, dsrc, ddst, 0, t >> 1, 2u, t & 1u, false);
  if (edge.any_full())  // upper rows of the points of parts in full mode
    for (uint32_t t = t0; t < nq; t += stride) {
      const uint32_t pt = divR4.div(t);
      if (edge.point_full(pt)) blur_quad(nbr, src, dst, half_q, pt, (uint32_t)R4, t - pt * (uint32_t)R4, true);
    }
}

// Lower slice then upper slice, as k_oslice on label-major rows: a vertex
// whose code is uniform c gives LO[lo] and LO[up + 1] from its total weight
// (vt.z) without a row load.  TAG 2 and 5 fix the vertex count.
template <int TAG>
__global__ void __launch_bounds__(kBlock) k_cslice(const uint8_t* __restrict__ bins, int L, int LP, int RS, int LC,
                                                   RowMap rm, int N, int d1, const float4* __restrict__ vtx,
                                                   const float* __restrict__ lo_v, const uint8_t* __restrict__ codes,
                                                   const float* __restrict__ up_v, const float* __restrict__ dv,
                                                   EdgeMap edge, float alpha, SliceArgs o) {
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
  int cj[NV];
#pragma unroll
  for (int j = 0; j < NV; ++j)
    if (j < nv) cj[j] = codes[(size_t)__float_as_int(vt[j].x) * LC + k0];
  const int part = edge.ppart(pix);
  const int mode = edge.mode(part);
  const int lo = bb & 15, up = bb >> 4;
  const bool full = mode > kMaxCombo;
  // LO[lo] (bin 0 is Tb) and LO[up + 1] (bin 10 is 0); up is lo or lo - 1, so
  // one pair of adjacent slots LO[i0 + 1], LO[i0 + 2] holds both
  const int i0 = min(max(lo, 1), 8) - 1;
  const size_t ip = (size_t)k0 * kOBins + i0;
  const bool a_hi = lo - 1 - i0 == 1;
  const bool u_hi = up - i0 == 1;
  float a[NV], u[NV];
#pragma unroll
  for (int j = 0; j < NV; ++j)
    if (j < nv) {
      if (cj[j] == kCodeRow) {
        const float2 pr = ld_pair(lo_v, (size_t)__float_as_int(vt[j].x) * RS + ip);
        a[j] = a_hi ? pr.y : pr.x;
        u[j] = u_hi ? pr.y : pr.x;
      } else {
        a[j] = lo <= cj[j] ? vt[j].z : 0.f;
        u[j] = up + 1 <= cj[j] ? vt[j].z : 0.f;
      }
    }
  if (full && up < 9)  // the upper rows were splatted and blurred in full: UP[up] (bins 0..8)
#pragma unroll
    for (int j = 0; j < NV; ++j)
      if (j < nv) u[j] = up_v[(size_t)__float_as_int(vt[j].x) * RS + (size_t)k0 * kOBins + up];
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

// The same slice with one lane per (pixel, quad of labels) (TAG 2 with at
// least as many quads as vertices): the pixel's bins of four labels are one
// 4-byte load, the codes of four labels at a vertex one 4-byte load, the
// output four values; ten pixels per wave at L = 21.
template <int TAG>
__global__ void __launch_bounds__(kBlock) k_cslice4(const uint8_t* __restrict__ bins, int L, int LP, int RS, int LC,
                                                    RowMap rm, int N, const float4* __restrict__ vtx,
                                                    const float* __restrict__ lo_v, const uint8_t* __restrict__ codes,
                                                    const float* __restrict__ up_v, const float* __restrict__ dv,
                                                    EdgeMap edge, float alpha, SliceArgs o) {
  constexpr int NV = TAG == 2 ? 3 : 6;
  const int nq = (L + 3) >> 2;
  long pix;
  int q, qs;
  rm.at(nq, pix, q, qs);
  if (pix < 0 || pix >= N) return;
  const int k0 = 4 * q, nk = min(4, L - k0);
  float* dst = o.out + (size_t)pix * L + k0;
  const float nm = o.norm ? o.norm[pix] : 1.f;
  float prev[4] = {0.f, 0.f, 0.f, 0.f};
  if (o.accumulate) {
    if (nk == 4) {
      const Float4u u = *reinterpret_cast<const Float4u*>(dst);
      prev[0] = u.x, prev[1] = u.y, prev[2] = u.z, prev[3] = u.w;
    } else {
      for (int i = 0; i < nk; ++i) prev[i] = dst[i];
    }
  }
  const uint32_t bw = *reinterpret_cast<const uint32_t*>(bins + (size_t)pix * LP + k0);
  float4 mine = make_float4(0.f, 0.f, 0.f, 0.f);
  if (q < NV) mine = vtx[(size_t)pix * NV + q];
  const int g0 = (int)(threadIdx.x & 63) - q;
  float4 vt[NV];
#pragma unroll
  for (int j = 0; j < NV; ++j)
    vt[j] = make_float4(__shfl(mine.x, g0 + j), __shfl(mine.y, g0 + j), __shfl(mine.z, g0 + j), 0.f);
  uint32_t cw[NV];
#pragma unroll
  for (int j = 0; j < NV; ++j)
    cw[j] = *reinterpret_cast<const uint32_t*>(codes + (size_t)__float_as_int(vt[j].x) * LC + k0);
  const int part = edge.ppart(pix);
  const int mode = edge.mode(part);
  const bool full = mode > kMaxCombo;
  float x[4];
#pragma unroll
  for (int kk = 0; kk < 4; ++kk) {
    x[kk] = 0.f;
    if (kk >= nk) continue;
    const int k = k0 + kk;
    const int bb = (int)((bw >> (8 * kk)) & 255u);
    const int lo = bb & 15, up = bb >> 4;
    const int i0 = min(max(lo, 1), 8) - 1;
    const size_t ip = (size_t)k * kOBins + i0;
    const bool a_hi = lo - 1 - i0 == 1;
    const bool u_hi = up - i0 == 1;
    float a[NV], u[NV];
#pragma unroll
    for (int j = 0; j < NV; ++j) {
      const int c = (int)((cw[j] >> (8 * kk)) & 255u);
      if (c == kCodeRow) {
        const float2 pr = ld_pair(lo_v, (size_t)__float_as_int(vt[j].x) * RS + ip);
        a[j] = a_hi ? pr.y : pr.x;
        u[j] = u_hi ? pr.y : pr.x;
      } else {
        a[j] = lo <= c ? vt[j].z : 0.f;
        u[j] = up + 1 <= c ? vt[j].z : 0.f;
      }
    }
    if (full && up < 9)
#pragma unroll
      for (int j = 0; j < NV; ++j) u[j] = up_v[(size_t)__float_as_int(vt[j].x) * RS + (size_t)k * kOBins + up];
    int ci = -1;
    if (mode >= 1 && mode <= kMaxCombo && up < 9)
      for (int c = 0; c < mode; ++c)
        if (edge.combo(part, c) == k * 16 + up + 1) ci = c;
    float acc = 0.f;
#pragma unroll
    for (int j = 0; j < NV; ++j) {
      const float lv = lo == 0 ? vt[j].z : a[j];
      acc = acc + lv * (-vt[j].y * alpha);
    }
#pragma unroll
    for (int j = 0; j < NV; ++j) {
      float uv = full ? (up == 9 ? vt[j].z : u[j]) : vt[j].z - (up < 9 ? u[j] : 0.f);
      if (ci >= 0) uv = uv + dv[(size_t)__float_as_int(vt[j].x) * kMaxCombo + ci];
      acc = acc + uv * (vt[j].y * alpha);
    }
    if (o.norm) acc = acc * nm;
    if (o.apply_scale) acc = o.scale * acc;
    if (o.post_mult != 1.f) acc = o.post_mult * acc;
    x[kk] = o.accumulate ? prev[kk] + acc : acc;
  }
  if (nk == 4)
    *reinterpret_cast<Float4u*>(dst) = Float4u{x[0], x[1], x[2], x[3]};
  else
    for (int i = 0; i < nk; ++i) dst[i] = x[i];
}

// tbs[orow[p]] = v[p + 1] (one stage of the total weight, point rows to
// ordered rows; row 0 stays 0)
__global__ void __launch_bounds__(kBlock) k_tb_rows(const float* __restrict__ v, const int* __restrict__ orow, int M,
                                                    float* __restrict__ tbs) {
  const int p = blockIdx.x * kBlock + threadIdx.x;
  if (p < M) tbs[orow[p]] = v[p + 1];
  if (p == 0) tbs[0] = 0.f;
}
// stored (row, label) pairs of one code buffer (rows 1..M, labels < L), added to *cnt
__global__ void __launch_bounds__(kBlock) k_count_rows(const uint8_t* __restrict__ codes, int M, int L, int LC,
                                                       unsigned long long* __restrict__ cnt) {
  __shared__ unsigned ws[kBlock / 64];
  unsigned n = 0;
  const size_t tot = (size_t)M * LC;
  for (size_t t = (size_t)blockIdx.x * kBlock + threadIdx.x; t < tot; t += (size_t)gridDim.x * kBlock) {
    const int k = (int)(t % LC);
    n += (k < L && codes[LC + t] == kCodeRow) ? 1u : 0u;
  }
  for (int o = 32; o > 0; o >>= 1) n += __shfl_down(n, o, 64);
  if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = n;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned s = 0;
    for (int w = 0; w < kBlock / 64; ++w) s += ws[w];
    atomicAdd(cnt, (unsigned long long)s);
  }
}
// per call: row 0 of both upper halves (read as zero by the full upper pass's
// blur; the plain filter's rows share the buffers) and row 0 of both code
// buffers (the zero row) in one launch
__global__ void __launch_bounds__(kBlock) k_ocoded_reset(float* __restrict__ u0, float* __restrict__ u1, int nrow,
                                                         uint8_t* __restrict__ c0, uint8_t* __restrict__ c1, int LC) {
  const int t = (int)(blockIdx.x * kBlock + threadIdx.x);
  if (t < nrow) {
    u0[t] = 0.f;
    u1[t] = 0.f;
  }
  if (t < LC) {
    c0[t] = kCodeZero;
    c1[t] = kCodeZero;
  }
}


}  // namespace

// total weight after the splat and after each blur axis, in the ordered row
// order (the same arithmetic as ensure_total_weight); code buffers whose row
// 0 (the zero row) reads kCodeZero
void Lattice::ensure_coded_tables(int L) {
  hipStream_t s = stream_;
  const int LC = ocode_stride(L);
  if (!have_tbs_) {
    tbs_.resize((size_t)(d_ + 2) * (M_ + 1));
    ensure_values(1);
    DBuf<float> ones(N_);
    k_fill_f<<<grid_for(N_), kBlock, 0, s>>>(ones.get(), (size_t)N_, 1.f);
    SplatArgs a;
    a.kind = SplatIn::kPlain;
    a.in = ones.get();
    const RowMap rm = rowmap(1);
    float* src = vA_.get();
    float* dst = vB_.get();
    k_splat<SplatIn::kPlain><<<rows_grid(nchunks_, 1), kBlock, 0, s>>>(a, 1, rm, sp_pix_.get(), sp_w_.get(),
                                                                       seg_.get(), chstart_.get(), chpoint_.get(),
                                                                       pstart_.get(), nchunks_, src, part_.get());
    if (nmulti_ > 0)
      k_combine<<<rows_grid(nmulti_, 1), kBlock, 0, s>>>(1, rm, multi_.get(), nmulti_, chstart_.get(),
                                                         pstart_.get(), part_.get(), src);
    k_tb_rows<<<grid_for(M_), kBlock, 0, s>>>(src, orow_.get(), M_, tbs_.get());
    for (int j = 0; j <= d_; ++j) {
      k_blur<false><<<rows_grid(M_, 1), kBlock, 0, s>>>(1, rm, M_, nbr_.get() + (size_t)j * M_ * 2, src, dst);
      std::swap(src, dst);
      k_tb_rows<<<grid_for(M_), kBlock, 0, s>>>(src, orow_.get(), M_, tbs_.get() + (size_t)(j + 1) * (M_ + 1));
    }
    check_launch("coded total weights");
    DCRF_HIP(hipStreamSynchronize(s));  // `ones` is released on return
    have_tbs_ = true;
  }
  const size_t need = (size_t)(M_ + 1) * LC;
  if (need > ccap_) {
    cA_.resize(need);
    cB_.resize(need);
    ccap_ = need;
  }
  if (nmulti_ > 0 && (size_t)npart_ * LC > pccap_) {
    pcode_.resize((size_t)npart_ * LC);
    ptb_.resize(npart_);
    pccap_ = (size_t)npart_ * LC;
  }
  // row 0 of both code buffers (the zero row) is set by every call's reset kernel
}

void Lattice::ordered_coded(const float* Q, const float* d_minmax, int L, const SliceArgs& out, OrderedBins* shared,
                            bool fresh) {
  DCRF_REQUIRE(L > 0, "ordered filter: label count must be positive");
  DCRF_REQUIRE(L <= 64, "ordered filter: at most 64 labels");
  ensure_total_weight();
  ensure_ordered_tables();  // before k_obin, which lists edge values into elist_
  const int row = orow_stride(L);
  const int LC = ocode_stride(L);
  ensure_values(2 * (size_t)row, false);  // LO and UP halves of each buffer
  ensure_coded_tables(L);
  const size_t half = (size_t)(M_ + 1) * row;
  hipStream_t s = stream_;
  const RowMap rm = rowmap(L);
  float* srcL = vA_.get();
  float* dstL = vB_.get();
  float* srcU = srcL + half;
  float* dstU = dstL + half;
  uint8_t* csrc = cA_.get();
  uint8_t* cdst = cB_.get();
  // row 0 of the upper halves and of the code buffers (a kernel, so that the call stays capturable)
  k_ocoded_reset<<<grid_for((size_t)std::max(row, LC)), kBlock, 0, s>>>(srcU, dstU, row, csrc, cdst, LC);
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
  const EdgeMap emr{edge, em.pimg, oppart_.n ? oppart_.get() : nullptr, nparts_};
  if (!shared || fresh) {
    KTimer tm("obin" + tag, s);
    k_zero_ints<<<1, kBlock, 0, s>>>(edge, (int)edge_ints(nparts_));
    const uint32_t nq = (uint32_t)(L + 3) / 4;
    k_obin<<<grid_for((size_t)N_ * nq), kBlock, 0, s>>>(Q, d_minmax, N_, L, FastDiv(nq), bins, LP,
                                                                        edge, pimg_.n ? pimg_.get() : nullptr,
                                                                        elist);
  }
  {
    KTimer tm("oedge" + tag, s);
    k_ocombo<<<nparts_, 1024, 0, s>>>(edge, nparts_, d_ + 1, elist, offset_.get(), bary_.get(), orow_.get(),
                                      dfix_.get());
    const size_t nd = (size_t)(M_ + 1) * kMaxCombo;
    k_odconv<<<grid_for(nd, kBlock, 2048), kBlock, 0, s>>>(edge, nparts_, nd, dfix_.get(), dvA_.get());
    check_launch("k_ocombo");
  }
  float* dsrc = dvA_.get();
  float* ddst = dvB_.get();
  // profiling: stored (row, label) pairs after the splat and after each axis
  const bool counting = prof_on() && Profiler::get().prefix.empty();
  DBuf<unsigned long long> nrows(counting ? d_ + 2 : 0);
  if (counting) DCRF_HIP(hipMemsetAsync(nrows.get(), 0, (d_ + 2) * sizeof(unsigned long long), s));
  auto count_rows = [&](const uint8_t* c, int stage) {
    if (counting)
      k_count_rows<<<grid_for((size_t)M_ * LC, kBlock, 1024), kBlock, 0, s>>>(c, M_, L, LC, nrows.get() + stage);
  };
  const long nlanes = (long)nchunks_ * ((L + 3) / 4);
  const unsigned nvb = grid_for((size_t)nlanes);
  {
    KTimer tm("csplat" + tag, s);
    auto launch = [&](unsigned grid) {
      auto go = [&](auto kern) {
        kern<<<grid, kBlock, 0, s>>>(bins, L, LP, row, LC, nlanes, sp_pix_.get(), sp_w_.get(), seg_.get(),
                                     chstart_.get(), chpoint_.get(), pstart_.get(), orow_.get(), srcL, csrc,
                                     part_.get(), pcode_.get(), ptb_.get());
      };
      if (d_ == 5)
        go(k_csplat<5>);
      else if (d_ == 2)
        go(k_csplat<2>);
      else
        go(k_csplat<0>);
    };
    const unsigned g = Tuner::get().pick(d_ == 5 ? "csplat5" : "csplat", nvb, grid_candidates(nvb), s, launch);
    launch(g);
    if (nmulti_ > 0) {
      k_ccombine<<<rows_grid(nmulti_, L), kBlock, 0, s>>>(L, row, LC, rm, multi_.get(), nmulti_, chstart_.get(),
                                                          pstart_.get(), part_.get(), pcode_.get(), ptb_.get(),
                                                          orow_.get(), srcL, csrc);
    }
    check_launch("k_csplat");
  }
  count_rows(csrc, 0);
  // the slice: d = 5 (six vertices) reads dense bin-major rows that the last
  // blur axis writes out, where a pixel's labels mostly read one run of the
  // row (bin 1 of every label); otherwise the coded rows, d = 2 with one lane
  // per quad of labels when a pixel's vertices fit its quads
  const int nqc = (L + 3) / 4;
  const RowMap rq = rowmap(nqc);
  const bool expand = d_ == 5;
  const bool quad = d_ == 2 && nqc >= d_ + 1;
  {  // the upper pass of parts in full mode (exits at once otherwise), in the slice's layout
    KTimer tm("osplat_up" + tag, s);
    auto up = [&](auto kern) {
      // a small grid: the launch exits at once unless some image is in full mode
      kern<<<std::min(nvb, 512u), kBlock, 0, s>>>(bins, L, LP, row, nlanes, sp_pix_.get(), sp_w_.get(), seg_.get(),
                                                   chstart_.get(), chpoint_.get(), pstart_.get(), nchunks_,
                                                   orow_.get(), srcU, part_.get(), em);
    };
    if (expand)
      up(k_osplat<100, float>);  // bin-major rows
    else
      up(k_osplat<102, float>);  // label-major rows
    if (nmulti_ > 0) {
      const int pv = kOBins * L;
      const RowMap rw = rowmap(pv);
      k_combine_if<float><<<std::min(rows_grid(nmulti_, pv), 256u), kBlock, 0, s>>>(
          pv, row, rw, multi_.get(), nmulti_, chstart_.get(), pstart_.get(), part_.get(), orow_.get(), srcU, em, 1);
    }
    check_launch("k_osplat upper");
  }
  const int R4 = row / 4;
  const size_t nquads = (size_t)M_ * R4;
  DCRF_REQUIRE(nquads < (1u << 31), "ordered filter: lattice too large for one blur launch");
  const FastDiv divR4((uint32_t)R4);
  const size_t half_q = half / 4;
  for (int j = 0; j <= d_; ++j) {  // forward order only (:621, :669)
    const int* nb = onbr_.get() + (size_t)j * M_ * 2;
    {
      KTimer tm((j == d_ ? "cblur_last" : "cblur") + tag, s);
      if (expand && j == d_)
        k_cblur_expand<<<rows_grid(M_, L), kBlock, 0, s>>>(
            L, row, LC, rm, (uint32_t)M_, nb, srcL, dstL, csrc, tbs_.get() + (size_t)j * (M_ + 1),
            tbs_.get() + (size_t)(j + 1) * (M_ + 1), dsrc, ddst, R4, divR4, (uint32_t)nquads, half_q, emr);
      else
        [&](auto kern) {
          kern<<<rows_grid(M_, nqc), kBlock, 0, s>>>(L, row, LC, rq, (uint32_t)M_, nb, srcL, dstL, csrc, cdst,
                                                     tbs_.get() + (size_t)j * (M_ + 1), dsrc, ddst, R4, divR4,
                                                     (uint32_t)nquads, half_q, emr, j == d_);
        }(d_ == 5 ? k_cblur_wc<5> : d_ == 2 ? k_cblur_wc<2> : k_cblur_wc<0>);
    }
    check_launch("k_cblur");
    std::swap(srcL, dstL);
    std::swap(srcU, dstU);
    std::swap(csrc, cdst);
    std::swap(dsrc, ddst);
    if (!(expand && j == d_)) count_rows(csrc, j + 1);
  }
  if (counting) {
    std::vector<unsigned long long> h(d_ + 2);
    DCRF_HIP(hipMemcpyAsync(h.data(), nrows.get(), h.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost, s));
    DCRF_HIP(hipStreamSynchronize(s));
    if (expand) h[d_ + 1] = (unsigned long long)M_ * L;  // the last axis writes every row out
    for (int j = 0; j <= d_ + 1; ++j) Profiler::get().count("rows" + tag + ":" + std::to_string(j), (double)h[j]);
  }
  KTimer tm((expand ? "oslice" : "cslice") + tag, s);
  auto slice = [&](auto kern) {
    kern<<<rows_grid(N_, L), kBlock, 0, s>>>(bins, L, LP, row, LC, rm, N_, d_ + 1, vtx_.get(), srcL, csrc, srcU, dsrc,
                                             em, lattice_alpha(d_), out);
  };
  auto slice4 = [&](auto kern) {
    kern<<<rows_grid(N_, nqc), kBlock, 0, s>>>(bins, L, LP, row, LC, rq, N_, vtx_.get(), srcL, csrc, srcU, dsrc, em,
                                               lattice_alpha(d_), out);
  };
  if (expand)
    k_oslice<5, float><<<rows_grid(N_, L), kBlock, 0, s>>>(bins, L, LP, row, rm, N_, d_ + 1, vtx_.get(), srcL, srcU,
                                                           dsrc, em, lattice_alpha(d_), out);
  else if (quad)
    slice4(k_cslice4<2>);
  else if (d_ == 2)
    slice(k_cslice<2>);
  else
    slice(k_cslice<0>);
  check_launch("k_cslice");
}

}  // namespace dcrf
