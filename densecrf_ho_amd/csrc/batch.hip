// Batched dense CRF: many images of uneven sizes solved together on one
// stream.  (The reference has one DenseCRF per image and no batching; each
// image here follows the per-image algorithm of crf.hip / lp.hip exactly.)
//
// Layout: the images' pixels are concatenated (image b owns pixels
// [pix[b], pix[b+1])), every field is that union, label-contiguous per pixel.
// Each pairwise term has ONE lattice, the disjoint union of the images'
// lattices, so every splat / blur / slice / ordered-filter launch covers the
// whole batch.  Per-pixel and reduction kernels run on a block map that gives
// each image its own range of blocks (ImageBlocks): a block reads its image's
// scalars, and per-block partials combine per image.  All control decisions
// of the solvers (line-search step, qp-gamma stopping sweep, LP init rounding,
// best-integer tracking) are taken per image on the device; the host only
// learns when every image has finished a qp-gamma loop.  The fixed parts of an
// iteration are captured once into HIP graphs and replayed.
#include <algorithm>
#include <cmath>
#include <map>
#include <numeric>

#include "batch.hpp"
#include "crf_kernels.hpp"

namespace dcrf {
namespace {

// outer-loop state of the batched LP, per image
constexpr int kO_BEST = 0, kO_PREV = 1, kO_NSC = 2;                       // doubles
constexpr int kOF_COUNT = 0, kOF_STOP = 1, kOF_COPY = 2, kOF_LC = 3, kOF_NFL = 4;  // ints
// the LP's per-image scalar and flag slots are those of lp.hip
constexpr int kB_LNSC = kL_NSC, kB_EINT = kL_EINT, kB_TMP = kL_TMP;

// per-image min / max of Q (rescale) as block partials
template <class SEG>
__global__ void __launch_bounds__(kBlock) kb_minmax_parts(SEG sg, int L, const float* __restrict__ Q,
                                                          float* __restrict__ parts) {
  __shared__ float smn[kBlock], smx[kBlock];
  const SegPos q = sg.pos();
  float mn = INFINITY, mx = -INFINITY;
  bool nan = false;
  const size_t e = (size_t)(q.p0 + q.n) * L;
  for (size_t i = (size_t)q.p0 * L + (size_t)q.lb * blockDim.x + threadIdx.x; i < e; i += (size_t)q.nb * blockDim.x) {
    const float v = Q[i];
    nan = nan || v != v;
    mn = fminf(mn, v);
    mx = fmaxf(mx, v);
  }
  smn[threadIdx.x] = nan ? NAN : mn;
  smx[threadIdx.x] = nan ? NAN : mx;
  __syncthreads();
  for (int o = kBlock / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) {
      const float a = smn[threadIdx.x], b = smn[threadIdx.x + o];
      const float c = smx[threadIdx.x], d = smx[threadIdx.x + o];
      smn[threadIdx.x] = (a != a || b != b) ? NAN : fminf(a, b);
      smx[threadIdx.x] = (c != c || d != d) ? NAN : fmaxf(c, d);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    parts[2 * blockIdx.x] = smn[0];
    parts[2 * blockIdx.x + 1] = smx[0];
  }
}

// out[i] = a[i] - b[i] per image (doubles at stride)
__global__ void kb_sub(double* __restrict__ a, const double* __restrict__ b, int stride, int nimg) {
  for (int i = threadIdx.x; i < nimg; i += blockDim.x) a[(size_t)i * stride] -= b[(size_t)i * stride];
}

// LP init (:543-549): images whose LP energy exceeds the integer energy start
// from max_rounding; the integer energy is the first best and previous value
__global__ void kb_lp_init_decide(const double* __restrict__ sc, double* __restrict__ ost, int* __restrict__ ofl,
                                  int nimg) {
  for (int i = threadIdx.x; i < nimg; i += blockDim.x) {
    const double eint = sc[(size_t)i * kB_LNSC + kB_EINT];
    const double elp = sc[(size_t)i * kB_LNSC + kB_TMP + 1];
    ost[(size_t)i * kO_NSC + kO_BEST] = eint;
    ost[(size_t)i * kO_NSC + kO_PREV] = eint;
    ofl[(size_t)i * kOF_NFL + kOF_COUNT] = 0;
    ofl[(size_t)i * kOF_NFL + kOF_STOP] = 0;
    ofl[(size_t)i * kOF_NFL + kOF_LC] = 0;
    ofl[(size_t)i * kOF_NFL + kOF_COPY] = elp > eint ? 1 : 0;
  }
}

// Q = onehot(lab) for the pixels of images whose copy flag is set
template <class SEG>
__global__ void __launch_bounds__(kBlock) kb_onehot_if(SEG sg, int L, const int* __restrict__ lab,
                                                       const int* __restrict__ ofl, float* __restrict__ Q) {
  const SegPos q = sg.pos();
  if (!ofl[(size_t)q.img * kOF_NFL + kOF_COPY]) return;
  const size_t e = (size_t)(q.p0 + q.n) * L;
  for (size_t i = (size_t)q.p0 * L + (size_t)q.lb * blockDim.x + threadIdx.x; i < e; i += (size_t)q.nb * blockDim.x) {
    const size_t pix = i / L;
    Q[i] = lab[pix] == (int)(i - pix * L) ? 1.f : 0.f;
  }
}

// dst = src for images whose copy flag is set
template <class SEG>
__global__ void __launch_bounds__(kBlock) kb_copy_if(SEG sg, int L, const int* __restrict__ ofl,
                                                     const float* __restrict__ src, float* __restrict__ dst) {
  const SegPos q = sg.pos();
  if (!ofl[(size_t)q.img * kOF_NFL + kOF_COPY]) return;
  const size_t e = (size_t)(q.p0 + q.n) * L;
  for (size_t i = (size_t)q.p0 * L + (size_t)q.lb * blockDim.x + threadIdx.x; i < e; i += (size_t)q.nb * blockDim.x)
    dst[i] = src[i];
}

// end of an outer LP step with best_int (:753-770): count steps whose integer
// energy changed by less than prox_energy_tol, stop an image after 5 in a
// row, and mark it for a copy Q -> best when its integer energy improved
__global__ void kb_lp_outer_decide(const double* __restrict__ sc, double* __restrict__ ost, int* __restrict__ ofl,
                                   float tol, int nimg) {
  for (int i = threadIdx.x; i < nimg; i += blockDim.x) {
    int* f = ofl + (size_t)i * kOF_NFL;
    double* o = ost + (size_t)i * kO_NSC;
    f[kOF_COPY] = 0;
    if (f[kOF_STOP]) continue;
    const double e = sc[(size_t)i * kB_LNSC + kB_EINT];
    if (fabs(e - o[kO_PREV]) < (double)tol)
      ++f[kOF_COUNT];
    else
      f[kOF_COUNT] = 0;
    o[kO_PREV] = e;
    if (f[kOF_COUNT] >= 5) {
      f[kOF_STOP] = 1;
      continue;
    }
    if (e < o[kO_BEST]) {
      f[kOF_COPY] = 1;
      o[kO_BEST] = e;
    }
  }
}

// the less-confident stop (:784-794) per image, after the best copy: stop an
// image when more than pct percent of its pixels are less confident, then
// clear its count for the next outer step
__global__ void kb_lp_less_confident_decide(const long* __restrict__ pix, int* __restrict__ ofl, float pct, int nimg) {
  for (int i = threadIdx.x; i < nimg; i += blockDim.x) {
    int* f = ofl + (size_t)i * kOF_NFL;
    if (!f[kOF_STOP] && double(f[kOF_LC]) / double(pix[i + 1] - pix[i]) * 100 > (double)pct) f[kOF_STOP] = 1;
    f[kOF_LC] = 0;
  }
}

__global__ void kb_fill_int(int* __restrict__ p, int n, int v) {
  for (int i = threadIdx.x; i < n; i += blockDim.x) p[i] = v;
}

}  // namespace

// ---------------------------------------------------------------------------
void BlockPlan::build(const std::vector<int>& per_image, hipStream_t s) {
  const int B = (int)per_image.size();
  first_h.assign(B + 1, 0);
  for (int b = 0; b < B; ++b) first_h[b + 1] = first_h[b] + std::max(1, per_image[b]);
  nblocks = first_h[B];
  std::vector<int> img_h(nblocks);
  for (int b = 0; b < B; ++b)
    for (int k = first_h[b]; k < first_h[b + 1]; ++k) img_h[k] = b;
  img.resize(nblocks);
  first.resize(B + 1);
  DCRF_HIP(hipMemcpyAsync(img.get(), img_h.data(), nblocks * sizeof(int), hipMemcpyHostToDevice, s));
  DCRF_HIP(hipMemcpyAsync(first.get(), first_h.data(), (B + 1) * sizeof(int), hipMemcpyHostToDevice, s));
  DCRF_HIP(hipStreamSynchronize(s));
}

GraphCache::~GraphCache() { clear(); }
void GraphCache::clear() {
  for (auto& kv : exec) (void)hipGraphExecDestroy(kv.second);
  exec.clear();
  seen.clear();
}
void GraphCache::run(long key, hipStream_t s, bool use, const std::vector<const void*>& buffers,
                     const std::function<void()>& body) {
  if (!use || prof_on()) {
    body();
    return;
  }
  if (buffers != captured_with) {  // some buffer moved: no graph may be replayed
    clear();
    captured_with = buffers;
  }
  auto it = exec.find(key);
  if (it == exec.end()) {
    if (!seen.count(key)) {  // first use runs eagerly: buffers are sized, tables built
      seen.insert(key);
      body();
      return;
    }
    hipGraph_t g = nullptr;
    DCRF_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeRelaxed));
    body();
    DCRF_HIP(hipStreamEndCapture(s, &g));
    hipGraphExec_t e = nullptr;
    const hipError_t rc = hipGraphInstantiate(&e, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    DCRF_HIP(rc);
    ++captures;
    it = exec.emplace(key, e).first;
  }
  DCRF_HIP(hipGraphLaunch(it->second, s));
}

Batch::Batch(int L, const std::vector<int>& npix) : L_(L), B_((int)npix.size()) {
  DCRF_REQUIRE(B_ > 0 && B_ <= 65535, "batch: 1..65535 images");
  DCRF_REQUIRE(L > 0 && L <= kMaxLabels, "batch: 1..64 labels");
  pix_h_.assign(B_ + 1, 0);
  for (int b = 0; b < B_; ++b) {
    DCRF_REQUIRE(npix[b] > 0, "batch: every image needs pixels");
    pix_h_[b + 1] = pix_h_[b] + npix[b];
  }
  N_ = pix_h_[B_];
  DCRF_REQUIRE(N_ * L < (1L << 31), "batch: more than 2^31 pixel-labels");
  DCRF_HIP(hipStreamCreateWithFlags(&s_, hipStreamNonBlocking));
  pix_.resize(B_ + 1);
  DCRF_HIP(hipMemcpyAsync(pix_.get(), pix_h_.data(), (B_ + 1) * sizeof(long), hipMemcpyHostToDevice, s_));
  // reduction / row-group plan: every image gets the block count the
  // one-image path uses for its reductions (kRedBlocks), so each image's
  // partial sums, and with them every line-search step, gap and energy, are
  // those of the one-image path bit for bit; sweep plan: one block per 256
  // pixels, as there
  std::vector<int> red(B_, kRedBlocks), sw(B_);
  for (int b = 0; b < B_; ++b) sw[b] = (int)((npix[b] + kBlock - 1) / kBlock);
  red_.build(red, s_);
  sweep_.build(sw, s_);
  U_.resize((size_t)N_ * L);
  DCRF_HIP(hipMemsetAsync(U_.get(), 0, U_.bytes(), s_));
  DCRF_HIP(hipStreamSynchronize(s_));
}

Batch::~Batch() {
  if (s_) {
    (void)hipStreamSynchronize(s_);
    graphs_.clear();
    pw_.clear();
    (void)hipStreamDestroy(s_);
  }
}

void Batch::set_unary(int b, const float* dU) {
  DCRF_REQUIRE(b >= 0 && b < B_, "batch: image index out of range");
  DCRF_HIP(hipMemcpyAsync(U_.get() + (size_t)pix_h_[b] * L_, dU, (size_t)(pix_h_[b + 1] - pix_h_[b]) * L_ * sizeof(float),
                          hipMemcpyDeviceToDevice, s_));
}

void Batch::add_pairwise(const std::vector<const float*>& feats, int d, float w, int ntype) {
  DCRF_REQUIRE((int)feats.size() == B_, "batch: one feature array per image");
  std::vector<std::unique_ptr<Lattice>> parts;
  std::vector<const Lattice*> raw;
  for (int b = 0; b < B_; ++b) {
    auto lat = std::make_unique<Lattice>(s_);
    lat->build(feats[b], (int)(pix_h_[b + 1] - pix_h_[b]), d);
    raw.push_back(lat.get());
    parts.push_back(std::move(lat));
  }
  auto u = Lattice::concat(raw, s_);
  parts.clear();
  u->set_bf16_rows(bf16_);
  auto p = std::make_unique<Potential>();
  potential_init_lattice(*p, std::move(u), w, ntype, s_);
  pw_.push_back(std::move(p));
  graphs_.clear();
}

void Batch::unary_init(float* Q) {
  k_softmax_neg<<<grid_for(N_), kBlock, 0, s_>>>(U_.get(), Q, (int)N_, L_);
  check_launch("k_softmax_neg");
}

std::vector<const void*> Batch::work_buffers() const {
  std::vector<const void*> v;
  for (auto& p : pw_) {
    p->lat->work_buffers(v);
    v.push_back(p->ctmp.get());
  }
  obins_.buffers(v);
  for (const void* p : {(const void*)Q_.get(), (const void*)gy_.get(), (const void*)psx_.get(),
                        (const void*)tmp_.get(), (const void*)rpix_.get(), (const void*)lab_.get(),
                        (const void*)parts_.get(), (const void*)sc_.get(), (const void*)fl_.get(),
                        (const void*)alpha_.get(), (const void*)sa_.get(), (const void*)gamma_.get(),
                        (const void*)gamma2_.get(), (const void*)Y_.get(), (const void*)lQ_.get(),
                        (const void*)curQ_.get(), (const void*)best_.get(), (const void*)mmparts_.get(),
                        (const void*)mm_.get(), (const void*)lsc_.get(), (const void*)lfl_.get()})
    v.push_back(p);
  return v;
}

ImageBlocks Batch::red_map() const { return red_.map(pix_.get()); }
ImageBlocks Batch::sweep_map() const { return sweep_.map(pix_.get()); }

void Batch::ensure_state() {
  const size_t NL = (size_t)N_ * L_;
  for (DBuf<float>* b : {&Q_, &gy_, &psx_, &tmp_, &rpix_}) b->resize(b == &rpix_ ? (size_t)N_ : NL);
  lab_.resize(N_);
  const int nparts = std::max(red_.nblocks, sweep_.nblocks);
  parts_.resize((size_t)std::max(kSweepBatchB, 3) * nparts + 16);
  static_assert(kSweepBatchB == kSweepBatch, "per-image sweep totals are kSweepBatch apart");
  swtot_.resize((size_t)B_ * kSweepBatchB);
  sc_.resize((size_t)B_ * kS_COUNT);
  fl_.resize((size_t)B_ * kF_COUNT);
}

// ---------------------------------------------------------------------------
// QP (qp_inference_super_pixels_non_convex without superpixel layers), one
// iteration for all images; per-image scalars at sc + b * kS_COUNT
// ---------------------------------------------------------------------------
std::vector<int> Batch::qp(const float* init, float* out, int max_iter, bool ref_stop) {
  DCRF_REQUIRE(max_iter > 0 || ref_stop, "batch qp: max_iter must be positive without the reference stopping rule");
  ensure_state();
  const size_t NL = (size_t)N_ * L_;
  hipStream_t s = s_;
  const int K = (int)pw_.size();
  DCRF_REQUIRE(K <= 7, "batch qp: at most 7 pairwise terms");
  const ImageBlocks rm = red_map();
  const unsigned nbr = (unsigned)red_.nblocks;
  const int* first = red_.first.get();
  const FastDiv divL(L_);
  float *Q = Q_.get(), *gy = gy_.get(), *psx = psx_.get(), *tmp = tmp_.get();
  int* lab = lab_.get();
  double* sc = sc_.get();
  double* parts = parts_.get();
  int* fl = fl_.get();

  if (init)
    DCRF_HIP(hipMemcpyAsync(Q, init, NL * sizeof(float), hipMemcpyDeviceToDevice, s));
  else
    unary_init(Q);
  DCRF_HIP(hipMemcpyAsync(gy, U_.get(), NL * sizeof(float), hipMemcpyDeviceToDevice, s));
  k_zero_scalars<<<1, 256, 0, s>>>(sc, B_ * kS_COUNT, fl, B_ * kF_COUNT);
  k_dot<<<nbr, kBlock, 0, s>>>(U_.get(), Q, L_, rm, parts);
  finalize_seg<1><<<B_, kBlock, 0, s>>>(parts, first, 0, sc + kS_TMP, kS_COUNT);
  for (int k = 0; k < K; ++k) {
    potential_apply(*pw_[k], Q, tmp, L_, false, false);
    k_gy_init_accum<<<nbr, kBlock, 0, s>>>(gy, tmp, Q, L_, rm, parts);
    finalize_seg<1><<<B_, kBlock, 0, s>>>(parts, first, 0, sc + kS_TMP + 1 + k, kS_COUNT);
  }
  k_qp_e0<<<1, 256, 0, s>>>(sc, K, 0, nullptr, 0, B_);
  k_argmin_rg<<<rg_grid(N_, L_, 8192), kBlock, 0, s>>>(gy, lab, (int)N_, L_);
  check_launch("batch qp init");

  auto iteration = [&]() {
    for (int k = 0; k < K; ++k) {
      Potential& p = *pw_[k];
      const bool pre = p.ntype == kNormSymmetric || p.ntype == kNormBefore;
      const bool post = p.ntype == kNormSymmetric || p.ntype == kNormAfter;
      SplatArgs a;
      a.kind = pre ? SplatIn::kSxScaled : SplatIn::kSx;
      a.in = Q;
      a.lab = lab;
      a.norm = p.norm.get();
      SliceArgs o;
      o.out = psx;
      o.norm = post ? p.norm.get() : nullptr;
      o.accumulate = k > 0;
      compat_slice(p, o, L_, s, [&](const SliceArgs& oo) { p.lat->filter(a, oo, L_, false); });
    }
    {
      KTimer tm("qp_dots", s);
      k_qp_dots<<<nbr, kBlock, 0, s>>>(U_.get(), Q, lab, psx, rm, divL, parts, fl);
    }
    finalize_seg<3><<<B_, kBlock, 0, s>>>(parts, first, 0, sc + kS_DOTS, kS_COUNT);
    k_qp_step<<<1, 256, 0, s>>>(sc, fl, 0, nullptr, nullptr, nullptr, nullptr, 0, B_);
    {
      KTimer tm("qp_update", s);
      k_qp_update_rg<<<nbr, kBlock, 0, s>>>(U_.get(), Q, gy, psx, lab, nullptr, 0, rm, L_, sc, parts, fl);
    }
    finalize_seg<1><<<B_, kBlock, 0, s>>>(parts, first, 0, sc + kS_EPART, kS_COUNT);
    k_qp_check<<<1, 256, 0, s>>>(sc, fl, ref_stop ? 1 : 0, 0, nullptr, 0, B_);
    check_launch("batch qp iteration");
  };

  const int cap_iters = max_iter > 0 ? max_iter : (1 << 30);
  const int group = 8;
  int issued = 0;
  std::vector<int> hf((size_t)B_ * kF_COUNT);
  while (issued < cap_iters) {
    const int n = std::min(group, cap_iters - issued);
    for (int i = 0; i < n; ++i) run_graph(kGraphQP, iteration);
    issued += n;
    DCRF_HIP(hipMemcpyAsync(hf.data(), fl, hf.size() * sizeof(int), hipMemcpyDeviceToHost, s));
    DCRF_HIP(hipStreamSynchronize(s));
    bool all = true;
    for (int b = 0; b < B_; ++b) all = all && hf[(size_t)b * kF_COUNT + kF_DONE] != 0;
    if (all) break;
  }
  DCRF_HIP(hipMemcpyAsync(out, Q, NL * sizeof(float), hipMemcpyDeviceToDevice, s));
  DCRF_HIP(hipStreamSynchronize(s));
  std::vector<int> iters(B_);
  for (int b = 0; b < B_; ++b) iters[b] = hf[(size_t)b * kF_COUNT + kF_ITERS];
  return iters;
}

// per-image labelling energy (unary + true pairwise) of device labels into
// dst[b * stride]
void Batch::int_energy_dev(const int* lab, double* dst, int stride) {
  hipStream_t s = s_;
  k_fill<<<grid_for(N_), kBlock, 0, s>>>(rpix_.get(), (size_t)N_, 0.f);
  for (auto& pp : pw_) {
    Potential& p = *pp;
    SplatArgs a;
    a.kind = SplatIn::kNotOnehot;
    a.lab = lab;
    SliceArgs o;
    o.out = tmp_.get();
    o.scale = -p.nn_w;
    o.apply_scale = true;
    p.lat->filter(a, o, L_, false);
    k_pairwise_true_accum<<<grid_for(N_), kBlock, 0, s>>>(tmp_.get(), lab, rpix_.get(), (int)N_, L_);
  }
  k_assign_energy<<<red_.nblocks, kBlock, 0, s>>>(U_.get(), lab, rpix_.get(), red_map(), L_, parts_.get());
  finalize_seg<1><<<B_, kBlock, 0, s>>>(parts_.get(), red_.first.get(), 0, dst, stride);
  check_launch("batch int energy");
}

// per-image LP energy (compute_energy_LP_higher_order without superpixels,
// src/densecrf.cpp:982-1019): U.Q - sum_k Q . ord_k(rescale(Q)), rescaled per
// image, into dst[b * stride] (scratch: same stride)
void Batch::lp_energy_dev(const float* Q, double* dst, double* scratch, int stride) {
  hipStream_t s = s_;
  const ImageBlocks rm = red_map();
  const int nb = red_.nblocks;
  const int* first = red_.first.get();
  k_dot<<<nb, kBlock, 0, s>>>(U_.get(), Q, L_, rm, parts_.get());
  finalize_seg<1><<<B_, kBlock, 0, s>>>(parts_.get(), first, 0, dst, stride);
  kb_minmax_parts<<<nb, kBlock, 0, s>>>(rm, L_, Q, mmparts_.get());
  k_lp_minmax_final<<<B_, 1024, 0, s>>>(mmparts_.get(), nb, first, mm_.get(), lfl_.get());
  for (size_t k = 0; k < pw_.size(); ++k) {
    Potential& p = *pw_[k];
    SliceArgs o;
    o.out = tmp_.get();
    compat_slice(p, o, L_, s,
                 [&](const SliceArgs& oo) { p.lat->ordered(Q, mm_.get(), L_, oo, &obins_, k == 0); });
    k_dot<<<nb, kBlock, 0, s>>>(Q, tmp_.get(), L_, rm, parts_.get());
    finalize_seg<1><<<B_, kBlock, 0, s>>>(parts_.get(), first, 0, scratch, stride);
    kb_sub<<<1, 256, 0, s>>>(dst, scratch, stride, B_);
  }
  check_launch("batch lp energy");
}

// ---------------------------------------------------------------------------
// LP (lp_inference_prox_super_pixels without superpixel layers, best_int),
// the same steps as CRF::lp for every image at once
// ---------------------------------------------------------------------------
std::vector<std::vector<int>> Batch::lp(const float* init, float* out, const LPParams& P) {
  DCRF_REQUIRE(P.best_int, "batch lp: only best_int (the default) is supported");
  DCRF_REQUIRE(P.qp_max_iter > 0 && P.fw_max_iter > 0 && P.prox_max_iter > 0, "batch lp: iteration counts must be positive");
  ensure_state();
  const size_t NL = (size_t)N_ * L_;
  const int N = (int)N_, L = L_, B = B_;
  hipStream_t s = s_;
  for (DBuf<float>* b : {&alpha_, &sa_, &gamma_, &gamma2_, &Y_, &lQ_, &curQ_, &best_}) b->resize(NL);
  mmparts_.resize(2 * (size_t)std::max(red_.nblocks, 1));
  mm_.resize(2 * (size_t)B);
  lsc_.resize((size_t)B * kL_NSC);
  lfl_.resize((size_t)B * kLF_NFL);
  ost_.resize((size_t)B * kO_NSC);
  ofl_.resize((size_t)B * kOF_NFL);
  const ImageBlocks rm = red_map(), sm = sweep_map();
  const int nbr = red_.nblocks, nbs = sweep_.nblocks;
  const int* rfirst = red_.first.get();
  double* sc = lsc_.get();
  int* fl = lfl_.get();
  int* ofl = ofl_.get();
  int* lab = lab_.get();
  k_zero_scalars<<<1, 256, 0, s>>>(sc, B * kL_NSC, fl, B * kLF_NFL);

  const float lambda = P.prox_reg_const;
  LPConst c;
  c.lambda = lambda;
  c.c_m1_over_M = (float)(-1.0 / (double)(float)L);
  c.c_2l_over_M = (float)(2 * lambda) / (float)L;
  c.c_1_over_M = (float)(1.0 / (double)(float)L);
  c.c_1m2_over_M = (float)(1 - 2.0 / (double)(float)L);
  c.qp_delta = P.qp_const;

  std::vector<int> hfl((size_t)B * kLF_NFL), hofl((size_t)B * kOF_NFL);
  auto read_flags = [&]() {
    DCRF_HIP(hipMemcpyAsync(hfl.data(), fl, hfl.size() * sizeof(int), hipMemcpyDeviceToHost, s));
    DCRF_HIP(hipStreamSynchronize(s));
  };
  auto check_err = [&]() {
    read_flags();
    for (int b = 0; b < B; ++b)
      if (hfl[(size_t)b * kLF_NFL + kLF_ERR])
        throw Error(kMaxvalNegative, "Maxval < 0, skipping (image " + std::to_string(b) + ")");
  };

  // ---- initialisation (:534-550) ------------------------------------------------
  if (L == 21)
    k_lp_project_px<21><<<nbr, kBlock, 0, s>>>(rm, init, lQ_.get(), lab, 0);
  else
    k_lp_project<<<nbr, kBlock, 0, s>>>(rm, L, init, lQ_.get(), lab, 0);
  DCRF_HIP(hipMemcpyAsync(best_.get(), lQ_.get(), NL * sizeof(float), hipMemcpyDeviceToDevice, s));
  int_energy_dev(lab, sc + kL_EINT, kL_NSC);
  lp_energy_dev(lQ_.get(), sc + kL_TMP + 1, sc + kL_TMP + 2, kL_NSC);
  check_err();
  kb_lp_init_decide<<<1, 256, 0, s>>>(sc, ost_.get(), ofl, B);
  kb_onehot_if<<<nbr, kBlock, 0, s>>>(rm, L, lab, ofl, lQ_.get());

  std::vector<std::vector<int>> sweeps(B);
  float* gcur = gamma_.get();
  float* gnext = gamma2_.get();
  hstatus_.resize(3 + (size_t)B);
  for (int it = 0; it < P.prox_max_iter; ++it) {
    k_fill<<<grid_for(NL), kBlock, 0, s>>>(gcur, NL, 1.f);
    DCRF_HIP(hipMemcpyAsync(curQ_.get(), lQ_.get(), NL * sizeof(float), hipMemcpyDeviceToDevice, s));
    k_fill<<<grid_for(NL), kBlock, 0, s>>>(alpha_.get(), NL, 0.f);
    for (int pit = 1; pit <= P.fw_max_iter; ++pit) {
      const int upd = pit > 1 ? 1 : 0;
      run_graph(kGraphY + upd, [&]() {
        KTimer tm("lp_Y", s);
        k_lp_Y<<<nbr, kBlock, 0, s>>>(rm, L, c, alpha_.get(), nullptr, sa_.get(), nullptr, sc, upd, U_.get(),
                                      curQ_.get(), Y_.get());
      });
      // qp-gamma sweeps in batches for all images, each image's batch lengths
      // planned on the device (k_lp_sweep_decide); images that stopped carry
      // their gamma over while the others go on.  The host reads the status
      // from pinned memory after each batch: a batch enqueued ahead would
      // copy every finished image's gamma once more when the loop has ended.
      constexpr int kSweepAhead = 1;
      k_lp_sweep_reset<<<1, 256, 0, s>>>(fl, B, std::min(kSweepFirst, P.qp_max_iter));
      while (true) {
        for (int ahead = 0; ahead < kSweepAhead; ++ahead) {
          const int parity = gcur == gamma_.get() ? 0 : 1;
          float* gi = gcur;
          float* go = gnext;
          run_graph(kGraphSweep + parity, [&]() {
            {
              KTimer tm("lp_sweeps", s);
              if (L == 21)
                k_lp_sweeps<21, 21><<<nbs, kBlock, 0, s>>>(sm, L, c, gi, go, Y_.get(), -1, 0, 1, parts_.get(), fl);
              else
                DCRF_MAXL_DISPATCH(L, (k_lp_sweeps<MAXL><<<nbs, kBlock, 0, s>>>(sm, L, c, gi, go, Y_.get(), -1, 0, 1,
                                                                               parts_.get(), fl)));
            }
            k_lp_sweep_sums<<<dim3(B, kSweepBatch), 1024, 0, s>>>(parts_.get(), nbs, sweep_.first.get(),
                                                                   swtot_.get(), fl, 1);
            k_lp_sweep_decide<<<1, 256, 0, s>>>(swtot_.get(), B, -1, P.qp_tol, P.qp_max_iter, sc, fl,
                                                hstatus_.get());
            {
              KTimer tm("lp_sweeps_replay", s);
              if (L == 21)
                k_lp_sweeps<21, 21><<<nbs, kBlock, 0, s>>>(sm, L, c, gi, go, Y_.get(), -1, 1, 1, parts_.get(), fl);
              else
                DCRF_MAXL_DISPATCH(L, (k_lp_sweeps<MAXL><<<nbs, kBlock, 0, s>>>(sm, L, c, gi, go, Y_.get(), -1, 1, 1,
                                                                               parts_.get(), fl)));
            }
          });
          std::swap(gcur, gnext);
        }
        DCRF_HIP(hipStreamSynchronize(s));
        if (hstatus_.get()[0] || P.qp_max_iter <= 0) break;
      }
      for (int b = 0; b < B; ++b) sweeps[b].push_back((int)((const volatile int*)hstatus_.get())[3 + b]);
      const int parity = gcur == gamma_.get() ? 0 : 1;
      float* gm = gcur;
      run_graph(kGraphPost + parity, [&]() {
        {
          KTimer tm("lp_betaQ", s);
          k_lp_betaQ<<<nbr, kBlock, 0, s>>>(rm, L, lambda, alpha_.get(), nullptr, gm, U_.get(), curQ_.get(),
                                            lQ_.get(), mmparts_.get());
          k_lp_minmax_final<<<B, 1024, 0, s>>>(mmparts_.get(), nbr, rfirst, mm_.get(), fl);
        }
        for (int k = 0; k < num_pairwise(); ++k) {
          Potential& p = *pw_[k];
          SliceArgs o;
          o.out = sa_.get();
          o.accumulate = k > 0;
          compat_slice(p, o, L, s, [&](const SliceArgs& oo) {
            p.lat->ordered(lQ_.get(), mm_.get(), L, oo, &obins_, k == 0);
          });
        }
        if (num_pairwise() == 0) k_fill<<<grid_for(NL), kBlock, 0, s>>>(sa_.get(), NL, 0.f);
        {
          KTimer tm("lp_gap", s);
          k_lp_gap<<<nbr, kBlock, 0, s>>>(rm, L, alpha_.get(), nullptr, sa_.get(), nullptr, lQ_.get(), parts_.get());
          finalize_seg<2><<<B, kBlock, 0, s>>>(parts_.get(), rfirst, 0, sc + kL_GAP, kL_NSC);
          k_lp_delta<<<1, 256, 0, s>>>(sc, lambda, nullptr, nullptr, 0, 0, B);
        }
      });
    }
    check_err();
    {
      KTimer tm("lp_project", s);
      if (L == 21)
        k_lp_project_px<21><<<nbr, kBlock, 0, s>>>(rm, lQ_.get(), lQ_.get(), lab, 1);
      else
        k_lp_project<<<nbr, kBlock, 0, s>>>(rm, L, lQ_.get(), lQ_.get(), lab, 1);
    }
    int_energy_dev(lab, sc + kL_EINT, kL_NSC);
    kb_lp_outer_decide<<<1, 256, 0, s>>>(sc, ost_.get(), ofl, P.prox_energy_tol, B);
    kb_copy_if<<<nbr, kBlock, 0, s>>>(rm, L, ofl, lQ_.get(), best_.get());
    if (P.less_confident_percent > 0) {
      k_less_confident<<<nbr, kBlock, 0, s>>>(rm, L, best_.get(), P.confidence_tol, ofl + kOF_LC, kOF_NFL);
      kb_lp_less_confident_decide<<<1, 256, 0, s>>>(rm.pix, ofl, P.less_confident_percent, B);
    }
    DCRF_HIP(hipMemcpyAsync(hofl.data(), ofl, hofl.size() * sizeof(int), hipMemcpyDeviceToHost, s));
    DCRF_HIP(hipStreamSynchronize(s));
    bool all = true;
    for (int b = 0; b < B; ++b) all = all && hofl[(size_t)b * kOF_NFL + kOF_STOP] != 0;
    if (all) break;
  }
  DCRF_HIP(hipMemcpyAsync(out, best_.get(), NL * sizeof(float), hipMemcpyDeviceToDevice, s));
  DCRF_HIP(hipStreamSynchronize(s));
  return sweeps;
}

void Batch::assignment_energy_true(const int* lab, double* host) {
  ensure_state();
  DBuf<double> e(B_);
  int_energy_dev(lab, e.get(), 1);
  DCRF_HIP(hipMemcpyAsync(host, e.get(), B_ * sizeof(double), hipMemcpyDeviceToHost, s_));
  DCRF_HIP(hipStreamSynchronize(s_));
}

}  // namespace dcrf
