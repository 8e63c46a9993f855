// Permutohedral lattice on the device: build, plain Gaussian filter and the
// ordered ("upper minus lower") filter.
//
// Reference: Permutohedral (include/permutohedral.h:50-88,
// src/permutohedral.cpp).  The lattice reproduces the reference's point ids
// exactly: ids are assigned in order of first occurrence of a key in the
// entry sequence pixel*(d+1)+remainder, which is the insertion order of the
// reference's HashTable (src/permutohedral.cpp:100-128).
#pragma once

#include <atomic>

#include <memory>
#include <string>
#include <vector>

#include "common.hpp"

namespace dcrf {

struct RestrictedState;
struct DcScratch;

// Input of a splat: value `k` of pixel `pix`.  Concrete forms are defined in
// lattice.hip; the enum selects the one used by a launch.
enum class SplatIn : int {
  kPlain = 0,      // in[pix*V+k]
  kScaled = 1,     // in[pix*V+k] * norm[pix]
  kSxScaled = 2,   // ((k==s[pix]) - Q[pix*V+k]) * norm[pix]   (QP direction)
  kNotOnehot = 3,  // (lab[pix]==k ? 0 : 1)                     (pairwise_energy_true)
  kSx = 4,         // ((k==s[pix]) - Q[pix*V+k])                 (no normalisation)
};

struct SplatArgs {
  SplatIn kind = SplatIn::kPlain;
  const float* in = nullptr;    // kPlain/kScaled/kSx*: N x V field
  const float* norm = nullptr;  // per-pixel normalisation
  const int* lab = nullptr;     // kSx*: argmin label; kNotOnehot: labelling
};

// What the slice does with the filtered value x of (pixel, k).
// Bins of one field shared by the ordered filters of several lattices over
// the same pixels (the pairwise terms of one CRF, or of one batch): the bins,
// the per-part edge flags and bookkeeping, and the listed edge values
struct OrderedBins {
  DBuf<uint8_t> bins;
  DBuf<int> edge;
  DBuf<int2> elist;
  void buffers(std::vector<const void*>& out) const {
    for (const void* p : {(const void*)bins.get(), (const void*)edge.get(), (const void*)elist.get()}) out.push_back(p);
  }
};

struct SliceArgs {
  float* out = nullptr;
  const float* norm = nullptr;  // multiply by norm[pix] (post-normalisation) if set
  float scale = 1.f;            // then multiply by `scale` (Potts: -w) if apply_scale
  bool apply_scale = false;
  bool accumulate = false;      // out = out + x instead of out = x
  float post_mult = 1.f;        // exact multiplier (1 or 2) applied last
};

// Which build of the reference the lattice follows: the SSE2 build compiled on
// x86-64 (0, the default: blocks of four pixels with padding lanes, round half
// to even, sseCompute for more than two values) or the portable build where
// __SSE__ is not defined (1: src/permutohedral.cpp:334-486, seqCompute for
// every value size).  Process-wide, read when a lattice is built.
std::atomic<int>& reference_path();

class Lattice {
 public:
  Lattice(hipStream_t s);
  ~Lattice();
  // Permutohedral::init (src/permutohedral.cpp:152-333) from device features
  // (N x d, feature-contiguous per pixel).
  void build(const float* d_feat, int N, int d);

  int N() const { return N_; }
  int d() const { return d_; }
  int M() const { return M_; }
  int Npad() const { return Np_; }

  // Permutohedral::compute (src/permutohedral.cpp:1079-1087) with the input
  // and output transforms of DenseKernel::filter fused in.  V<=2 follows the
  // reference's seqCompute rounding (:488-540), V>2 its sseCompute (:998-1058).
  void filter(const SplatArgs& in, const SliceArgs& out, int V, bool reverse);

  // seqCompute_upper_minus_lower_ord (src/permutohedral.cpp:590-701) on the
  // rescaled field q = (Q - qmin) / qmax (qmax == 0 -> no division), read on
  // the fly from Q; then out (+)= scale * result.
  // With `shared`, the bins live there: `fresh` computes them from Q (the
  // first lattice of a group), otherwise the ones of the same Q and {min,
  // max} computed by an earlier call of the group are used.
  void ordered(const float* Q, const float* d_minmax, int L, const SliceArgs& out, OrderedBins* shared = nullptr,
               bool fresh = true);

  // The disjoint union of several lattices of the same dimension (one per
  // image): pixels and points are renumbered image after image, so filtering
  // the union equals filtering every part.  The ordered filter then takes one
  // {min, max} pair per part.
  static std::unique_ptr<Lattice> concat(const std::vector<const Lattice*>& parts, hipStream_t s);

  // storage of the ordered filter's lattice rows: float (default) or
  // bfloat16 (arithmetic stays float; half the row bytes)
  void set_bf16_rows(bool on) { bf16_ = on; }
  bool bf16_rows() const { return bf16_; }

  // addresses of the work buffers a filter call writes (value rows, chunk
  // partials, ordered-filter bins and totals); they change when a call with
  // wider rows reallocates them, which invalidates captured launch sequences
  void work_buffers(std::vector<const void*>& out) const {
    for (const void* p : {(const void*)vA_.get(), (const void*)vB_.get(), (const void*)part_.get(),
                          (const void*)bins_.get(), (const void*)tb_.get(), (const void*)edge_.get(),
                          (const void*)vtx_.get(), (const void*)elist_.get(), (const void*)dfix_.get(),
                          (const void*)dvA_.get(), (const void*)dvB_.get(), (const void*)orow_.get(),
                          (const void*)onbr_.get(), (const void*)oppart_.get(), (const void*)tbs_.get(),
                          (const void*)cA_.get(), (const void*)cB_.get(), (const void*)pcode_.get(),
                          (const void*)ptb_.get()})
      out.push_back(p);
  }

  // seqCompute_upper_minus_lower_ord_restricted (src/permutohedral.cpp:702-880):
  // the ordered filter at the pixels pI (ascending, nI of them, device
  // pointer) with values rq (nI x V); the other pixels take their values from
  // ext (N x V, read only when `store`).  `store` keeps the splat of the other
  // pixels and the active points for later calls without it.  Adds into out
  // (nI x V).  Not re-entrant: the state lives in the lattice.
  void ordered_restricted(const float* rq, const int* pI, int nI, const float* ext, int V, bool store, float* out);

  // Permutohedral::compute_upper_minus_lower_dc (src/permutohedral.cpp:881-997)
  // on one value per pixel: out (N, device) += the upper pass of [mh, high)
  // into [low, ml), -= the lower pass of [low, ml) into [mh, high).
  void compute_upper_minus_lower_dc(float* out, int low, int ml, int mh, int high);
  // the same step with caller-held scratch (zero outside a step), enqueued only
  void dc_step(float* out, int low, int ml, int mh, int high, DcScratch& sc);

  // Permutohedral::gradient (src/permutohedral.cpp:1095-1178): df (N x d,
  // device) = d/df a^T K b for a, b N x V.
  void gradient(const float* a, const float* b, int V, float* df);

  // copies of the lattice tables for parity checks (host pointers, may be null)
  void export_tables(int* offset, int* rank, float* bary, short* keys, int* nbr) const;

  // device tables (read by solver kernels)
  const int* offset() const { return offset_.get(); }
  const float* bary() const { return bary_.get(); }
  hipStream_t stream() const { return stream_; }
  size_t bytes_values(int V) const { return (size_t)(M_ + 1) * V * sizeof(float); }

 private:
  void ensure_values(size_t floats, bool zero_row0 = true);
  const float* splat_blur(const SplatArgs& in, int V, bool reverse, bool seq, const std::string& tag);
  template <class T>
  void ordered_t(const float* Q, const float* d_minmax, int L, const SliceArgs& out, OrderedBins* shared,
                 bool fresh);
  // the same with the lower rows held as per-(point, label) codes and the
  // rows of mixed labels only (float rows; ordered_coded.hip)
  void ordered_coded(const float* Q, const float* d_minmax, int L, const SliceArgs& out, OrderedBins* shared,
                     bool fresh);
  void ensure_coded_tables(int L);
  bool bf16_ = false;
  void ensure_total_weight();
  hipStream_t stream_;
  int N_ = 0, d_ = 0, Np_ = 0, M_ = 0;
  bool scalar_ = false;  // built as the reference's portable build (reference_path() == 1)
  int nchunks_ = 0, nmulti_ = 0, npart_ = 0;
  DBuf<int> offset_, rank_;
  DBuf<float> bary_;
  DBuf<int16_t> keys_;
  DBuf<int> nbr_;
  // splat gather structure: entries sorted by (point id, pixel)
  DBuf<int> sp_pix_;
  DBuf<float> sp_w_;
  DBuf<int> seg_;      // M+1
  DBuf<int> chstart_;  // M+1: first chunk of each point
  DBuf<int> chpoint_;  // point of each chunk
  DBuf<int> multi_;    // points with more than one chunk
  DBuf<int> pstart_;   // M+1: first partial-sum slot of each multi-chunk point
  // value buffers (row 0 is the zero row standing for absent neighbours)
  DBuf<float> vA_, vB_, part_;
  size_t vcap_ = 0;
  // ordered filter: blurred total weight per point and the bin-edge flag
  DBuf<float> tb_;
  DBuf<int> edge_;
  DBuf<uint8_t> bins_;
  // ordered slice: each (pixel, vertex) as {point row, weight, total weight},
  // built at the lattice's first ordered call
  void ensure_ordered_tables();
  bool have_otab_ = false;
  DBuf<float4> vtx_;
  // Row order of the ordered state: each part's points sorted by the Morton
  // code of their keys, so lattice neighbours sit close in memory (and in one
  // XCD's L2).  orow_: point -> row (1-based, row 0 is the zero row); onbr_:
  // the blur neighbours in rows ((d+1) x M pairs, row - 1, -1 if absent);
  // oppart_: part of each row (union lattices).
  DBuf<int> orow_, onbr_;
  DBuf<uint16_t> oppart_;
  std::vector<int> part_pt_off_;  // first point of each part and M (union lattices)
  // bin-edge correction of the ordered filter's upper pass: listed edge
  // values, fixed-point sums, correction rows (ping-pong through the blur)
  DBuf<int2> elist_;
  DBuf<unsigned long long> dfix_;
  DBuf<float> dvA_, dvB_;
  DBuf<uint16_t> pimg_;   // part of each pixel (union lattices only)
  DBuf<uint16_t> ppimg_;  // part of each lattice point (union lattices only)
  int nparts_ = 1;
  bool have_tb_ = false;
  // coded ordered filter: the total weight after the splat and after each
  // blur axis in row order ((d+2) x (M+1)); per-(row, label) codes of the two
  // row buffers; codes and total weights of the chunk partials
  DBuf<float> tbs_;
  bool have_tbs_ = false;
  DBuf<uint8_t> cA_, cB_, pcode_;
  DBuf<float> ptb_;
  size_t ccap_ = 0, pccap_ = 0;
  std::shared_ptr<RestrictedState> rs_;  // restricted ordered filter state
};

}  // namespace dcrf
