// Batched dense CRF (many images of uneven sizes, one lattice union per
// pairwise term, per-image solver control on the device); see batch.hip.
#pragma once

#include <functional>
#include <map>
#include <memory>
#include <set>
#include <vector>

#include "crf.hpp"
#include "crf_kernels.hpp"

namespace dcrf {

// blocks of a batched launch: image b owns blocks [first[b], first[b+1])
struct BlockPlan {
  int nblocks = 0;
  std::vector<int> first_h;
  DBuf<int> img, first;
  void build(const std::vector<int>& per_image, hipStream_t s);
  ImageBlocks map(const long* pix) const { return ImageBlocks{img.get(), first.get(), pix}; }
};

// Fixed launch sequences captured once into HIP graphs and replayed.  A
// sequence runs eagerly the first time (so every buffer it touches is
// allocated and every table built), is captured the second time, and from
// then on replays.  Sequences that read a double-buffered array are keyed by
// which buffer is current.
//
// A graph holds the device addresses it was captured with.  `buffers` is the
// list of addresses of every buffer a captured sequence may touch; when it
// differs from the list at capture time (a buffer was reallocated, e.g. a
// lattice's value rows grew for the ordered filter) every graph is dropped
// and captured again, so no replay can reach freed memory.
struct GraphCache {
  std::map<long, hipGraphExec_t> exec;
  std::set<long> seen;
  std::vector<const void*> captured_with;
  long captures = 0;
  ~GraphCache();
  void clear();
  void run(long key, hipStream_t s, bool use, const std::vector<const void*>& buffers,
           const std::function<void()>& body);
};

class Batch {
 public:
  Batch(int L, const std::vector<int>& npix);
  ~Batch();
  int L() const { return L_; }
  int B() const { return B_; }
  long N() const { return N_; }
  const std::vector<long>& pixel_offsets() const { return pix_h_; }
  hipStream_t stream() const { return s_; }
  void set_graphs(bool on) {
    use_graph_ = on;
    graphs_.clear();
  }
  long graph_captures() const { return graphs_.captures; }
  // ordered-filter lattice rows in bfloat16 (true) or float
  void set_bf16_rows(bool on) {
    bf16_ = on;
    for (auto& p : pw_) p->lat->set_bf16_rows(on);
    graphs_.clear();
  }

  void set_unary(int b, const float* dU);                                          // device, N_b x L
  void add_pairwise(const std::vector<const float*>& feats, int d, float w, int ntype);  // device, per image
  int num_pairwise() const { return (int)pw_.size(); }
  Potential& pairwise(int k) { return *pw_.at(k); }
  void unary_init(float* Q);

  // solvers over the union (device pointers); init may be null (softmax(-U))
  std::vector<int> qp(const float* init, float* out, int max_iter, bool ref_stop);
  // returns per image the qp-gamma sweep count of every inner step
  std::vector<std::vector<int>> lp(const float* init, float* out, const LPParams& P);
  void assignment_energy_true(const int* lab, double* host);  // per image

 private:
  static constexpr int kSweepBatchB = 16;
  static constexpr long kGraphQP = 1, kGraphY = 2, kGraphPost = 10, kGraphSweep = 100;
  void ensure_state();
  // addresses of every buffer a captured solver sequence may touch
  std::vector<const void*> work_buffers() const;
  void run_graph(long key, const std::function<void()>& body) {
    graphs_.run(key, s_, use_graph_, work_buffers(), body);
  }
  ImageBlocks red_map() const;
  ImageBlocks sweep_map() const;
  void int_energy_dev(const int* lab, double* dst, int stride);
  void lp_energy_dev(const float* Q, double* dst, double* scratch, int stride);

  int L_, B_;
  long N_ = 0;
  hipStream_t s_ = nullptr;
  std::vector<long> pix_h_;
  DBuf<long> pix_;
  BlockPlan red_, sweep_;
  DBuf<float> U_;
  std::vector<std::unique_ptr<Potential>> pw_;
  bool use_graph_ = true;
  bool bf16_ = false;
  GraphCache graphs_;
  // QP state
  DBuf<float> Q_, gy_, psx_, tmp_, rpix_;
  OrderedBins obins_;  // bins of the field the ordered filters of all pairwise terms share
  DBuf<int> lab_;
  DBuf<double> parts_, sc_, swtot_;  // swtot_: per-image totals of a sweep batch
  DBuf<int> fl_;
  // LP state
  DBuf<float> alpha_, sa_, gamma_, gamma2_, Y_, lQ_, curQ_, best_, mmparts_, mm_;
  DBuf<double> lsc_, ost_;
  DBuf<int> lfl_, ofl_;
  HBuf<int> hstatus_;  // the sweep decision's status (pinned host memory)
};

}  // namespace dcrf
