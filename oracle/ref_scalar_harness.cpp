// C ABI over the reference's Permutohedral compiled WITHOUT its SSE code
// (src/permutohedral.cpp:334-486, the portable init, and compute() through
// seqCompute for every value size, :1059-1064): what the reference computes
// when built where __SSE__ is not defined.  TEST INFRASTRUCTURE ONLY: built
// from the reference sources into oracle/_ref/ (see oracle/Makefile), in the
// development container only; the fixtures it writes are what the tests use.
//
// Protected members are reached by deriving from the class; no reference code
// is restated here.

#include <algorithm>

#include "permutohedral.h"

namespace {

struct ScalarLattice : public Permutohedral {
  int M() const { return M_; }
  void get(int* offset, int* rank, float* bary, int* nbr, int n) const {
    const int d1 = d_ + 1;
    for (int i = 0; i < n * d1; ++i) {
      offset[i] = offset_[i];
      rank[i] = rank_[i];
      bary[i] = barycentric_[i];
    }
    for (int i = 0; i < d1 * M_; ++i) {
      nbr[2 * i] = blur_neighbors_[i].n1;
      nbr[2 * i + 1] = blur_neighbors_[i].n2;
    }
  }
};

Eigen::MatrixXf as_mat(const float* p, int rows, int cols) {
  Eigen::MatrixXf m(rows, cols);
  std::copy(p, p + (size_t)rows * cols, m.data());
  return m;
}

}  // namespace

extern "C" {

void* refs_lattice_new(const float* feat, int N, int d) {
  ScalarLattice* L = new ScalarLattice();
  L->init(as_mat(feat, d, N));
  return L;
}
void refs_lattice_free(void* h) { delete (ScalarLattice*)h; }
int refs_lattice_M(void* h) { return ((ScalarLattice*)h)->M(); }
void refs_lattice_get(void* h, int* offset, int* rank, float* bary, int* nbr, int n) {
  ((ScalarLattice*)h)->get(offset, rank, bary, nbr, n);
}
void refs_lattice_compute(void* h, const float* in, float* out, int V, int N, int reverse) {
  Eigen::MatrixXf i = as_mat(in, V, N), o;
  ((ScalarLattice*)h)->compute(o, i, reverse != 0);
  std::copy(o.data(), o.data() + (size_t)V * N, out);
}

void refs_lattice_ordered(void* h, const float* q, float* out, int V, int N) {
  Eigen::MatrixXf i = as_mat(q, V, N), o(V, N);
  o.fill(0);
  ((ScalarLattice*)h)->compute_upper_minus_lower_ord(o, i);
  std::copy(o.data(), o.data() + (size_t)V * N, out);
}

}  // extern "C"
