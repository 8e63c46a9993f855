// Drives the C++ interface (include/densecrf_ho_amd.hpp) the way the
// reference's example_inference does (src/example_inference.cpp:15-88):
// build a DenseCRF2D with a spatial and a bilateral Potts kernel, initialise
// from the unaries, run the QP and then the LP, print the energies.
//
//   api_example W H L unary.f32 image.u8 out_qp.f32 out_lp.f32
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <stdexcept>
#include <vector>

#include "densecrf_ho_amd.hpp"

using namespace dcrf_ho;

template <class T>
static std::vector<T> read_file(const char* p, size_t n) {
  std::vector<T> v(n);
  std::ifstream f(p, std::ios::binary);
  if (!f.read(reinterpret_cast<char*>(v.data()), n * sizeof(T))) throw std::runtime_error(std::string("read ") + p);
  return v;
}
static void write_mat(const char* p, const MatrixXf& m) {
  std::ofstream f(p, std::ios::binary);
  f.write(reinterpret_cast<const char*>(m.data()), m.size() * sizeof(float));
}

int main(int argc, char** argv) {
  if (argc != 8) {
    std::fprintf(stderr, "usage: %s W H L unary.f32 image.u8 out_qp.f32 out_lp.f32\n", argv[0]);
    return 2;
  }
  const int W = std::atoi(argv[1]), H = std::atoi(argv[2]), L = std::atoi(argv[3]), N = W * H;
  std::vector<float> u = read_file<float>(argv[4], (size_t)N * L);
  std::vector<unsigned char> im = read_file<unsigned char>(argv[5], (size_t)N * 3);
  MatrixXf unary(L, N);
  std::copy(u.begin(), u.end(), unary.data());

  DenseCRF2D crf(W, H, L);
  crf.setUnaryEnergy(new ConstUnaryEnergy(unary));
  crf.addPairwiseGaussian(3, 3, new PottsCompatibility(3));
  crf.addPairwiseBilateral(50, 50, 15, 15, 15, im.data(), new PottsCompatibility(5));
  MatrixXf Q = crf.unary_init();
  MatrixXf Qqp = crf.qp_inference_super_pixels_non_convex(Q);
  LP_inf_params params;
  params.prox_max_iter = 2;
  MatrixXf Qlp = crf.lp_inference_prox_super_pixels(Qqp, params);
  std::printf("qp_energy %.6f\n", crf.assignment_energy_true(crf.currentMap(Qqp)));
  std::printf("lp_energy %.6f\n", crf.assignment_energy_true(crf.currentMap(Qlp)));
  std::printf("qp_value %.6f\n", crf.compute_LR_QP_value(Qqp, MatrixXf(L, N)));
  write_mat(argv[6], Qqp);
  write_mat(argv[7], Qlp);

  // the reference's exception for an ordered filter input outside [0,1]
  MatrixXf f(2, N);
  for (int j = 0; j < H; ++j)
    for (int i = 0; i < W; ++i) {
      f(0, j * W + i) = i / 3.f;
      f(1, j * W + i) = j / 3.f;
    }
  PairwisePotential pw(f, new PottsCompatibility(3));
  MatrixXf bad = Q, out;
  bad(0, 0) = 1.5f;
  try {
    pw.apply_upper_minus_lower_ord(out, bad);
    std::printf("no exception\n");
  } catch (const std::runtime_error& e) {
    std::printf("exception %s\n", e.what());
  }

  // unrolled mean field with a Matrix compatibility term and its adjoint
  DenseCRF2D mf(W, H, L);
  mf.setUnaryEnergy(unary);
  mf.addPairwiseGaussian(3, 3, new PottsCompatibility(3));
  MatrixXf m(L, L);
  for (int j = 0; j < L; ++j)
    for (int i = 0; i < L; ++i) m(i, j) = i == j ? -5.f : 0.1f * (float)((i + 2 * j) % 5);
  mf.addPairwiseBilateral(50, 50, 15, 15, 15, im.data(), new MatrixCompatibility(m));
  MatrixXf Qmf = mf.meanfield(3);
  MatrixXf dQ(L, N);
  for (int i = 0; i < N; ++i)
    for (int l = 0; l < L; ++l) dQ(l, i) = l == i % L ? -1.f / Qmf(l, i) : 0.f;
  std::vector<double> dparams;
  MatrixXf dU = mf.meanfield_backward(dQ, &dparams);
  std::printf("meanfield_params %d\n", (int)dparams.size());
  std::printf("meanfield_dU00 %.9g\n", dU(0, 0));

  // a batch of two images (the same image twice) against the one-image QP
  DenseCRFBatch batch(L, {W, W}, {H, H});
  batch.setUnaryEnergy(0, unary);
  batch.setUnaryEnergy(1, unary);
  batch.addPairwiseGaussian(3, 3, 3);
  batch.addPairwiseBilateral(50, 50, 15, 15, 15, {im.data(), im.data()}, 5);
  MatrixXf Qb = batch.qp(5);
  MatrixXf Q5 = crf.qp_inference_super_pixels_non_convex(Q, 5);
  double diff = 0;
  for (int b = 0; b < 2; ++b)
    for (int i = 0; i < N; ++i)
      for (int l = 0; l < L; ++l) diff = std::max(diff, (double)std::abs(Qb(l, b * N + i) - Q5(l, i)));
  std::printf("batch_qp_diff %.9g\n", diff);

  // the other filters of PairwisePotential and Permutohedral on the spatial
  // features: the restricted filter with every pixel equals the ordered
  // filter when N % 4 == 0; the brute-force and divide-and-conquer filters
  // agree on a sorted order; the kernel gradient of a DIAG kernel has d values
  MatrixXf q01(L, N);
  for (int i = 0; i < N; ++i)
    for (int l = 0; l < L; ++l) q01(l, i) = (float)((i * 7 + l * 3) % 11) / 10.f;
  std::vector<int> all(N);
  for (int i = 0; i < N; ++i) all[i] = i;
  MatrixXf ord, rest(L, N);
  pw.apply_upper_minus_lower_ord(ord, q01);
  pw.apply_upper_minus_lower_ord_restricted(rest, q01, all, q01, true);
  double rdiff = 0, omax = 0;
  for (int i = 0; i < N; ++i)
    for (int l = 0; l < L; ++l) {
      rdiff = std::max(rdiff, (double)std::abs(ord(l, i) - rest(l, i)));
      omax = std::max(omax, (double)std::abs(ord(l, i)));
    }
  std::printf("restricted_vs_ordered %.9g\n", rdiff / omax);
  MatrixXi ind(L, N);
  for (int l = 0; l < L; ++l)
    for (int i = 0; i < N; ++i) ind(l, i) = (i * (2 * l + 1)) % N;  // pixel orders (any indices in [0, N))
  MatrixXf bf, dc;
  pw.apply_upper_minus_lower_bf(bf, ind);
  pw.apply_upper_minus_lower_dc(dc, ind);
  std::printf("bf_mid %.9g\n", (double)bf(0, N / 2));
  std::printf("dc_mid %.9g\n", (double)dc(0, N / 2));
  PairwisePotential kp(f, new PottsCompatibility(3), DIAG_KERNEL);
  std::vector<float> kg = kp.kernelGradient(q01, Q);
  std::printf("kernel_gradient_size %d\n", (int)kg.size());
  std::vector<float> df((size_t)N * 2);
  Permutohedral lat;
  lat.init(f);
  lat.gradient(df.data(), q01.data(), Q.data(), L);
  double dfs = 0;
  for (float v : df) dfs += std::abs(v);
  std::printf("lattice_gradient_abs %.9g\n", dfs);

  // LogisticUnaryEnergy (src/unary.cpp:53-68): the unary L f of F = 6 per-pixel
  // features; setUnaryEnergy(L, f) runs the QP on the same unary as a
  // ConstUnaryEnergy holding get(), and setParameters reaches the next solve
  const int F = 6;
  MatrixXf feat(F, N), Lw(L, F);
  for (int i = 0; i < N; ++i) {
    feat(0, i) = 1.f;
    feat(1, i) = (i % W) / (float)W;
    feat(2, i) = (i / W) / (float)H;
    for (int c = 0; c < 3; ++c) feat(3 + c, i) = im[(size_t)i * 3 + c] / 255.f;
  }
  for (int k = 0; k < F; ++k)
    for (int l = 0; l < L; ++l) Lw(l, k) = 0.5f * (float)(((l * 5 + k * 3) % 7) - 3);
  LogisticUnaryEnergy* lu = new LogisticUnaryEnergy(Lw, feat);
  const MatrixXf Ul = lu->get();
  double umax = 0;
  for (long i = 0; i < Ul.size(); ++i) umax = std::max(umax, (double)std::abs(Ul.data()[i]));
  std::printf("logistic_get_max %.9g\n", umax);
  std::printf("logistic_params %d\n", (int)lu->parameters().rows());
  std::printf("logistic_gradient_size %d\n", (int)lu->gradient(q01).rows());
  DenseCRF2D lc(W, H, L), cc(W, H, L);
  lc.setUnaryEnergy(lu);
  cc.setUnaryEnergy(new ConstUnaryEnergy(Ul));
  for (DenseCRF2D* c : {&lc, &cc}) {
    c->addPairwiseGaussian(3, 3, new PottsCompatibility(3));
    c->addPairwiseBilateral(50, 50, 15, 15, 15, im.data(), new PottsCompatibility(5));
  }
  const MatrixXf ql = lc.qp_inference_super_pixels_non_convex(lc.unary_init(), 5);
  const MatrixXf qc = cc.qp_inference_super_pixels_non_convex(cc.unary_init(), 5);
  double ldiff = 0;
  for (long i = 0; i < ql.size(); ++i) ldiff = std::max(ldiff, (double)std::abs(ql.data()[i] - qc.data()[i]));
  std::printf("logistic_vs_const_qp_diff %.9g\n", ldiff);
  VectorXf p2 = lc.getUnaryEnergy()->parameters();
  for (long i = 0; i < p2.rows(); ++i) p2(i) *= 2.f;
  lc.getUnaryEnergy()->setParameters(p2);
  const MatrixXf u2 = lc.unary_init(), u1 = cc.unary_init();
  double udiff = 0;
  for (long i = 0; i < u2.size(); ++i) udiff = std::max(udiff, (double)std::abs(u2.data()[i] - u1.data()[i]));
  std::printf("logistic_set_parameters_changes_unary %d\n", udiff > 1e-3 ? 1 : 0);

  // addPairwiseEnergy(PairwisePotential*) (src/densecrf.cpp:161-168) adds the same
  // term as addPairwiseEnergy(features, function): equal filtered fields
  DenseCRF a(N, L), b(N, L);
  a.setUnaryEnergy(unary);
  b.setUnaryEnergy(unary);
  a.addPairwiseEnergy(f, new MatrixCompatibility(m));
  b.addPairwiseEnergy(new PairwisePotential(f, new MatrixCompatibility(m), DIAG_KERNEL, NORMALIZE_SYMMETRIC));
  const MatrixXf ma = a.meanfield(2), mb = b.meanfield(2);
  double pdiff = 0;
  for (long i = 0; i < ma.size(); ++i) pdiff = std::max(pdiff, (double)std::abs(ma.data()[i] - mb.data()[i]));
  std::printf("potential_object_vs_features_diff %.9g\n", pdiff);
  std::printf("potential_object_params %d\n", (int)b.pairwise_parameters(0).size());
  return 0;
}
