/*
 * densecrf_ho_amd: C ABI of the MI355X (gfx950) dense-CRF engine.
 *
 * Drop-in boundary for the hot path of thwjoy/densecrf_ho: the
 * Permutohedral lattice filter, PairwisePotential, and the DenseCRF QP / LP
 * solvers with Pn superpixel terms.  Each entry point names the reference
 * interface it replaces (file:line in the reference repository).
 *
 * Conventions
 *  - Every function returns DCRF_OK (0) or an error code; dcrf_last_error()
 *    returns the message of the calling thread's last failure.  The messages
 *    of the reference's exceptions are kept ("Value out of bounds",
 *    "Maxval < 0, skipping").
 *  - Layouts are the reference's Eigen column-major ones:
 *      fields (Q, unaries, filter in/out)  L x N, element [pixel*L + label]
 *      features                             d x N, element [pixel*d + k]
 *      labellings                           N int32
 *      image                                N x 3 uint8 RGB, pixel = y*W + x
 *  - `on_device` != 0 means the pointers are device (HIP) pointers on the
 *    current device; otherwise they are host pointers and the call copies.
 *  - Calls are synchronous: results are complete when the call returns.
 *  - A handle is used by one thread at a time (like the reference objects,
 *    whose ordered filter mutates member buffers).
 */
#ifndef DENSECRF_HO_AMD_H
#define DENSECRF_HO_AMD_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DCRF_OK 0
#define DCRF_INVALID_ARGUMENT 1
#define DCRF_OUT_OF_BOUNDS 2   /* std::runtime_error("Value out of bounds"), src/pairwise.cpp:218-219 */
#define DCRF_MAXVAL_NEGATIVE 3 /* std::runtime_error("Maxval < 0, skipping"), src/eigen_utils.cpp:138 */
#define DCRF_HIP_ERROR 4
#define DCRF_INTERNAL 5

/* NormalizationType, include/pairwise.h:34-39 */
#define DCRF_NO_NORMALIZATION 0
#define DCRF_NORMALIZE_BEFORE 1
#define DCRF_NORMALIZE_AFTER 2
#define DCRF_NORMALIZE_SYMMETRIC 3

/* LabelCompatibility kinds, include/labelcompatibility.h:51-92.  Parameters:
 * Potts w (1 value); Diagonal v (L values); Matrix m (L*L, column-major; stored
 * as (m + m^T)/2 like the MatrixCompatibility constructor,
 * src/labelcompatibility.cpp:94-97). */
#define DCRF_COMPAT_POTTS 0
#define DCRF_COMPAT_DIAGONAL 1
#define DCRF_COMPAT_MATRIX 2

typedef struct dcrf_lattice_s* dcrf_lattice_t;
typedef struct dcrf_pairwise_s* dcrf_pairwise_t;
typedef struct dcrf_crf_s* dcrf_crf_t;
typedef struct dcrf_batch_s* dcrf_batch_t;

const char* dcrf_last_error(void);
/* library version string */
const char* dcrf_version(void);

/* Which build of the reference the lattices follow.  The reference's
 * Permutohedral compiles one of two paths (src/permutohedral.cpp:36-40):
 *   0  the SSE2 build, what x86-64 compilers produce (the default): init in
 *      blocks of four pixels with padding lanes and round-half-even
 *      (:152-333), compute() through sseCompute for more than two values
 *      (:998-1058);
 *   1  the portable build, where __SSE__ is not defined: init one pixel at a
 *      time with ties rounded down (:334-486), compute() through seqCompute
 *      for every value size (:488-540, :1059-1064).
 * Process-wide; read when a lattice is built (dcrf_lattice_create, and the
 * potentials of dcrf_pairwise_create / dcrf_crf_add_pairwise*).  A lattice
 * keeps the path it was built with.  Only the lattice follows the switch:
 * column sums (renormalize, feasible_Q, qp-gamma) keep the order of Eigen's
 * 4-float packets either way.  Returns nonzero for other values. */
int dcrf_set_reference_path(int portable);
int dcrf_reference_path(void);

/* ------------------------------------------------------------------------ */
/* Permutohedral lattice (include/permutohedral.h:50-88)                     */
/* ------------------------------------------------------------------------ */

/* Permutohedral::init(const MatrixXf& features), src/permutohedral.cpp:152-333.
 * Builds the lattice of N feature vectors of dimension d (1 <= d <= 12). */
int dcrf_lattice_create(const float* features, int N, int d, int on_device, dcrf_lattice_t* out);
int dcrf_lattice_destroy(dcrf_lattice_t lat);
/* number of lattice points (Permutohedral::M_) and padded pixel count
 * (N rounded up to a multiple of 4, the rows of offset/rank/barycentric) */
int dcrf_lattice_info(dcrf_lattice_t lat, int* M, int* Npad);
/* copies of offset_, rank_, barycentric_ ((d+1) x Npad each), the lattice
 * keys (d x M, int16) and blur_neighbors_ ((d+1) x M pairs n1,n2; -1 absent),
 * src/permutohedral.cpp:65-67,311-330.  Host pointers; any may be NULL. */
int dcrf_lattice_export(dcrf_lattice_t lat, int* offset, int* rank, float* barycentric, short* keys, int* neighbors);
/* Permutohedral::compute(MatrixXf& out, const MatrixXf& in, bool reverse),
 * include/permutohedral.h:80-81, src/permutohedral.cpp:1079-1087.
 * in/out are V x N; out may alias in. */
int dcrf_lattice_compute(dcrf_lattice_t lat, const float* in, float* out, int V, int reverse, int on_device);
/* Permutohedral::compute_upper_minus_lower_ord(MatrixXf& out, const MatrixXf& Q),
 * include/permutohedral.h:83, src/permutohedral.cpp:590-701 (10-bin ordered
 * filter, no normalisation, no compatibility).  Q is V x N with values in [0,1]. */
int dcrf_lattice_compute_upper_minus_lower_ord(dcrf_lattice_t lat, const float* Q, float* out, int V, int on_device);
/* Permutohedral::gradient(float* df, const float* a, const float* b, int value_size),
 * include/permutohedral.h:87, src/permutohedral.cpp:1095-1178: df (d x N) = d/df of
 * a^T K b for a, b (V x N), K the unnormalised lattice filter. */
/* Permutohedral::compute_upper_minus_lower_dc(MatrixXf& out, int low, int middle_low,
 * int middle_high, int high), include/permutohedral.h:82, src/permutohedral.cpp:881-997
 * and 1066-1070: one value per pixel; out (N) gets += the pass of [middle_high, high)
 * sliced into [low, middle_low) and -= the pass of [low, middle_low) sliced into
 * [middle_high, high), blurred over those pixels' lattice points only. */
int dcrf_lattice_compute_upper_minus_lower_dc(dcrf_lattice_t lat, float* out, int low, int middle_low,
                                              int middle_high, int high, int on_device);
int dcrf_lattice_gradient(dcrf_lattice_t lat, float* df, const float* a, const float* b, int V, int on_device);
/* Permutohedral::compute_upper_minus_lower_ord_restricted(MatrixXf& rout, const MatrixXf& rQ,
 * const std::vector<int>& pI, const MatrixXf& Q, bool store), include/permutohedral.h:84-85,
 * src/permutohedral.cpp:702-880, 1074-1078.  The ordered filter at the nI pixels pI
 * (strictly ascending) with values rQ (V x nI); every other pixel takes its values from
 * Q (V x N), which is read only when store != 0.  A call with store keeps the splat of
 * the other pixels and the lattice points that pI touches in the handle; later calls
 * without store reuse both (their pI pixels are added to the stored splat, and only the
 * stored points are blurred).  rout (V x nI) is ADDED to, as the reference does.  rQ and
 * Q must lie in [0,1] (DCRF_OUT_OF_BOUNDS otherwise); a call without store before any
 * call with store, or with another V, fails.  Not re-entrant per handle. */
int dcrf_lattice_compute_upper_minus_lower_ord_restricted(dcrf_lattice_t lat, float* rout, const float* rQ,
                                                           const int* pI, int nI, const float* Q, int V, int store,
                                                           int on_device);

/* ------------------------------------------------------------------------ */
/* PairwisePotential with Potts compatibility (include/pairwise.h:63-95)     */
/* ------------------------------------------------------------------------ */

/* PairwisePotential(features, new PottsCompatibility(w), DIAG_KERNEL, ntype),
 * src/pairwise.cpp:200-202 with DenseKernel::initLattice (:40-57) */
int dcrf_pairwise_create(const float* features, int N, int d, float potts_w, int ntype, int on_device,
                         dcrf_pairwise_t* out);
int dcrf_pairwise_destroy(dcrf_pairwise_t pw);
/* PairwisePotential::apply / applyTranspose, src/pairwise.cpp:203-213; in/out are L x N */
int dcrf_pairwise_apply(dcrf_pairwise_t pw, const float* in, float* out, int L, int transpose, int on_device);
/* PairwisePotential::apply_upper_minus_lower_ord, src/pairwise.cpp:214-228 */
int dcrf_pairwise_apply_upper_minus_lower_ord(dcrf_pairwise_t pw, const float* Q, float* out, int L, int on_device);
/* PairwisePotential::apply_upper_minus_lower_ord_restricted(rout, rQ, pI, Q, store),
 * include/pairwise.h:78-79, src/pairwise.cpp:229-239: rout (L x nI) = 0, the restricted
 * ordered filter (as dcrf_lattice_compute_upper_minus_lower_ord_restricted), then the
 * compatibility (Potts: -w). */
/* KernelType of the potential (0 CONST_KERNEL, the default here as in
 * PairwisePotential's constructor, include/pairwise.h:40-44,71; 1 DIAG_KERNEL;
 * 2 FULL_KERNEL); set before any kernel parameters. */
int dcrf_pairwise_set_kernel_type(dcrf_pairwise_t pw, int ktype);
/* The LabelCompatibility of a standalone potential (PairwisePotential(features, function,
 * ...), src/pairwise.cpp:186-202; include/labelcompatibility.h:51-92): Potts (params[0]),
 * Diagonal (L values) or Matrix (L*L column-major, stored as (m + m^T)/2).  Its filters
 * then take fields of L labels (Potts: any). */
int dcrf_pairwise_set_compatibility(dcrf_pairwise_t pw, int compat_type, const float* params, int L);
/* PairwisePotential::kernelParameters / setKernelParameters, src/pairwise.cpp:396-401
 * (DenseKernel::parameters / setParameters, :136-161): 0, d or d*d values (DIAG: the
 * feature scales, ones at first; FULL: a d x d column-major matrix, the identity at
 * first).  Setting rebuilds the lattice and the normalisation on diag(v) f or V f. */
int dcrf_pairwise_kernel_nparams(dcrf_pairwise_t pw);
int dcrf_pairwise_kernel_parameters(dcrf_pairwise_t pw, float* out);
int dcrf_pairwise_set_kernel_parameters(dcrf_pairwise_t pw, const float* v, int n);
/* PairwisePotential::kernelGradient(b, Q), src/pairwise.cpp:402-407: the gradient of
 * b^T K compatibility(Q) w.r.t. the kernel parameters (DenseKernel::gradient and
 * featureGradient, :90-122, :170-181); b and Q are L x N, out (host) holds
 * dcrf_pairwise_kernel_nparams values. */
int dcrf_pairwise_kernel_gradient(dcrf_pairwise_t pw, const float* b, const float* Q, int L, float* out,
                                  int on_device);
/* PairwisePotential::apply_bf(out, Q), include/pairwise.h:73, src/pairwise.cpp:368-383:
 * the dense Gaussian exp(-|f_c - f_b|^2) over all pixel pairs (no lattice, no
 * normalisation), then the compatibility.  O(N^2 L): for small inputs. */
int dcrf_pairwise_apply_bf(dcrf_pairwise_t pw, const float* Q, float* out, int L, int on_device);
/* PairwisePotential::apply_upper_minus_lower_bf(out, ind), include/pairwise.h:76,
 * src/pairwise.cpp:294-320: per label l, pixels taken in the order ind (L x N, entry
 * (l, c) at ind[c*L + l]); out(l, c) = sum_{b>c} K - sum_{b<c} K, left in that sorted
 * order; then the compatibility. */
int dcrf_pairwise_apply_upper_minus_lower_bf(dcrf_pairwise_t pw, const int* ind, float* out, int L, int on_device);
/* PairwisePotential::apply_upper_minus_lower_bf_ord(out, ind, Q), include/pairwise.h:77,
 * src/pairwise.cpp:321-353: as above with the split by value, out(l, c) =
 * sum_{Q_b <= Q_c} K - sum_{Q_b >= Q_c} K. */
/* PairwisePotential::apply_upper_minus_lower_dc(out, ind), include/pairwise.h:75,
 * src/pairwise.cpp:241-293: per label a lattice on the features in the order ind and
 * the divide-and-conquer over the sorted positions (blocks of <= 30 by brute force with
 * alpha 1/0.6, larger ones halved with the lattice step between the halves); out (L x N)
 * in the sorted order, then the compatibility.  It runs one launch per node (about N/15
 * per label), so N * L is limited to 2^20 (DCRF_INVALID_ARGUMENT above). */
int dcrf_pairwise_apply_upper_minus_lower_dc(dcrf_pairwise_t pw, const int* ind, float* out, int L, int on_device);
int dcrf_pairwise_apply_upper_minus_lower_bf_ord(dcrf_pairwise_t pw, const int* ind, const float* Q, float* out,
                                                 int L, int on_device);
int dcrf_pairwise_apply_upper_minus_lower_ord_restricted(dcrf_pairwise_t pw, float* rout, const float* rQ,
                                                         const int* pI, int nI, const float* Q, int L, int store,
                                                         int on_device);

/* ------------------------------------------------------------------------ */
/* DenseCRF (include/densecrf.h:73-194)                                      */
/* ------------------------------------------------------------------------ */

/* DenseCRF::DenseCRF(int N, int M), src/densecrf.cpp:139 (1 <= L <= 64) */
int dcrf_crf_create(int N, int L, dcrf_crf_t* out);
int dcrf_crf_destroy(dcrf_crf_t crf);
/* DenseCRF::setUnaryEnergy(const MatrixXf& unary), src/densecrf.cpp:261-263 */
int dcrf_crf_set_unary(dcrf_crf_t crf, const float* unary, int on_device);
/* DenseCRF::setUnaryEnergy(const MatrixXf& L, const MatrixXf& f), src/densecrf.cpp:264-266:
 * the unary becomes LogisticUnaryEnergy(Lw, f).get() = Lw f, computed on the device
 * (dcrf_logistic_unary); Lw is L x F, f is F x N. */
int dcrf_crf_set_unary_logistic(dcrf_crf_t crf, const float* Lw, const float* f, int F, int on_device);
/* LogisticUnaryEnergy::get(), src/unary.cpp:56-58 (include/unary.h:49-57): U (L x N) =
 * Lw (L x F) f (F x N), column-major (U[p*L + l], Lw[k*L + l], f[p*F + k]); fp32 matrix
 * cores (v_mfma_f32_16x16x4_f32), fp32 accumulation.  1 <= L <= 64. */
int dcrf_logistic_unary(const float* Lw, const float* f, int L, int F, int N, float* U, int on_device);
/* LogisticUnaryEnergy::gradient(b), src/unary.cpp:64-68: g = b f^T (L x F, column-major,
 * i.e. the reference's result resized to L*F x 1), b L x N.  fp32 matrix-core partial
 * sums over blocks of 4096 pixels, added in block order in FP64. */
int dcrf_logistic_unary_gradient(const float* b, const float* f, int L, int F, int N, float* g, int on_device);
/* DenseCRF::addPairwiseEnergy(features, new PottsCompatibility(w), DIAG_KERNEL, ntype),
 * src/densecrf.cpp:157-169 (also keeps the NO_NORMALIZATION copy used by
 * the energies; here both share one lattice) */
int dcrf_crf_add_pairwise(dcrf_crf_t crf, const float* features, int d, float potts_w, int ntype, int on_device);
/* DenseCRF::addPairwiseEnergy(features, function, DIAG_KERNEL, ntype) with
 * function a PottsCompatibility, DiagonalCompatibility or MatrixCompatibility
 * (src/densecrf.cpp:157-169, src/labelcompatibility.cpp:47-129); `params` is a
 * host array in the layout described at DCRF_COMPAT_POTTS. */
int dcrf_crf_add_pairwise_compat(dcrf_crf_t crf, const float* features, int d, int compat_type,
                                 const float* params, int ntype, int on_device);
/* PairwisePotential::parameters() / setParameters() of potential k
 * (src/pairwise.cpp:382-387; layouts src/labelcompatibility.cpp:61-68, 88-96,
 * 109-121: Potts [w], Diagonal v[L], Matrix the upper triangle, row by row,
 * L*(L+1)/2 values).  `params` is a host array. */
int dcrf_crf_pairwise_nparams(dcrf_crf_t crf, int k, int* n);
int dcrf_crf_pairwise_parameters(dcrf_crf_t crf, int k, float* params);
int dcrf_crf_pairwise_set_parameters(dcrf_crf_t crf, int k, const float* params);
/* PairwisePotential::gradient(b, Q) of potential k, src/pairwise.cpp:390-395:
 * the compatibility's gradient of b against the filtered Q, in the layout of
 * parameters() (host array `grad`, FP64 accumulation rounded to float). */
int dcrf_crf_pairwise_gradient(dcrf_crf_t crf, int k, const float* b, const float* Q, float* grad, int on_device);
/* DenseCRF2D::addPairwiseGaussian(sx, sy, new PottsCompatibility(w)), src/densecrf.cpp:170-178 */
int dcrf_crf_add_pairwise_gaussian(dcrf_crf_t crf, int W, int H, float sx, float sy, float potts_w);
/* DenseCRF2D::addPairwiseBilateral(sx, sy, sr, sg, sb, im, new PottsCompatibility(w)),
 * src/densecrf.cpp:179-190; `image` is a host pointer to H*W*3 RGB bytes */
int dcrf_crf_add_pairwise_bilateral(dcrf_crf_t crf, int W, int H, float sx, float sy, float sr, float sg, float sb,
                                    const unsigned char* image, float potts_w);
/* Superpixel layer from a region map: region[pixel] in [0,R) (host pointer);
 * exp_of_superpixels_ (e) and constants_ (c) per region.  No reference
 * counterpart: DenseCRF2D::addSuperPixel (src/densecrf.cpp:193-251) computes
 * the map with mean-shift and then stores exactly these three things.
 * Regions with no pixel are dropped. */
int dcrf_crf_add_superpixel_layer(dcrf_crf_t crf, const int* region, int R, const float* e, const float* c);
/* PairwisePotential::apply / applyTranspose of potential k, src/pairwise.cpp:203-213;
 * no_norm != 0 applies the NO_NORMALIZATION copy (src/densecrf.cpp:164-168). */
int dcrf_crf_apply(dcrf_crf_t crf, int k, const float* in, float* out, int transpose, int no_norm, int on_device);
/* PairwisePotential::apply_upper_minus_lower_ord(out, Q), src/pairwise.cpp:214-228:
 * fails with DCRF_OUT_OF_BOUNDS when Q leaves [0,1]. */
int dcrf_crf_apply_upper_minus_lower_ord(dcrf_crf_t crf, int k, const float* Q, float* out, int on_device);
/* PairwisePotential::apply_upper_minus_lower_ord_restricted of potential k,
 * src/pairwise.cpp:229-239, with that potential's compatibility (Potts, Diagonal or
 * Matrix); rout is L x nI, Q is L x N. */
/* The KernelType of potential k (0 CONST_KERNEL, 1 DIAG_KERNEL, 2 FULL_KERNEL): the
 * kernel_type argument of DenseCRF::addPairwiseEnergy (src/densecrf.cpp:157-160, which
 * builds PairwisePotential(features, function, kernel_type, ntype), include/pairwise.h:
 * 40-44,71).  DIAG_KERNEL after adding, as the reference's default; set before any
 * kernel parameters.  With the initial (unit) parameters every type filters alike; the
 * type fixes the kernel parameters and their gradient: 0, d or d*d values
 * (dcrf_crf_pairwise_kernel_nparams). */
int dcrf_crf_pairwise_set_kernel_type(dcrf_crf_t crf, int k, int ktype);
/* DenseCRF::addPairwiseEnergy(PairwisePotential* potential), src/densecrf.cpp:161-168: adds
 * the potential of pw (its features, normalisation, compatibility, kernel type and kernel
 * parameters) as the CRF's next term, with the NO_NORMALIZATION Potts copy of its
 * parameters()(0) taken now.  pw itself is not changed; the reference takes ownership of
 * the object, so the C++ layer destroys pw afterwards. */
int dcrf_crf_add_pairwise_potential(dcrf_crf_t crf, dcrf_pairwise_t pw);
int dcrf_crf_pairwise_kernel_nparams(dcrf_crf_t crf, int k, int* n);
/* PairwisePotential::kernelGradient(b, Q) of potential k: dcrf_crf_pairwise_kernel_nparams
 * values into out (host). */
int dcrf_crf_pairwise_kernel_gradient(dcrf_crf_t crf, int k, const float* b, const float* Q, float* out,
                                      int on_device);
/* The brute-force filters of potential k (apply_bf, apply_upper_minus_lower_bf,
 * apply_upper_minus_lower_bf_ord, src/pairwise.cpp:294-383) with its compatibility;
 * layouts as the dcrf_pairwise_ forms, L x N fields. */
int dcrf_crf_apply_bf(dcrf_crf_t crf, int k, const float* Q, float* out, int on_device);
int dcrf_crf_apply_upper_minus_lower_bf(dcrf_crf_t crf, int k, const int* ind, float* out, int on_device);
int dcrf_crf_apply_upper_minus_lower_dc(dcrf_crf_t crf, int k, const int* ind, float* out, int on_device);
int dcrf_crf_apply_upper_minus_lower_bf_ord(dcrf_crf_t crf, int k, const int* ind, const float* Q, float* out,
                                            int on_device);
int dcrf_crf_apply_upper_minus_lower_ord_restricted(dcrf_crf_t crf, int k, float* rout, const float* rQ,
                                                    const int* pI, int nI, const float* Q, int store, int on_device);
/* DenseCRF::unary_init(), src/densecrf.cpp:321-325 */
int dcrf_crf_unary_init(dcrf_crf_t crf, float* Q, int on_device);

/* Per-iteration record of the QP (host arrays owned by the caller). */
typedef struct {
  int iterations;  /* out: iterations entered (the reference's loop counter) */
  int updates;     /* out: iterations that updated Q (all but a final num==0 || denom==0 stop) */
  int capacity;    /* in: length of num/denom/step (energy holds capacity+1) */
  double* energy;  /* energy before the first and after each iteration */
  double* num;
  double* denom;
  double* step;
} dcrf_qp_stats;

/* DenseCRF::qp_inference_super_pixels_non_convex(init), src/densecrf.cpp:340-442.
 * ref_stop != 0: the reference's rule (stop when the energy decreases by at
 * most 100, src/densecrf.cpp:390), with max_iter as an extra cap (0 = none);
 * ref_stop == 0: exactly max_iter iterations unless the line search breaks
 * (denom == 0 or num == 0, :419).  stats may be NULL. */
int dcrf_crf_qp(dcrf_crf_t crf, const float* init, float* out, int max_iter, int ref_stop, int on_device,
                dcrf_qp_stats* stats);

/* LP_inf_params, include/densecrf.h:43-69 (defaults: dcrf_lp_params_default) */
typedef struct {
  float prox_reg_const, dual_gap_tol, prox_energy_tol;
  int prox_max_iter, fw_max_iter, qp_max_iter;
  float qp_tol, qp_const;
  int best_int, accel_prox, work_set_size, approx_fw_iter; /* last three read nowhere by the reference */
  float less_confident_percent, confidence_tol;
} dcrf_lp_params;
/* LP_inf_params::LP_inf_params(), src/densecrf.cpp:287-302 */
void dcrf_lp_params_default(dcrf_lp_params* p);

typedef struct {
  int outer;          /* out: proximal iterations */
  int inner;          /* out: Frank-Wolfe steps recorded */
  int capacity;       /* in: length of the arrays below */
  int* qp_sweeps;     /* per FW step */
  double* dual_gap;   /* per FW step */
  double* delta;      /* per FW step */
  double* int_energy; /* per proximal iteration */
  double init_lp_energy, init_int_energy;
} dcrf_lp_stats;

/* DenseCRF::lp_inference_prox_super_pixels(init, params), src/densecrf.cpp:526-799 */
int dcrf_crf_lp(dcrf_crf_t crf, const float* init, float* out, const dcrf_lp_params* params, int on_device,
                dcrf_lp_stats* stats);

/* DenseCRF::currentMap(Q), src/densecrf.cpp:929-938 */
int dcrf_crf_current_map(dcrf_crf_t crf, const float* Q, int* labels, int on_device);
/* DenseCRF::assignment_energy_true(l), src/densecrf.cpp:873-885 */
int dcrf_crf_assignment_energy_true(dcrf_crf_t crf, const int* labels, int on_device, double* energy);
/* DenseCRF::unaryEnergy(l) / pairwise_energy_true(l, term), include/densecrf.h:150-153,
 * src/densecrf.cpp:887-927 (the reference's public debugging functions): per-pixel
 * float energies into out (N); term -1 sums every pairwise term (no-norm Potts copies). */
int dcrf_crf_unary_energy(dcrf_crf_t crf, const int* labels, float* out, int on_device);
int dcrf_crf_pairwise_energy_true(dcrf_crf_t crf, const int* labels, int term, float* out, int on_device);
/* DenseCRF::assignment_energy_higher_order(l), src/densecrf.cpp:844-870 */
int dcrf_crf_assignment_energy_higher_order(dcrf_crf_t crf, const int* labels, int on_device, double* energy);
/* DenseCRF::compute_energy_higher_order(Q), include/densecrf.h:162, src/densecrf.cpp:960-980:
 * U.Q + sum_k Q.(K_k Q) (FP64 dots) + sum(z .* c) + multiplySuperPixels((1 - z) .* c, Q - 1)
 * with the z_labels_ of the last QP (the region term needs a QP run first when superpixel
 * layers exist; DCRF_INVALID_ARGUMENT otherwise). */
int dcrf_crf_energy_higher_order(dcrf_crf_t crf, const float* Q, int on_device, double* energy);
/* DenseCRF::compute_energy_LP_higher_order(Q), src/densecrf.cpp:982-1037 */
int dcrf_crf_lp_energy(dcrf_crf_t crf, const float* Q, int on_device, double* energy);
/* DenseCRF::computeUCondGrad(Us, Q), src/densecrf.cpp:803-830 (protected in the
 * reference; the LP's superpixel conditional gradient) */
int dcrf_crf_ucondgrad(dcrf_crf_t crf, const float* Q, float* Us, int on_device);
/* DenseCRF::compute_LR_QP_value(Q, 0), src/densecrf.cpp:940-957 */
int dcrf_crf_qp_value(dcrf_crf_t crf, const float* Q, int on_device, double* energy);

/* ------------------------------------------------------------------------ */
/* Unrolled mean field and its adjoint.  The fork has no inference loop: the */
/* update is the classic one built from its pieces, Q0 = expAndNormalize(-U) */
/* (src/densecrf.cpp:321-336), then `iterations` times                       */
/*   Q <- softmax(-U - sum_k PairwisePotential::apply_k(Q))                  */
/* (src/pairwise.cpp:203-208).  The forward keeps Q_0..Q_T and every K_k Q_t */
/* on the device; the backward runs the hand-written reverse pass through    */
/* the softmax, the compatibilities and the transposed filters              */
/* (applyTranspose, src/pairwise.cpp:209-213).                              */
/* ------------------------------------------------------------------------ */
int dcrf_crf_meanfield(dcrf_crf_t crf, int iterations, float* Q, int on_device);
/* dQ = dL/dQ_T of the last dcrf_crf_meanfield -> dU = dL/dU (N x L) and the
 * gradient of every potential's parameters (parameters() layout, potentials
 * back to back, host FP64 array; may be null). */
int dcrf_crf_meanfield_backward(dcrf_crf_t crf, const float* dQ, float* dU, double* dparams, int on_device);
/* meanfield_backward fails (DCRF_INVALID_ARGUMENT) when the unary, a pairwise term or its
 * parameters changed after the last dcrf_crf_meanfield: its checkpoints belong to that
 * forward.  dcrf_crf_meanfield_id numbers the forward runs (1, 2, ...; 0 before the first),
 * so a caller that kept an earlier forward's number can tell that it was overwritten.  No
 * reference counterpart. */
int dcrf_crf_meanfield_id(dcrf_crf_t crf, long long* id);
/* loss = -sum_i log Q(labels[i], i) over pixels with labels[i] in [0,L) (FP64)
 * and dQ = dloss/dQ.  No reference counterpart: the upstream loss of the
 * unrolled-inference gradient. */
int dcrf_crf_cross_entropy(dcrf_crf_t crf, const float* Q, const int* labels, float* dQ, double* loss,
                           int on_device);

/* ------------------------------------------------------------------------ */
/* Batches: many images of uneven sizes solved together.  No reference       */
/* counterpart (the reference makes one DenseCRF per image and runs them one */
/* after another); every image follows the per-image DenseCRF algorithm      */
/* (qp_inference_super_pixels_non_convex, lp_inference_prox_super_pixels     */
/* with best_int; no superpixel layers).  Fields of a batch are the images'  */
/* fields concatenated: image b owns pixels [offsets[b], offsets[b+1]).      */
/* ------------------------------------------------------------------------ */
int dcrf_batch_create(int L, int B, const int* npix, dcrf_batch_t* out);
/* images of W[b] x H[b] pixels (needed for the 2-D kernels below) */
int dcrf_batch_create_2d(int L, int B, const int* W, const int* H, dcrf_batch_t* out);
int dcrf_batch_destroy(dcrf_batch_t b);
/* offsets[0..B] (long long) of every image's first pixel in the batch fields */
int dcrf_batch_pixel_offsets(dcrf_batch_t b, long long* offsets);
/* replay the fixed parts of every solver iteration from HIP graphs (default on) */
int dcrf_batch_set_graphs(dcrf_batch_t b, int on);
/* storage of the ordered filter's lattice rows (arithmetic is float either
 * way): DCRF_VALUES_F32 (default) or DCRF_VALUES_BF16 (half the bytes;
 * results move by the rounding of the stored rows).  No reference
 * counterpart.  The CRF form applies to the terms added so far. */
#define DCRF_VALUES_F32 0
#define DCRF_VALUES_BF16 1
int dcrf_batch_set_value_type(dcrf_batch_t b, int type);
int dcrf_crf_set_value_type(dcrf_crf_t crf, int type);
/* number of HIP graphs captured so far (a graph is captured again whenever a
 * buffer it refers to was reallocated) */
int dcrf_batch_graph_captures(dcrf_batch_t b, long long* n);
/* setUnaryEnergy of image `img` (N_img x L) */
int dcrf_batch_set_unary(dcrf_batch_t b, int img, const float* unary, int on_device);
/* addPairwiseEnergy(features[img], PottsCompatibility(w), DIAG_KERNEL, ntype) for every image */
int dcrf_batch_add_pairwise(dcrf_batch_t b, const float* const* features, int d, float potts_w, int ntype,
                            int on_device);
/* addPairwiseGaussian / addPairwiseBilateral for every image (images: host RGB per image) */
int dcrf_batch_add_pairwise_gaussian(dcrf_batch_t b, float sx, float sy, float potts_w);
int dcrf_batch_add_pairwise_bilateral(dcrf_batch_t b, float sx, float sy, float sr, float sg, float sb,
                                      const unsigned char* const* images, float potts_w);
/* unary_init of every image into the batch field Q */
int dcrf_batch_unary_init(dcrf_batch_t b, float* Q, int on_device);
/* QP of every image (init null: unary_init); iterations[b] (may be null) */
int dcrf_batch_qp(dcrf_batch_t b, const float* init, float* out, int max_iter, int ref_stop, int on_device,
                  int* iterations);
/* LP of every image; sweeps[b * cap + j] = qp-gamma sweeps of inner step j (may be null) */
int dcrf_batch_lp(dcrf_batch_t b, const float* init, float* out, const dcrf_lp_params* params, int on_device,
                  int* sweeps, int cap);
/* assignment_energy_true of every image's labelling (labels: batch pixel order) */
int dcrf_batch_assignment_energy_true(dcrf_batch_t b, const int* labels, int on_device, double* energies);

/* ------------------------------------------------------------------------ */
/* Device memory and kernel timing.  No counterpart in the reference: they  */
/* let a caller without another GPU allocator keep inputs resident in HBM   */
/* (on_device = 1) and time each kernel on the stream it runs on.           */
/* ------------------------------------------------------------------------ */
int dcrf_malloc(void** ptr, size_t bytes);
int dcrf_free(void* ptr);
/* kind: 0 host->device, 1 device->host, 2 device->device */
int dcrf_memcpy(void* dst, const void* src, size_t bytes, int kind);
int dcrf_device_synchronize(void);
/* enable(1) clears the record and times every kernel launch with HIP events;
 * report writes lines "name count total_ms" into buf (NUL-terminated) and
 * returns the length it needs (call with len 0 to size the buffer). */
int dcrf_profile_enable(int on);
/* Time only the launches whose name starts with `prefix` (NULL or "" for
 * all), so that timing one kernel costs the others nothing.  No reference
 * counterpart. */
int dcrf_profile_filter(const char* prefix);
/* Launch-shape autotuner (no reference counterpart): grid-stride kernels whose
 * results do not depend on the grid are timed over a few candidate grids at
 * their first launch for a problem size and keep the fastest.  enable(0)
 * turns it off (also DCRF_AUTOTUNE=0); report writes "kernel@size_class grid
 * ms | candidate:ms ..." lines like dcrf_profile_report. */
int dcrf_autotune_enable(int on);
int dcrf_autotune_report(char* buf, int len);
int dcrf_profile_report(char* buf, int len);

/* ---- file formats of the reference's example pipeline (host only, no GPU) ---------- */
/* load_unary (src/file_storage.cpp:262-291 over ProbImage::decompress and boostToProb,
 * external/ProbImage/probimage.cpp:64-78, 140-165): the .c_unary file as -log
 * probabilities, L x N label-contiguous.  out may be NULL to query W, H, L; cap is the
 * number of floats out holds. */
int dcrf_io_load_unary(const char* path, float* out, size_t cap, int* W, int* H, int* L);
/* load_image (src/file_storage.cpp:197-222): an uncompressed 24 / 32 bpp BMP as RGB,
 * pixel (x, y) at (x + y W) * 3.  out may be NULL to query W and H. */
int dcrf_io_load_image(const char* path, unsigned char* out, size_t cap, int* W, int* H);
/* save_map(Q, size, path, "MSRC") (src/file_storage.cpp:332-380): the argmax labelling
 * in the MSRC colours, PNG (or binary PPM for a ".ppm" path). */
int dcrf_io_save_map(const float* Q, int W, int H, int L, const char* path);

/* The per-region terms of DenseCRF2D::addSuperPixel (src/densecrf.cpp:193-251) for a
 * segmentation given as region ids in [0, R) (the reference computes it by mean-shift):
 * e[r] = constant * exp(-sd_r / normaliser), c[r] = constant, with the reference's grey
 * value read from img[i], img[i+1], img[i+2].  Host only; feed the result to
 * dcrf_crf_add_superpixel_layer. */
int dcrf_superpixel_terms(const unsigned char* img, int N, const int* region, int R, float constant, float normaliser,
                          float* e, float* c);

/* less_confident_pixels(indices, Q, tol), include/densecrf.h:41, src/densecrf.cpp:457-462:
 * the pixels (ascending) whose largest value of Q (L x N) is <= tol, e.g. the pixel set
 * of the restricted ordered filter.  indices must hold N entries; *count receives how
 * many were written. */
int dcrf_less_confident_pixels(const float* Q, int N, int L, float tol, int* indices, int* count, int on_device);
/* No counterpart in the reference.  dcrf_less_confident_pixels for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc): it calls no hipMalloc / hipFree and no allocating library
 * routine.  Q (L x N) and indices (N entries) must be device pointers (host pointers are
 * refused); *count is a host int.  Its only working memory is a fixed device array of
 * 128 KiB per GPU, held until the process ends, so N is at most 16,777,216 (larger N is
 * refused with DCRF_INVALID_ARGUMENT).  Every call on one GPU shares that array: calls
 * that overlap on one GPU race.  It runs on the null stream and waits for it; Q must be
 * complete before the call (the engine's own entry points return with their results
 * complete). */
int dcrf_less_confident_pixels_noalloc(const float* Q, int N, int L, float tol, int* indices, int* count);

#ifdef __cplusplus
}
#endif

#endif /* DENSECRF_HO_AMD_H */
