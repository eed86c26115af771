/* uw12g_testing.h -- test hooks of libuw12g (no reference counterpart).
 *
 * Exposes the strided FP64 GEMM building block so that its MFMA kernel can be
 * checked in isolation against a one-thread-per-output kernel and the host.
 * Every index of every operand uses the two-level map
 *   off(i) = (i mod n0) s0 + (i div n0) s1   (n0 == 0: off(i) = i s0).
 */
#ifndef UW12G_TESTING_H
#define UW12G_TESTING_H

#include <stddef.h>

#include "uw12g.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  long long s0, s1;
  unsigned n0;
} uw12g_map2;

/* C <- alpha sum_k A(m,k) ks(k) B(k,n) + beta C for every batch, on device
 * `device`; host arrays are copied in and C is copied back.  mode 0 runs the
 * FP64 MFMA kernel, 1 the one-thread-per-output FP64 kernel, 2 the FP32 MFMA
 * form (operands rounded to FP32, FP32 sums per k chunk), 3 / 4 the BF16x3 /
 * BF16x6 split forms, 5 the FP16x3 split form; ks may be NULL.  With the split
 * forms, adding 8 / 16 / 32 to the mode keeps A / B / C in FP32 storage on the
 * device (the host values rounded to float; C is returned widened), and adding
 * 64 splits A once into a kept plane buffer that the product then reads (the
 * path of a product whose A recurs). */
int uw12g_test_gemm(int device, int M, int N, int K, int batch, double alpha, double beta, const double* A,
                    size_t nA, uw12g_map2 Am, uw12g_map2 Ak, long long Ab, const double* B, size_t nB,
                    uw12g_map2 Bk, uw12g_map2 Bn, long long Bb, double* C, size_t nC, uw12g_map2 Cm,
                    uw12g_map2 Cn, long long Cb, const double* ks, size_t nks, uw12g_map2 Km, int mode);

/* Average time in milliseconds of `reps` launches of the same GEMM on arrays
 * already in device memory (after one untimed launch), on a stream of its own;
 * prec 0 runs the FP64 matrix cores, 1 the FP32 form, 2 / 3 the BF16x3 /
 * BF16x6 split forms, 4 the FP16x3 split form; + 8 / 16 / 32: A / B / C
 * hold FP32 values (split forms).  nA, nB, nC and nks are the element counts
 * of the arrays: a descriptor that would reach past one of them is refused
 * (-2) before anything is launched. */
int uw12g_time_gemm(int device, int M, int N, int K, int batch, double alpha, double beta, const double* A,
                    size_t nA, uw12g_map2 Am, uw12g_map2 Ak, long long Ab, const double* B, size_t nB,
                    uw12g_map2 Bk, uw12g_map2 Bn, long long Bb, double* C, size_t nC, uw12g_map2 Cm,
                    uw12g_map2 Cn, long long Cb, const double* ks, size_t nks, uw12g_map2 Km, int prec, int reps,
                    double* ms);

/* Host staging size (doubles per buffer) of uw12g_setup_base_integrals_fn on
 * this context; small values force many shell groups (tests). */
int uw12g_test_set_stream_cap(uw12g_context* ctx, size_t doubles);

/* Doubles allowed for one blocked intermediate of the same-spin terms on this
 * context (K~, X_i / Y_i, GW / GV blocks; L and Xi use twice / half of it);
 * small values force many blocks (tests).  0 restores the default. */
int uw12g_test_set_block_cap(uw12g_context* ctx, size_t doubles);

/* Largest set of forward intermediates (doubles) a differentiation-form tape
 * keeps on this context; past it uw12g_forward keeps nothing and
 * uw12g_adjoint recomputes them (small values force that path; tests).  0
 * restores the default, a quarter of the device memory. */
int uw12g_test_set_tape_cap(uw12g_context* ctx, size_t doubles);

/* Comma-separated stage names whose contractions the mixed form also runs in
 * FP64 on this context (error attribution); NULL or "" clears. */
int uw12g_test_set_exact_stages(uw12g_context* ctx, const char* prefixes);

/* Force the GEMM tile shape (64 or 128 rows x 64 or 128 columns) for every
 * later launch in the process; 256, 256 forces the 256 x 256 tiles of the
 * two-plane split forms (BF16x3, FP16x3) and leaves the other forms' choice
 * automatic; 0, 0 restores the automatic choice (tests and tile-choice
 * measurements). */
int uw12g_test_force_tile(int bm, int bn);

/* BF16 split forms: on != 0 (the default) splits the operands of each product
 * once into bfloat16 plane tensors in the workspace; 0 splits them inside the
 * GEMM kernel, for every later launch in the process (tests of both paths).
 * max_bytes > 0 caps the plane workspace of one product, so that larger
 * products run in pieces; 0 restores the default cap (16 GiB). */
int uw12g_test_set_bf16_planes(int on, long long max_bytes);

/* What the split forms keep of the rotated V integrals for the t~ dX product,
 * for every later evaluation in the process: 1 (the default) its plane tensors,
 * or its row maxima when those do not fit; 2 its row maxima only; 0 nothing
 * (every product splits it again).  All give the same results (tests). */
int uw12g_test_set_kept_gt(int mode);

/* Host evaluation of the two-level map (checks the fast division). */
long long uw12g_test_map2(uw12g_map2 m, unsigned i);

#ifdef __cplusplus
}
#endif

#endif /* UW12G_TESTING_H */
