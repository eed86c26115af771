/* uw12g.h -- C ABI of the MI355X (gfx950) UW12 energy + Fock library.
 *
 * Drop-in for the libuw12 public API (zwill22/libuw12, libuw12.hpp:12-171):
 * the same inputs (FP64, host, column-major, the reference's packing), the same
 * validation, the same results; the work runs in HIP on one GPU.  Every entry
 * point returns 0 on success or a negative status, never throws, and leaves a
 * thread-local message for uw12g_last_error().
 *
 * Status codes: -1 invalid argument (the reference throws std::runtime_error /
 * std::logic_error for the same inputs), -2 unsupported size, -3 HIP / device
 * error, -4 numerical failure (eigensolver), -5 out of device memory.
 *
 * Layouts (as in the reference, libuw12.cpp:12-123):
 *   X3     n_ao(n_ao+1)/2 x n_df, row = mu(mu+1)/2 + nu for nu <= mu
 *   X2     n_df x n_df (symmetric to 1e-10)
 *   X3_ri  (n_ao n_ri) x n_df, row = r + n_ri mu
 *   S      (n_ao+n_ri)^2 overlap of the AO then RI functions
 *   C      n_ao x (n_orb n_spin), alpha columns then beta; first n_occ are occupied
 *   occ    n_occ[0] alpha values then n_occ[1] beta values, all > 0
 *   fock   n_ao x (n_ao n_spin), alpha block then beta
 */
#ifndef UW12G_H
#define UW12G_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct uw12g_context uw12g_context;
typedef struct uw12g_base_integrals uw12g_base_integrals;
typedef struct uw12g_abs_projectors uw12g_abs_projectors;

/* Arithmetic of the bulk contractions.  FP64 is the plain form.  UW12G_FP32 runs
 * every contraction of the evaluation on FP32 matrix cores (operands rounded to
 * FP32, FP32 sums over k ranges of at most 4096, ranges combined in FP64);
 * storage, the metric eigensolves and the energy reductions stay FP64.
 * UW12G_MIXED runs the large intermediate tensors of the Fock terms that way
 * and keeps in FP64 the contractions whose results enter the energies directly
 * (orbital transforms, t, X, the energy-only products) and those that
 * accumulate into the Fock matrix.  UW12G_MIXED_BF16X3 and UW12G_MIXED_BF16X6
 * are the mixed form with those intermediates as BF16 split products on the
 * BF16 matrix cores (FP32 sums as above): each FP64 operand split into two
 * bfloat16 planes and three plane products (about 3 * 2^-16 relative error
 * per product), or into three planes and six products (about 2^-22).
 * UW12G_MIXED_FP16X3 splits each operand row, scaled by a power of two, into
 * two IEEE half planes and runs three products on the FP16 matrix cores
 * (about 3 * 2^-22 per product at the cost of BF16x3). */
typedef enum {
  UW12G_FP64 = 0,
  UW12G_FP32 = 1,
  UW12G_MIXED = 2,
  UW12G_MIXED_BF16X3 = 3,
  UW12G_MIXED_BF16X6 = 4,
  UW12G_MIXED_FP16X3 = 5
} uw12g_precision;

/* Term order of terms_out[] */
enum {
  UW12G_E2_DIRECT = 0,
  UW12G_E2_INDIRECT = 1,
  UW12G_E3_DIRECT = 2,
  UW12G_E3_INDIRECT = 3,
  UW12G_E4_DIRECT = 4,
  UW12G_E4_INDIRECT = 5,
  UW12G_N_TERMS = 6
};

/* Thread-local description of the last failure on this thread. */
const char* uw12g_last_error(void);

/* One HIP device + stream + device caches.  No reference counterpart: the
 * reference keeps its state inside BaseIntegrals / Integrals objects. */
int uw12g_context_create(int device, uw12g_context** out);
void uw12g_context_destroy(uw12g_context* ctx);

/* Replaces uw12::setup_base_integrals(X3, X2, X3_ri, n_ao, n_df, n_ri, copy)
 * (libuw12.hpp:42-50, libuw12.cpp:22-36) and, with X3_ri == NULL and
 * n_ri == 0, setup_base_integrals(X3, X2, n_ao, n_df, copy) (libuw12.hpp:64-70).
 * Host arrays are copied to the device; the DF metric X2 is eigendecomposed in
 * FP64 at creation (base_integrals.cpp:155-166). */
int uw12g_setup_base_integrals(uw12g_context* ctx, const double* X3, const double* X2, const double* X3_ri,
                               size_t n_ao, size_t n_df, size_t n_ri, uw12g_base_integrals** out);
/* Same, for arrays that already live in device memory of the context's GPU. */
int uw12g_setup_base_integrals_device(uw12g_context* ctx, const double* X3_dev, const double* X2,
                                      const double* X3_ri_dev, size_t n_ao, size_t n_df, size_t n_ri,
                                      uw12g_base_integrals** out);

/* Three-index integrals of one DF shell, for uw12g_setup_base_integrals_fn:
 * write the column-major n_rows x df_sizes[shell] block of shell `shell` to
 * `out` (n_rows = n_ao (n_ao + 1) / 2 packed AO pairs, row mu (mu + 1) / 2 + nu;
 * or n_ao n_ri, row r + n_ri mu, for the RI callback) and return 0. */
typedef int (*uw12g_three_index_fn)(void* user, size_t shell, double* out);

/* Replaces the reference's function-based BaseIntegrals constructors
 * (base_integrals.hpp:52-72 with integral_functions.hpp:16-55): the
 * three-index integrals are requested one DF shell at a time and streamed to
 * the device in groups of whole shells, so the caller never holds X3 / X3_ri
 * whole.  X2 is the two-index matrix (the TwoIndexFn result); three_index_ri
 * NULL with n_ri = 0 gives the no-RI form (:64-72).  A callback returning
 * nonzero fails the call with -6. */
int uw12g_setup_base_integrals_fn(uw12g_context* ctx, const double* X2, uw12g_three_index_fn three_index,
                                  uw12g_three_index_fn three_index_ri, void* user, const size_t* df_sizes,
                                  size_t n_shells, size_t n_ao, size_t n_df, size_t n_ri, uw12g_base_integrals** out);

/* Three-index integrals of the DF functions [df0, df0 + n), produced on the
 * device, for uw12g_setup_base_integrals_gen: enqueue on `stream` (the
 * context's hipStream_t) the writes of the column-major block
 *   out[mu' + m (nu + n_ao j)] = (mu' nu | x | df0 + j),  nu < n_ao, j < n,
 * with m = n_ao (with_ri == 0: the AO functions mu' < n_ao) or m = n_ao + n_ri
 * (with_ri != 0: the AO rows, then row n_ao + r = (r nu | x | df0 + j) of RI
 * function r).  `out` is device memory of the context's GPU; return 0, or
 * nonzero to fail the calling entry point with -6. */
typedef int (*uw12g_three_index_gen_fn)(void* user, size_t df0, size_t n, int with_ri, double* out, void* stream);

/* Replaces the function-based BaseIntegrals constructor with its storage flags
 * (base_integrals.hpp:52-62, store_ao_integrals / store_ri_integrals; blocks
 * as in integral_functions.hpp:16-55) for integrals produced on the device by
 * `gen`, with X2 the two-index matrix.  stored != 0: the library builds and
 * keeps its three-index tensor, as the other setup calls do.  stored == 0:
 * only the DF metric's eigenvectors are kept, and every evaluation requests
 * again the blocks it needs (the reference's non-storing, direct mode,
 * direct_utils.cpp:349-436), so no three-index tensor of the size of X3_ri is
 * ever resident.  Evaluations with non-stored W / V integrals compute the
 * opposite-spin terms (indirect_term == 0 or scale_same_spin == 0; -2
 * otherwise) and take UW12G_FP64 or one of the mixed forms. */
int uw12g_setup_base_integrals_gen(uw12g_context* ctx, const double* X2, uw12g_three_index_gen_fn gen, void* user,
                                   size_t n_ao, size_t n_df, size_t n_ri, int stored, uw12g_base_integrals** out);

void uw12g_base_integrals_destroy(uw12g_base_integrals* ints);
/* Number of DF metric eigenvectors kept (|lambda| >= 1e-10), or negative status. */
long uw12g_base_integrals_kept(const uw12g_base_integrals* ints);

/* Replaces uw12::setup_abs_projectors(S, n_ao, n_ri) (libuw12.hpp:80,
 * ri_utils.hpp:45-74): eigenvalues |s| <= max(1e-6, 1e-8 max|s|) are dropped. */
int uw12g_setup_abs_projectors(uw12g_context* ctx, const double* S, size_t n_ao, size_t n_ri,
                               uw12g_abs_projectors** out);
void uw12g_abs_projectors_destroy(uw12g_abs_projectors* abs);
/* Copy the (n_ao+n_ri)^2 projector S^+ (AO first) to host memory. */
int uw12g_abs_projectors_get(const uw12g_abs_projectors* abs, double* out);

/* Replaces uw12::uw12_energy (libuw12.hpp:129-140, libuw12.cpp:156-182).
 * orbitals/occupations are the arrays that setup_orbitals (libuw12.hpp:90-92)
 * and setup_occupations (libuw12.hpp:101,111-113) wrap. */
int uw12g_energy(uw12g_context* ctx, const uw12g_base_integrals* W, const uw12g_base_integrals* V,
                 const uw12g_base_integrals* WV, const uw12g_abs_projectors* abs, const double* C, size_t n_ao,
                 size_t n_orb, size_t n_spin, const double* occ, const size_t* n_occ, const size_t* n_active,
                 double scale_opp_spin, double scale_same_spin, double* energy);

/* Replaces uw12::uw12_fock(double* fock, ...) (libuw12.hpp:157-169,
 * libuw12.cpp:184-214). */
int uw12g_fock(uw12g_context* ctx, double* fock, const uw12g_base_integrals* W, const uw12g_base_integrals* V,
               const uw12g_base_integrals* WV, const uw12g_abs_projectors* abs, const double* C, size_t n_ao,
               size_t n_orb, size_t n_spin, const double* occ, const size_t* n_occ, const size_t* n_active,
               double scale_opp_spin, double scale_same_spin, double* energy);

/* Replaces uw12::form_fock (src/uw12.hpp:129-142, src/uw12.cpp:22-119): explicit
 * indirect_term flag, optional Fock (fock == NULL: energy only) and the six
 * energy pieces the reference prints (src/uw12.cpp:67-69,86-88,102-104). */
int uw12g_form_fock(uw12g_context* ctx, const uw12g_base_integrals* W, const uw12g_base_integrals* V,
                    const uw12g_base_integrals* WV, const uw12g_abs_projectors* abs, const double* C, size_t n_ao,
                    size_t n_orb, size_t n_spin, const double* occ, const size_t* n_occ, const size_t* n_active,
                    int indirect_term, double scale_opp_spin, double scale_same_spin, uw12g_precision precision,
                    double* fock, double* terms_out, double* energy);

/* Size in doubles of the static device scratch array of uw12g_form_fock_noalloc
 * (128 MiB per GPU the library runs on, held until the process ends). */
#define UW12G_NOALLOC_SCRATCH_DOUBLES ((size_t)1 << 24)

/* uw12g_form_fock without device allocation.  No counterpart in the reference:
 * it is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments, results and errors as uw12g_form_fock with stored integrals; every
 * device buffer of the evaluation -- intermediates, split-K partial sums, and
 * U^T Gamma of W / V when an earlier call has not kept it on the integrals --
 * is carved from one static device array of UW12G_NOALLOC_SCRATCH_DOUBLES
 * doubles per GPU, and the call makes no hipMalloc / hipFree.  An evaluation
 * whose working set exceeds the array fails with -2 (the message names the
 * size it needed) at the first buffer that does not fit, before that buffer is
 * used.  All calls on one GPU share its array: calls that overlap on one GPU
 * (from several threads or contexts) race.  Non-stored integrals: -2. */
int uw12g_form_fock_noalloc(uw12g_context* ctx, const uw12g_base_integrals* W, const uw12g_base_integrals* V,
                            const uw12g_base_integrals* WV, const uw12g_abs_projectors* abs, const double* C,
                            size_t n_ao, size_t n_orb, size_t n_spin, const double* occ, const size_t* n_occ,
                            const size_t* n_active, int indirect_term, double scale_opp_spin, double scale_same_spin,
                            uw12g_precision precision, double* fock, double* terms_out, double* energy);

/* Differentiation form, forward half: the energy and its six pieces as
 * uw12g_form_fock (fock == NULL), keeping the forward intermediates (the W/V
 * orbital transforms, (P|ij), t_AB, S+ Phi_V, X_AB) on the context.  *tape
 * identifies them for uw12g_adjoint.  No reference counterpart: the reference
 * computes energy and Fock in one pass (src/uw12.cpp:22-119). */
int uw12g_forward(uw12g_context* ctx, const uw12g_base_integrals* W, const uw12g_base_integrals* V,
                  const uw12g_base_integrals* WV, const uw12g_abs_projectors* abs, const double* C, size_t n_ao,
                  size_t n_orb, size_t n_spin, const double* occ, const size_t* n_occ, const size_t* n_active,
                  int indirect_term, double scale_opp_spin, double scale_same_spin, uw12g_precision precision,
                  double* terms_out, double* energy, unsigned long long* tape);

/* Differentiation form, adjoint half: F = dE/dD (n_ao x n_ao per spin, alpha
 * then beta, symmetric) of the forward call `tape`, from its kept
 * intermediates; the result equals uw12g_form_fock's Fock on the same inputs
 * (all-electron: the exact derivative; frozen core: the reference's averaged
 * form, four_electron.cpp:63-76).  When the intermediates would exceed a
 * quarter of the device memory, or the integrals are non-stored, the forward
 * keeps none and the adjoint recomputes them (regenerating the three-index
 * blocks for non-stored integrals); such a tape stays valid across later
 * evaluations.  Otherwise returns -4 when a later fused or forward evaluation
 * on the context has replaced the intermediates.  Destroying an integral or
 * projector handle of the forward call invalidates its tape (-4). */
int uw12g_adjoint(uw12g_context* ctx, unsigned long long tape, double* fock);

/* Progress lines on stderr at the major stages of every evaluation on this
 * context when level > 0; each line first waits for the GPU work queued before
 * it.  Counterpart of the reference's print_level argument of uw12_energy /
 * uw12_fock (libuw12.hpp:129-140), which prints the energy pieces. */
int uw12g_context_set_verbosity(uw12g_context* ctx, int level);

/* Kernel-level timing of the most recent evaluation on this context:
 * names[i] / milliseconds[i] for up to cap entries; returns the count.
 * Enabled by uw12g_context_set_profiling(ctx, 1).  No reference counterpart. */
int uw12g_context_set_profiling(uw12g_context* ctx, int on);
int uw12g_context_profile(const uw12g_context* ctx, const char** names, double* milliseconds,
                          double* flops, double* bytes, int cap);

#ifdef __cplusplus
}
#endif

#endif /* UW12G_H */
