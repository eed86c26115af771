// Internal types of libuw12g shared by its host translation units: the
// context, integral and projector objects behind the C ABI handles, and the
// launch helpers of the evaluation.
#pragma once

#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>

#include <chrono>
#include <cstdio>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "../../include/uw12g.h"
#include "kernels.h"
#include "uw12g_internal.h"

namespace uw12g {

// Device buffer owned by RAII.
struct DBuf {
  double* p = nullptr;
  size_t n = 0;
  DBuf() = default;
  DBuf(const DBuf&) = delete;
  DBuf& operator=(const DBuf&) = delete;
  ~DBuf() { reset(); }
  void reset() {
    if (p) {
      unregister_alloc(p);
      (void)hipFree(p);
    }
    p = nullptr;
    n = 0;
  }
  double* get(size_t count) {
    if (count > n || !p) {
      reset();
      if (hipMalloc(&p, std::max<size_t>(count, 1) * sizeof(double)) != hipSuccess)
        throw Error(-5, "out of device memory (" + std::to_string(count * 8.0 / 1e9) + " GB requested)");
      n = std::max<size_t>(count, 1);
      register_alloc(p, n * sizeof(double));
    }
    return p;
  }
};

}  // namespace uw12g

using uw12g::DBuf;

struct uw12g_context {
  int device = 0;
  // doubles allowed for one blocked intermediate of the indirect terms (K~,
  // Q~, GW/GV blocks): 1/16 of the device memory, at most 16 GiB
  size_t block_cap = size_t(1) << 28;
  size_t block_cap_default = size_t(1) << 28;
  int prec = 0;  // arithmetic of the bulk contractions during an evaluation (GemmDesc::prec)
  std::vector<std::string> exact_extra;  // stage names also kept FP64 in the mixed form (tests)
  int verbosity = 0;  // > 0: progress lines on stderr at the major stages of an evaluation
  std::chrono::steady_clock::time_point t_eval;
  size_t stream_cap = size_t(1) << 25;  // doubles per host staging buffer of the per-shell setup
  // Differentiation form: arguments of the last forward call.  Its
  // intermediates live in the named buffers until the next fused or forward
  // evaluation on this context (n_eval counts those).
  struct Tape {
    uint64_t id = 0, n_eval = 0;
    const uw12g_base_integrals *W = nullptr, *V = nullptr, *WV = nullptr;
    const uw12g_abs_projectors* P = nullptr;
    std::vector<double> C, occ;
    std::vector<size_t> n_occ, n_active;
    size_t n_ao = 0, n_orb = 0, n_spin = 0;
    bool indirect = false;
    double k0 = 0.0, k1 = 0.0;
    int prec = 0;
    // true: nothing is kept; the adjoint recomputes every intermediate (a fused
    // evaluation), so the tape does not depend on the context's buffers
    bool recompute = false;
  } tape;
  // Largest set of forward intermediates (doubles) a tape keeps; past it the
  // forward keeps nothing and the adjoint recomputes.  A quarter of the device
  // memory by default.
  size_t tape_cap = 0, tape_cap_default = 0;
  uint64_t n_eval = 0, n_tapes = 0;
  // Integral and projector handles still alive on this context.  Destroying the
  // context while some exist only marks it (doomed); the last handle's destroy
  // then deletes it, so a handle never outlives the stream and buffers it uses.
  int live = 0;
  bool doomed = false;
  hipStream_t stream = nullptr;
  rocblas_handle rb = nullptr;
  uw12g::Workspace gemm_ws;
  std::map<std::string, std::unique_ptr<DBuf>> bufs;
  uw12g::Profiler prof;

  // Named device buffers reused across evaluations; growing or releasing one
  // first drains the stream so no queued kernel still reads the old memory.
  // "blk1".."blk4" are scratch arenas shared by the consecutive blocked stages
  // of the indirect terms (4e, 3e energy, 3e Fock i/j, 3e Fock k): each stage
  // overwrites them in stream order, and keeping them allocated avoids a
  // multi-GiB allocate/free per evaluation.
  double* buf(const std::string& name, size_t n) {
    if (arena) return carve(name, n);
    auto& b = bufs[name];
    if (!b) b.reset(new DBuf());
    if (b->p && n > b->n) UW12G_HIP(hipStreamSynchronize(stream));
    return b->get(n);
  }
  void release(const std::string& name) {
    if (arena) return;
    auto it = bufs.find(name);
    if (it == bufs.end()) return;
    UW12G_HIP(hipStreamSynchronize(stream));
    bufs.erase(it);
  }

  // No-allocation mode (uw12g_form_fock_noalloc): while `arena` is set, named
  // buffers and the GEMM workspace are carved, 256-byte aligned, from the
  // device's static scratch array instead of allocated.  A name asked for again
  // with at most its carved size gets the same region; a larger request gets a
  // new one.  Exhausting the arena throws Error(-2) before the buffer is used.
  double* arena = nullptr;
  size_t arena_cap = 0, arena_used = 0;
  std::map<std::string, std::pair<double*, size_t>> arena_bufs;
  uw12g::Workspace arena_ws;
  uw12g::Workspace* ws_active = &gemm_ws;
  double* carve(const std::string& name, size_t n) {
    n = std::max<size_t>(n, 1);
    auto it = arena_bufs.find(name);
    if (it != arena_bufs.end() && n <= it->second.second) return it->second.first;
    const size_t off = (arena_used + 31) / 32 * 32;
    if (off + n > arena_cap)
      throw uw12g::Error(-2, "no-allocation scratch exhausted: this evaluation needs more than the " +
                                 std::to_string(arena_cap) + " doubles of the static scratch array (buffer '" +
                                 name + "' asks for " + std::to_string(n) + " more at offset " +
                                 std::to_string(off) + ")");
    arena_used = off + n;
    arena_bufs[name] = {arena + off, n};
    return arena + off;
  }
  ~uw12g_context() {
    prof.clear();
    bufs.clear();
    if (rb) rocblas_destroy_handle(rb);
    if (stream) (void)hipStreamDestroy(stream);
  }
};

struct uw12g_base_integrals {
  uw12g_context* ctx = nullptr;
  int N = 0, Nd = 0, Nr = 0, M = 0, D = 0;
  DBuf gamma;  // M N D
  DBuf sign;   // D
  std::vector<double> sign_host;
  // Gamma rotated onto the kept eigenvectors of one ABS overlap: U^T Gamma
  // (r N D, built on the first evaluation that pairs these integrals with the
  // projector whose id is gt_id, replaced when another projector comes)
  mutable DBuf gt;
  mutable uint64_t gt_id = 0;
  // gt as the A operand of the t~ dX product Xi, kept across evaluations for the gt of projector
  // id gtp_id / gtr_id: split into plane tensors (split_a_planes) for GEMM precision gtp_prec, or,
  // when those would not fit, the largest magnitude of each of its rows (the FP16 row scales' bound)
  mutable DBuf gtp;
  mutable uint64_t gtp_id = 0;
  mutable int gtp_prec = -1;
  mutable DBuf gtr;
  mutable uint64_t gtr_id = 0;
  // Non-stored integrals (uw12g_setup_base_integrals_gen with stored == 0):
  // no gamma; the DF metric's kept eigenvectors T[B, A] = Q[B, A] / sqrt|lambda_A|
  // (Nd x D) and the device generator, called again by every evaluation.
  bool stored = true;
  DBuf T;
  uw12g_three_index_gen_fn gen = nullptr;
  void* gen_user = nullptr;
};

struct uw12g_abs_projectors {
  uw12g_context* ctx = nullptr;
  int N = 0, Nr = 0, M = 0;
  int r = 0;       // kept eigenvectors of S
  uint64_t id = 0;  // distinct for every projector created in the process
  DBuf Sp;    // M x M, S+ = U diag(dinv) U^T
  DBuf U;     // M x r kept eigenvectors
  DBuf dinv;  // r: 1 / eigenvalue
};

namespace uw12g {

inline void require(bool ok, const std::string& msg, int code = -1) {
  if (!ok) throw Error(code, msg);
}

inline void set_device(uw12g_context* ctx) { UW12G_HIP(hipSetDevice(ctx->device)); }

// GEMM launch in the context's current contraction precision (see uw12g.cpp).
void run_gemm(uw12g_context* ctx, const GemmDesc& d, const char* stage, bool exact = false);
// The GEMM arithmetic run_gemm uses for a stage (GemmDesc::prec).
int stage_prec(uw12g_context* ctx, const char* stage, bool exact = false);
// C^T = B^T A^T: the same product with rows and columns (and operands) exchanged.
GemmDesc transposed(const GemmDesc& d);

template <class F>
void timed(uw12g_context* ctx, const char* stage, const char* kernel, double flops, double bytes, F&& f) {
  if (!ctx->prof.on) {
    f();
    return;
  }
  ProfileRec rec;
  ctx->prof.stage = stage;
  ctx->prof.begin(rec, ctx->stream);
  f();
  ctx->prof.end(rec, ctx->stream, kernel, flops, bytes);
}

inline double spin_factor(int ns, int s, int t, double k0, double k1) {  // four_electron_utils.cpp:9-30
  if (ns == 1) return 2.0 * (k0 + k1);
  return s == t ? k1 : k0;
}

// A progress line on stderr once the stream has finished the work queued so
// far (verbosity > 0 only: the wait serialises the host with the GPU).
template <class... Args>
void progress(uw12g_context* ctx, const char* fmt, Args... args) {
  if (ctx->verbosity <= 0) return;
  UW12G_HIP(hipStreamSynchronize(ctx->stream));
  const double t = std::chrono::duration<double>(std::chrono::steady_clock::now() - ctx->t_eval).count();
  size_t free_b = 0, total_b = 0;
  (void)hipMemGetInfo(&free_b, &total_b);
  std::fprintf(stderr, "[uw12g] %8.2fs %6.1f GB free  ", t, free_b / 1e9);
  if constexpr (sizeof...(Args) == 0) std::fputs(fmt, stderr);
  else std::fprintf(stderr, fmt, args...);
  std::fprintf(stderr, "\n");
  std::fflush(stderr);
}

// Per-spin operands of one evaluation (device pointers into the context's buffers).
struct SpinData {
  int o = 0, a = 0, c = 0;
  double* Co = nullptr;  // N x o (device)
  double* wk = nullptr;  // o   (1 + [k active]) when a > 0 else 1
  double* phiW = nullptr;
  double* phiV = nullptr;  // N o D: AO rows of Gamma Co
  double* psiW = nullptr;
  double* psiV = nullptr;  // Mr o D: (U^T Gamma) Co, the ABS rows rotated onto S's kept eigenvectors
  double* o2W = nullptr;
  double* o2V = nullptr;  // a o D
  double* phitV = nullptr;  // Mr a D_V: diag(1/s) Psi_V[:, act]
  double* t = nullptr;      // D_W x D_V
  double* x = nullptr;      // D_W x D_V
};

// Evaluation over non-stored W / V integrals (uw12g_stream.cpp): the
// opposite-spin terms, with the three-index blocks requested again from the
// integrals' device generators.  Adds the energy terms into Eacc and the
// unsymmetrised Fock matrices into Fraw (null: energy only).
void evaluate_streamed(uw12g_context* ctx, const uw12g_base_integrals* W, const uw12g_base_integrals* V,
                       const uw12g_base_integrals* WV, const uw12g_abs_projectors* P, const std::vector<SpinData>& sd,
                       double k0, double k1, double* Fraw, double* Eacc, double* partial);

// U^T Gamma for the projector P, built once per (integrals, projector) pair (see uw12g.cpp).
void ensure_rotated(uw12g_context* ctx, const uw12g_base_integrals* X, const uw12g_abs_projectors* P);

// What one evaluate() call does: the fused energy (+ optional Fock) pass, the
// forward half of the differentiation form (energy only; the W/V transforms
// Phi, O2, t, S+ Phi_V and X stay in the context's named buffers), or its
// adjoint half (Fock only, reading those buffers instead of recomputing them
// and skipping the energy-only contractions of the indirect terms).
enum EvalMode { kFused = 0, kForward = 1, kAdjoint = 2 };

}  // namespace uw12g
