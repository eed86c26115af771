// libuw12g: C ABI + host orchestration of the UW12 energy / Fock evaluation on
// one MI355X.  See include/uw12g.h for the contract.
//
// Data layout in HBM (per base-integrals object X in {W, V, WV}):
//   Gamma_X[mu' + M alpha + M N A]  (M = n_ao + n_ri, N = n_ao, A < D_X)
//     = sum_B G0_X[mu', alpha, B] Q_X[B, A] / sqrt|lambda_A|,  s_X[A] = sign(lambda_A)
//   with G0[mu' < N] the unpacked AO block of X3 and G0[N + r] = X3_ri[r + n_ri alpha].
//   Every DF-metric contraction of the reference, sum_A (..|A~) lambda_A (A~|..),
//   equals sum_A (..|Gamma_A) s_A (Gamma_A|..) for any factorisation of X2
//   (base_integrals.cpp:155-166 uses P2 = Q pinv(Lambda)); the balanced form keeps
//   |Gamma| bounded where 1/lambda would not be.
// Per evaluation and spin s (o occupied, a active, c = o - a core):
//   Phi_X[mu', k, A]   = sum_alpha Gamma_X[mu', alpha, A] Co[alpha, k]    (mu' k | X | A)
//   O2_X[i, k, A]      = sum_mu Ca[mu, i] Phi_X[mu, k, A]                  (i k | X | A)
//   PhiT_V[mu', i, B]  = sum_nu' S+[mu', nu'] Phi_V[nu', c+i, B]
// and every energy / Fock term is a strided GEMM over these (gemm_f64.hip).
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <atomic>
#include <cfloat>
#include <chrono>
#include <cstdio>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/uw12g.h"
#include "kernels.h"
#include "uw12g_internal.h"
#include "uw12g_impl.h"

namespace uw12g {

static thread_local std::string g_last_error;

void set_last_error(const std::string& msg) { g_last_error = msg; }

// Static device scratch of uw12g_form_fock_noalloc: each GPU the library runs
// on holds its own copy until the process ends.  Host code never uses the
// array's name as an address: scratch_address() returns the device address
// hipGetSymbolAddress gave for the device, taken once per device when the
// first context on it is created.
static __device__ __align__(256) double scratch[UW12G_NOALLOC_SCRATCH_DOUBLES];

namespace {
std::mutex& scratch_mu() {
  static std::mutex* mu = new std::mutex();
  return *mu;
}
std::map<int, double*>& scratch_ptrs() {
  static std::map<int, double*>* m = new std::map<int, double*>();
  return *m;
}
}  // namespace

// With `device` current: look up (once per device) the address of its scratch array.
void init_scratch(int device) {
  std::lock_guard<std::mutex> lk(scratch_mu());
  if (scratch_ptrs().count(device)) return;
  double* p = nullptr;
  UW12G_HIP(hipGetSymbolAddress(reinterpret_cast<void**>(&p), HIP_SYMBOL(scratch)));
  if (!p) throw Error(-3, "hipGetSymbolAddress returned no address for the scratch array");
  register_alloc(p, sizeof(double) * UW12G_NOALLOC_SCRATCH_DOUBLES);
  scratch_ptrs()[device] = p;
}

double* scratch_address(int device) {
  std::lock_guard<std::mutex> lk(scratch_mu());
  auto it = scratch_ptrs().find(device);
  return it == scratch_ptrs().end() ? nullptr : it->second;
}

namespace {
// Never destroyed: contexts held in static storage of a caller are destroyed
// at process exit in an order this library does not control, and their buffers
// unregister here.
std::mutex& alloc_mu() {
  static std::mutex* mu = new std::mutex();
  return *mu;
}
std::map<uintptr_t, size_t>& allocs() {
  static std::map<uintptr_t, size_t>* m = new std::map<uintptr_t, size_t>();
  return *m;
}
}  // namespace

void register_alloc(const void* p, size_t bytes) {
  std::lock_guard<std::mutex> lk(alloc_mu());
  allocs()[reinterpret_cast<uintptr_t>(p)] = bytes;
}

void unregister_alloc(const void* p) {
  std::lock_guard<std::mutex> lk(alloc_mu());
  allocs().erase(reinterpret_cast<uintptr_t>(p));
}

void check_extent(const void* base, int64_t max_off, const char* what, int elem) {
  if (!base) throw Error(-2, std::string(what) + ": null operand");
  if (max_off < 0) throw Error(-2, std::string(what) + ": negative extent");
  std::lock_guard<std::mutex> lk(alloc_mu());
  const uintptr_t b = reinterpret_cast<uintptr_t>(base);
  auto& m = allocs();
  auto it = m.upper_bound(b);
  if (it == m.begin()) return;  // not library-owned
  --it;
  if (b >= it->first + it->second) return;  // not inside a library allocation
  const uintptr_t end = b + uintptr_t(max_off + 1) * uintptr_t(elem);
  if (end > it->first + it->second)
    throw Error(-2, std::string(what) + ": operand extent exceeds its allocation (" +
                        std::to_string(end - it->first) + " > " + std::to_string(it->second) + " bytes)");
}

double* Workspace::get(size_t n, hipStream_t stream) {
  if (carve) return carve(owner, n);
  const size_t need = std::max<size_t>(n, 1) * sizeof(double);
  if (need > bytes) {
    if (ptr) {
      UW12G_HIP(hipStreamSynchronize(stream));
      unregister_alloc(ptr);
      (void)hipFree(ptr);
    }
    ptr = nullptr;
    bytes = 0;
    if (hipMalloc(&ptr, need) != hipSuccess) throw Error(-5, "out of device memory (workspace)");
    register_alloc(ptr, need);
    bytes = need;
  }
  return static_cast<double*>(ptr);
}

Workspace::~Workspace() {
  if (ptr) {
    unregister_alloc(ptr);
    (void)hipFree(ptr);
  }
}


void Profiler::begin(ProfileRec& r, hipStream_t s) {
  UW12G_HIP(hipEventCreate(&r.start));
  UW12G_HIP(hipEventCreate(&r.stop));
  UW12G_HIP(hipEventRecord(r.start, s));
}

void Profiler::end(ProfileRec& r, hipStream_t s, const std::string& kernel, double flops, double bytes) {
  UW12G_HIP(hipEventRecord(r.stop, s));
  r.name = stage + "|" + kernel;
  r.flops = flops;
  r.bytes = bytes;
  recs.push_back(r);
}

void Profiler::clear() {
  for (auto& r : recs) {
    if (r.start) (void)hipEventDestroy(r.start);
    if (r.stop) (void)hipEventDestroy(r.stop);
  }
  recs.clear();
}

}  // namespace uw12g

using namespace uw12g;


namespace {
// A handle created on ctx: counted so that the context outlives it.
void adopt_child(uw12g_context* ctx) { ++ctx->live; }
// A handle of ctx destroyed: delete the context if it was destroyed meanwhile.
void release_child(uw12g_context* ctx) {
  if (--ctx->live == 0 && ctx->doomed) {
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    delete ctx;
  }
}
}  // namespace

namespace uw12g {

// GEMM launch (timed per launch when profiling is on) in the context's
// current contraction precision.
// `exact` contractions stay FP64 in the mixed form: those whose results feed
// the energies directly (orbital transforms, t, X, the energy-only products)
// and those that accumulate into the Fock matrix (long sums of terms that
// cancel).  The mixed form runs the rest -- the large intermediate tensors of
// the Fock terms (K~, L, R, Xi, G, X_i / Y_i) -- on FP32 matrix cores.
int stage_prec(uw12g_context* ctx, const char* stage, bool exact) {
  const bool mixed = ctx->prec >= UW12G_MIXED;
  if (mixed && !exact)
    for (const std::string& x : ctx->exact_extra)
      if (x == stage) exact = true;
  // GEMM arithmetic of the reduced contractions: 1 FP32, 2 / 3 BF16 with two / three planes, 4 FP16 planes
  const int reduced = ctx->prec == UW12G_MIXED_BF16X3   ? 2
                      : ctx->prec == UW12G_MIXED_BF16X6 ? 3
                      : ctx->prec == UW12G_MIXED_FP16X3 ? 4
                                                        : 1;
  return mixed ? (exact ? 0 : reduced) : ctx->prec;
}

void run_gemm(uw12g_context* ctx, const GemmDesc& d, const char* stage, bool exact) {
  if (d.M <= 0 || d.N <= 0 || d.batch <= 0) return;
  ctx->prof.stage = stage;
  GemmDesc e = d;
  e.prec = stage_prec(ctx, stage, exact);
  if (e.prec < 2) e.planesA = nullptr;
  gemm_f64(e, ctx->stream, ctx->ws_active, &ctx->prof);
}

// Product d of `stage` whose A operand is X's rotated Gamma, in a form over plane tensors: A's
// planes are made on first use and kept on X across evaluations (d.planesA), or, when they would
// take more than a quarter of the free device memory, the FP16 form keeps the largest magnitude
// of each row of A (d.hintA: the same row scales without a pass over A).  Nothing in the
// allocation-free mode.
int g_kept_gt = 1;

void use_kept_gt(uw12g_context* ctx, const uw12g_base_integrals* X, GemmDesc& d, const char* stage) {
  if (g_kept_gt == 0 || ctx->arena || X->gt_id == 0 || d.A != X->gt.p || d.ks || d.batch != 1) return;
  GemmDesc e = d;
  e.prec = stage_prec(ctx, stage);
  const size_t bytes = plane_cache_bytes(e);
  if (bytes == 0) return;
  if (g_kept_gt == 1 && X->gtp_id == X->gt_id && X->gtp_prec == e.prec && X->gtp.n * 8 >= bytes) {
    d.planesA = X->gtp.p;
    return;
  }
  size_t free_b = 0, total_b = 0;
  UW12G_HIP(hipMemGetInfo(&free_b, &total_b));
  if (g_kept_gt == 1 && (bytes <= X->gtp.n * 8 || bytes <= free_b / 4)) {
    X->gtp_id = 0;
    if (X->gtp.p && bytes > X->gtp.n * 8) UW12G_HIP(hipStreamSynchronize(ctx->stream));
    double* buf = X->gtp.get((bytes + 7) / 8);
    ctx->prof.stage = stage;
    split_a_planes(e, buf, ctx->stream, &ctx->prof);
    X->gtp_id = X->gt_id;
    X->gtp_prec = e.prec;
    d.planesA = buf;
    return;
  }
  // FP16 form, A(m, k) at m + M k (a column-major M x K matrix): its row maxima
  if (e.prec != 4 || d.Am.n0 != 0xffffffffu || d.Am.s0 != 1 || d.Ak.n0 != 0xffffffffu || d.Ak.s0 != d.M) return;
  if (X->gtr_id != X->gt_id || X->gtr.n < size_t(d.M)) {
    X->gtr_id = 0;
    if (X->gtr.p && size_t(d.M) > X->gtr.n) UW12G_HIP(hipStreamSynchronize(ctx->stream));
    abs_max_rows(d.A, d.M, d.K, X->gtr.get(size_t(d.M)), ctx->stream);
    X->gtr_id = X->gt_id;
  }
  d.hintA = X->gtr.p;
  d.hAb = 0;
}

// The same product with the roles of the operands exchanged, C^T = B^T A^T:
// rows and columns of C swap, and with them the operands, their maps, batch
// strides, storage flags and row-scale bounds (the k scale stays on k).
GemmDesc transposed(const GemmDesc& d) {
  GemmDesc t = d;
  t.M = d.N;
  t.N = d.M;
  t.A = d.B;
  t.Am = d.Bn;
  t.Ak = d.Bk;
  t.Ab = d.Bb;
  t.B = d.A;
  t.Bk = d.Ak;
  t.Bn = d.Am;
  t.Bb = d.Ab;
  t.Cm = d.Cn;
  t.Cn = d.Cm;
  t.a32 = d.b32;
  t.b32 = d.a32;
  t.hintA = d.hintB;
  t.hAb = d.hBb;
  t.hintB = d.hintA;
  t.hBb = d.hAb;
  t.planesA = nullptr;  // A's kept planes do not describe the new A
  return t;
}

// U^T Gamma for the projector P: Gamma's AO+RI rows rotated onto the r kept
// eigenvectors U of the ABS overlap.  With S+ = U diag(1/s) U^T every
// contraction over the ABS index runs over those r rows, S+ becoming the
// diagonal 1/s (ri_utils.hpp:45-74 drops the other M - r directions).  Built
// in FP64 once per (integrals, projector) pair and kept on the integrals.
// In the no-allocation mode a tensor not already kept is formed in the arena
// for the one call (named `name`) and not kept.
const double* rotated(uw12g_context* ctx, const uw12g_base_integrals* X, const uw12g_abs_projectors* P,
                      const char* name) {
  if (X->gt_id == P->id) return X->gt.p;
  const int M = P->M, N = X->N, Mr = P->r, D = X->D;
  double* out = nullptr;
  if (ctx->arena) {
    out = ctx->buf(name, size_t(std::max(Mr, 1)) * N * std::max(D, 1));
  } else {
    X->gt_id = 0;
    if (X->gt.p && size_t(Mr) * N * D > X->gt.n) UW12G_HIP(hipStreamSynchronize(ctx->stream));
    out = X->gt.get(size_t(std::max(Mr, 1)) * N * std::max(D, 1));
  }
  if (Mr > 0 && D > 0) {
    const int prec = ctx->prec;
    ctx->prec = 0;
    GemmDesc g;  // Gt[m, alpha, A] = sum_mu' U[mu', m] Gamma[mu', alpha, A]
    g.M = Mr;
    g.N = N;
    g.K = M;
    g.batch = D;
    g.A = P->U.p;
    g.Am = lin(M);
    g.Ak = lin(1);
    g.B = X->gamma.p;
    g.Bk = lin(1);
    g.Bn = lin(M);
    g.Bb = int64_t(M) * N;
    g.C = out;
    g.Cm = lin(1);
    g.Cn = lin(Mr);
    g.Cb = int64_t(Mr) * N;
    run_gemm(ctx, g, "abs_rotate");
    ctx->prec = prec;
  }
  if (!ctx->arena) X->gt_id = P->id;
  return out;
}

void ensure_rotated(uw12g_context* ctx, const uw12g_base_integrals* X, const uw12g_abs_projectors* P) {
  (void)rotated(ctx, X, P, "gt");
}

}  // namespace uw12g

namespace {

// Active j per block of the pair-symmetric L contraction of the same-spin 3e Fock
constexpr int kPairBlock = 16;

template <class F>
int guard(F&& f) {
  try {
    f();
    return 0;
  } catch (const Error& e) {
    g_last_error = e.what();
    return e.code;
  } catch (const std::exception& e) {
    g_last_error = e.what();
    return -1;
  } catch (...) {
    g_last_error = "unknown error";
    return -1;
  }
}





// FP64 symmetric eigendecomposition on the device (rocSOLVER syevd, lower
// triangle, ascending eigenvalues -- the LAPACK routine behind Armadillo's
// eig_sym used at linalg.hpp:1112-1145).  A (device, n x n) is overwritten by
// the eigenvectors; eigenvalues are returned on the host.
std::vector<double> eig_device(uw12g_context* ctx, double* A, int n) {
  std::vector<double> w(n);
  if (n == 0) return w;
  DBuf D, E;
  D.get(n);
  E.get(n);
  int* info = nullptr;
  UW12G_HIP(hipMalloc(&info, sizeof(int)));
  const rocblas_status st =
      rocsolver_dsyevd(ctx->rb, rocblas_evect_original, rocblas_fill_lower, n, A, n, D.p, E.p, info);
  int hinfo = 0;
  UW12G_HIP(hipMemcpyAsync(&hinfo, info, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
  UW12G_HIP(hipMemcpyAsync(w.data(), D.p, sizeof(double) * n, hipMemcpyDeviceToHost, ctx->stream));
  UW12G_HIP(hipStreamSynchronize(ctx->stream));
  (void)hipFree(info);
  if (st != rocblas_status_success) throw Error(-4, "rocsolver_dsyevd failed (status " + std::to_string(st) + ")");
  if (hinfo != 0) throw Error(-4, "eigen-decomposition failed");
  return w;
}

bool is_symmetric(const double* X, size_t n, double tol) {
  for (size_t j = 0; j < n; ++j)
    for (size_t i = j + 1; i < n; ++i)
      if (std::fabs(X[i + n * j] - X[j + n * i]) > tol) return false;
  return true;
}

// Source of the three-index input, one block of DF columns at a time: host
// arrays, device arrays, or per-shell callbacks (base_integrals.hpp:52-72,
// integral_functions.hpp:16-55).  For callbacks, blocks are groups of whole
// shells.
struct ThreeIndexSource {
  const double* X3 = nullptr;  // packed AO rows x n_df (host or device)
  const double* X3ri = nullptr;
  bool device = false;
  uw12g_three_index_fn fn = nullptr, fn_ri = nullptr;
  void* user = nullptr;
  std::vector<size_t> shells;  // DF shell sizes (callbacks)
  uw12g_three_index_gen_fn gen = nullptr;  // device generator
  bool stored = true;                      // false: keep T, not Gamma (device generator only)
};

struct PinnedBuf {  // page-locked host staging (hipHostMalloc)
  double* p = nullptr;
  ~PinnedBuf() {
    if (p) (void)hipHostFree(p);
  }
  void get(size_t n) {
    UW12G_HIP(hipHostMalloc(reinterpret_cast<void**>(&p), sizeof(double) * std::max<size_t>(n, 1), hipHostMallocDefault));
  }
};

void setup_base(uw12g_context* ctx, const ThreeIndexSource& src, const double* X2, size_t n_ao, size_t n_df,
                size_t n_ri, uw12g_base_integrals** out) {
  require(ctx != nullptr, "null context");
  require(out != nullptr, "null output handle");
  require(X2 != nullptr, "null integral array");
  const bool cb = src.fn != nullptr, gen = src.gen != nullptr;
  require(cb || gen || src.X3 != nullptr, "null integral array");
  require(n_ao > 0, "Number of ao orbitals is zero");
  require(n_df > 0, "Number of density-fitting basis functions is zero");
  if (cb) require((src.fn_ri == nullptr) == (n_ri == 0), "three_index_ri and n_ri must both be given or both be absent");
  else if (!gen) require((src.X3ri == nullptr) == (n_ri == 0), "X3_ri and n_ri must both be given or both be absent");
  require(n_ao < 46341 && n_df < (1u << 30) && n_ri < (1u << 30), "size out of range", -2);
  if (cb) {
    size_t tot = 0;
    for (size_t sz : src.shells) {
      require(sz > 0, "empty density-fitting shell");
      tot += sz;
    }
    require(tot == n_df, "density-fitting shell sizes do not add up to n_df");
  }
  require(is_symmetric(X2, n_df, 1e-10), "matrix must be symmetric");
  set_device(ctx);
  std::unique_ptr<uw12g_base_integrals> I(new uw12g_base_integrals());
  I->ctx = ctx;
  const int N = int(n_ao), Nd = int(n_df), Nr = int(n_ri), M = N + Nr;
  I->N = N;
  I->Nd = Nd;
  I->Nr = Nr;
  I->M = M;
  hipStream_t s = ctx->stream;
  // ---- DF metric: eig on device; keep |lambda| >= 1e-10 (pinv tolerance, linalg.hpp:469-481)
  DBuf Q;
  Q.get(size_t(Nd) * Nd);
  UW12G_HIP(hipMemcpyAsync(Q.p, X2, sizeof(double) * Nd * Nd, hipMemcpyHostToDevice, s));
  std::vector<double> lam = eig_device(ctx, Q.p, Nd);
  std::vector<int> keep;
  std::vector<double> scale;
  for (int A = 0; A < Nd; ++A)
    if (std::fabs(lam[A]) >= 1e-10) {
      keep.push_back(A);
      scale.push_back(1.0 / std::sqrt(std::fabs(lam[A])));
      I->sign_host.push_back(lam[A] > 0 ? 1.0 : -1.0);
    }
  const int D = int(keep.size());
  I->D = D;
  DBuf T;
  T.get(size_t(Nd) * std::max(D, 1));
  I->sign.get(std::max(D, 1));
  if (D > 0) {
    int* keep_d = nullptr;
    DBuf sc;
    sc.get(D);
    UW12G_HIP(hipMalloc(&keep_d, sizeof(int) * D));
    UW12G_HIP(hipMemcpyAsync(keep_d, keep.data(), sizeof(int) * D, hipMemcpyHostToDevice, s));
    UW12G_HIP(hipMemcpyAsync(sc.p, scale.data(), sizeof(double) * D, hipMemcpyHostToDevice, s));
    UW12G_HIP(hipMemcpyAsync(I->sign.p, I->sign_host.data(), sizeof(double) * D, hipMemcpyHostToDevice, s));
    gather_scale_cols(Q.p, Nd, keep_d, sc.p, T.p, D, s);
    UW12G_HIP(hipStreamSynchronize(s));
    (void)hipFree(keep_d);
  }
  Q.reset();
  if (!src.stored) {  // non-stored: T and the generator are all an evaluation needs
    I->stored = false;
    I->gen = src.gen;
    I->gen_user = src.user;
    I->T.get(size_t(Nd) * std::max(D, 1));
    if (D > 0) UW12G_HIP(hipMemcpyAsync(I->T.p, T.p, sizeof(double) * Nd * D, hipMemcpyDeviceToDevice, s));
    I->gamma.get(1);
    UW12G_HIP(hipStreamSynchronize(s));
    adopt_child(ctx);
    *out = I.release();
    return;
  }
  // ---- Gamma = G0 T, streamed over DF column blocks of the input
  const size_t MN = size_t(M) * N;
  I->gamma.get(MN * std::max(D, 1));
  if (D == 0) {
    adopt_child(ctx);
    *out = I.release();
    return;
  }
  const size_t P = size_t(N) * (N + 1) / 2, PR = size_t(N) * Nr;
  // column blocks [A0, A0 + na): at most `budget` doubles of G0 per block; with
  // callbacks whole shells, and the host staging capped at 2^25 doubles per buffer
  const size_t budget = size_t(1) << 28;
  int cA = int(std::max<size_t>(1, std::min<size_t>(Nd, budget / std::max<size_t>(MN, 1))));
  std::vector<std::pair<int, int>> blocks;  // (A0, na)
  std::vector<size_t> shell_first;           // callbacks: first shell of each block
  if (cb) {
    const size_t cap = std::max<size_t>(1, ctx->stream_cap / std::max<size_t>(P + PR, 1));
    cA = int(std::min<size_t>(size_t(cA), cap));
    size_t A = 0, sh = 0;
    while (sh < src.shells.size()) {
      const size_t first = sh, A0 = A;
      size_t na = 0;
      do {
        na += src.shells[sh];
        A += src.shells[sh];
        ++sh;
      } while (sh < src.shells.size() && na + src.shells[sh] <= size_t(cA));
      blocks.emplace_back(int(A0), int(na));
      shell_first.push_back(first);
    }
  } else {
    for (int A0 = 0; A0 < Nd; A0 += cA) blocks.emplace_back(A0, std::min(cA, Nd - A0));
  }
  int wmax = 0;
  for (auto& b : blocks) wmax = std::max(wmax, b.second);
  DBuf G0, st3, stri;
  G0.get(MN * wmax);
  if (!src.device && !gen) {
    st3.get(P * wmax);
    if (Nr > 0) stri.get(PR * wmax);
  }
  PinnedBuf h3, hri;  // callbacks fill these; each block's copy completes before the next fill
  if (cb) {
    h3.get(P * wmax);
    if (Nr > 0) hri.get(PR * wmax);
  }
  for (size_t bi = 0; bi < blocks.size(); ++bi) {
    const int A0 = blocks[bi].first, na = blocks[bi].second;
    const double* x3 = nullptr;
    const double* xri = nullptr;
    if (gen) {
      const int rc = src.gen(src.user, size_t(A0), size_t(na), Nr > 0 ? 1 : 0, G0.p, s);
      require(rc == 0, "three-index generator failed for DF functions " + std::to_string(A0) + ".." +
                           std::to_string(A0 + na), -6);
    } else if (src.device) {
      x3 = src.X3 + P * A0;
      xri = Nr > 0 ? src.X3ri + PR * A0 : nullptr;
    } else if (!cb) {
      UW12G_HIP(hipMemcpyAsync(st3.p, src.X3 + P * A0, sizeof(double) * P * na, hipMemcpyHostToDevice, s));
      if (Nr > 0)
        UW12G_HIP(hipMemcpyAsync(stri.p, src.X3ri + PR * A0, sizeof(double) * PR * na, hipMemcpyHostToDevice, s));
      x3 = st3.p;
      xri = stri.p;
    } else {
      UW12G_HIP(hipStreamSynchronize(s));  // the previous block's copies have drained the staging
      size_t col = 0;
      for (size_t sh = shell_first[bi]; col < size_t(na); ++sh) {
        const int rc = src.fn(src.user, sh, h3.p + P * col);
        require(rc == 0, "three_index callback failed for shell " + std::to_string(sh), -6);
        if (Nr > 0) {
          const int rr = src.fn_ri(src.user, sh, hri.p + PR * col);
          require(rr == 0, "three_index_ri callback failed for shell " + std::to_string(sh), -6);
        }
        col += src.shells[sh];
      }
      UW12G_HIP(hipMemcpyAsync(st3.p, h3.p, sizeof(double) * P * na, hipMemcpyHostToDevice, s));
      if (Nr > 0) UW12G_HIP(hipMemcpyAsync(stri.p, hri.p, sizeof(double) * PR * na, hipMemcpyHostToDevice, s));
      x3 = st3.p;
      xri = stri.p;
    }
    if (!gen) assemble_gamma(x3, xri, G0.p, N, Nr, na, s);
    GemmDesc g;
    g.M = int(MN);
    g.N = D;
    g.K = na;
    g.alpha = 1.0;
    g.beta = A0 == 0 ? 0.0 : 1.0;
    g.A = G0.p;
    g.Am = lin(1);
    g.Ak = lin(int64_t(MN));
    g.B = T.p + A0;
    g.Bk = lin(1);
    g.Bn = lin(Nd);
    g.C = I->gamma.p;
    g.Cm = lin(1);
    g.Cn = lin(int64_t(MN));
    gemm_f64(g, s, &ctx->gemm_ws);
  }
  UW12G_HIP(hipStreamSynchronize(s));
  adopt_child(ctx);
  *out = I.release();
}

// ------------------------------------------------------------------ eval
void evaluate(uw12g_context* ctx, const uw12g_base_integrals* W, const uw12g_base_integrals* V,
              const uw12g_base_integrals* WV, const uw12g_abs_projectors* P, const double* C, size_t n_ao,
              size_t n_orb, size_t n_spin, const double* occ, const size_t* n_occ, const size_t* n_active,
              bool indirect_term, double k0, double k1, double* fock, double* terms, double* energy,
              EvalMode mode = kFused) {
  require(ctx && W && V && WV && P, "null handle");
  require(C != nullptr && occ != nullptr && n_occ != nullptr && n_active != nullptr, "null input array");
  require(n_spin == 1 || n_spin == 2, "Number of spin channels must be 1 or 2");
  const int ns = int(n_spin), N = int(n_ao);
  require(W->N == N && V->N == N && WV->N == N && P->N == N, "Inconsistent number of atomic orbitals");
  require(W->Nr > 0 && V->Nr > 0, "ri integrals requested, but no function provided to calculate");
  require(W->Nr == P->Nr && V->Nr == P->Nr, "Total number of basis functions is inconsistent");
  require(W->ctx == ctx && V->ctx == ctx && WV->ctx == ctx && P->ctx == ctx, "objects belong to another context");
  set_device(ctx);
  hipStream_t st = ctx->stream;
  const int M = P->M, Mr = P->r, DW = W->D, DV = V->D, DWV = WV->D;
  const bool calc_fock = fock != nullptr;
  const bool do_ind = indirect_term && k1 != 0.0;
  const double e = ns == 1 ? 2.0 : 1.0;
  ctx->prof.clear();
  ctx->t_eval = std::chrono::steady_clock::now();
  const bool streamed = !W->stored || !V->stored;
  if (streamed) {
    require(mode == kFused, "the differentiation form needs stored W and V integrals", -2);
    require(!do_ind, "the same-spin terms need stored W and V integrals (indirect_term = 0 or scale_same_spin = 0)",
            -2);
    require(ctx->prec != UW12G_FP32, "the FP32 form needs stored W and V integrals", -2);
  }
  // U^T Gamma of W and V (kept on the integrals; in the no-allocation mode formed in the arena
  // unless already kept)
  const double* gtp[2] = {nullptr, nullptr};
  if (!streamed) {
    require(WV->stored, "non-stored WV integrals need non-stored W and V integrals", -2);
    gtp[0] = rotated(ctx, W, P, "gtW");
    gtp[1] = rotated(ctx, V, P, "gtV");
  } else {
    require(!ctx->arena, "the no-allocation entry point needs stored integrals", -2);
  }

  // ---- host: occupation-weighted orbitals and active block (utils.hpp:235-310)
  std::vector<SpinData> sd(ns);
  std::vector<std::vector<double>> Co_h(ns), wk_h(ns);
  size_t occ_off = 0;
  for (int s = 0; s < ns; ++s) {
    const int o = int(n_occ[s]), a = int(n_active[s]);
    require(size_t(o) <= n_orb, "more occupied orbitals than orbitals provided");
    require(a <= o, "n_active exceeds the number of occupied orbitals");
    const double g = ns == 1 ? 2.0 : 1.0;
    sd[s].o = o;
    sd[s].a = a;
    sd[s].c = o - a;
    Co_h[s].resize(size_t(N) * std::max(o, 1));
    for (int k = 0; k < o; ++k) {
      const double w = occ[occ_off + k];
      require(w > 0, "occupations must be positive");
      const double scl = std::sqrt(w / g);
      for (int m = 0; m < N; ++m) Co_h[s][m + size_t(N) * k] = C[m + size_t(N) * (size_t(s) * n_orb + k)] * scl;
    }
    occ_off += o;
    wk_h[s].assign(std::max(o, 1), 1.0);
    if (a > 0)
      for (int k = o - a; k < o; ++k) wk_h[s][k] = 2.0;
  }
  double* Eacc = ctx->buf("E", 8);
  double* partial = ctx->buf("partial", kMaxPartials);
  fill(Eacc, 8, 0.0, st);
  double* Fraw = nullptr;
  if (calc_fock) {
    Fraw = ctx->buf("Fraw", size_t(N) * N * ns);
    fill(Fraw, int64_t(N) * N * ns, 0.0, st);
  }
  for (int s = 0; s < ns; ++s) {
    const int o = sd[s].o;
    sd[s].Co = ctx->buf("Co" + std::to_string(s), size_t(N) * std::max(o, 1));
    sd[s].wk = ctx->buf("wk" + std::to_string(s), std::max(o, 1));
    UW12G_HIP(hipMemcpyAsync(sd[s].Co, Co_h[s].data(), sizeof(double) * N * std::max(o, 1), hipMemcpyHostToDevice, st));
    UW12G_HIP(hipMemcpyAsync(sd[s].wk, wk_h[s].data(), sizeof(double) * std::max(o, 1), hipMemcpyHostToDevice, st));
  }
  const size_t NN = size_t(N) * N;
  auto results = [&] {
    double Eh[8];
    UW12G_HIP(hipMemcpyAsync(Eh, Eacc, sizeof(double) * 8, hipMemcpyDeviceToHost, st));
    if (calc_fock) {
      double* Fs = ctx->buf("Fsym", NN * ns);
      for (int s = 0; s < ns; ++s) symmetrise_add(Fraw + NN * s, Fs + NN * s, N, 0, st);
      UW12G_HIP(hipMemcpyAsync(fock, Fs, sizeof(double) * NN * ns, hipMemcpyDeviceToHost, st));
    }
    UW12G_HIP(hipStreamSynchronize(st));
    for (int i = 0; i < 6; ++i) terms[i] = Eh[i];
    if (energy) {
      double tot = 0.0;
      for (int i = 0; i < 6; ++i) tot += terms[i];
      *energy = tot;
    }
  };
  if (streamed) {
    evaluate_streamed(ctx, W, V, WV, P, sd, k0, k1, Fraw, Eacc, partial);
    results();
    return;
  }

  // ================= two-electron term (two_electron.cpp:11-152) =================
  {
    double* DCa = ctx->buf("DCa", NN * ns);
    for (int s = 0; s < ns; ++s) {
      const int a = sd[s].a;
      const double* Ca = sd[s].Co + size_t(N) * sd[s].c;
      GemmDesc g;  // DCa = Ca Ca^T
      g.M = N;
      g.N = N;
      g.K = a;
      g.A = Ca;
      g.Am = lin(1);
      g.Ak = lin(N);
      g.B = Ca;
      g.Bk = lin(N);
      g.Bn = lin(1);
      g.C = DCa + NN * s;
      g.Cm = lin(1);
      g.Cn = lin(N);
      run_gemm(ctx, g, "2e_density", true);
    }
    // X_D[s]_A = sum_j (jj|A~)  (integrals.cpp:347-378); HBM-bound pass over Gamma_WV
    double* XD = ctx->buf("XD", size_t(std::max(DWV, 1)) * ns);
    // partial sums of the two passes: gemv_t keeps at most ~4096 + D_WV per vector
    const int64_t gpart_n = std::max<int64_t>(2 * (4096 + int64_t(DWV)), calc_fock ? gemv_n_scratch(NN, ns, DWV) : 0);
    double* gpart = ctx->buf("gemv_part", size_t(gpart_n));
    timed(ctx, "2e_XD", "gemv_t_kernel", 2.0 * NN * DWV * ns, 8.0 * NN * DWV, [&] {
      gemv_t(WV->gamma.p, int64_t(NN), int64_t(NN), DCa, int64_t(NN), ns, XD, DWV, DWV, gpart, gpart_n, st);
    });
    // E2_direct and the Fock's DF vector on the device (no host round trip)
    double f2[2][2] = {{0.0, 0.0}, {0.0, 0.0}};
    for (int s = 0; s < ns; ++s)
      for (int t = 0; t < ns; ++t) f2[s][t] = spin_factor(ns, s, t, k0, k1);
    double* vd = calc_fock ? ctx->buf("XDv", size_t(std::max(DWV, 1)) * ns) : nullptr;
    if (DWV > 0) xd_combine(XD, DWV, ns, f2, WV->sign.p, vd, Eacc + UW12G_E2_DIRECT, st);
    if (calc_fock && DWV > 0) {
      timed(ctx, "2e_fock", "gemv_n_kernel", 2.0 * NN * DWV * ns, 8.0 * NN * DWV, [&] {
        gemv_n(WV->gamma.p, int64_t(NN), int64_t(NN), vd, DWV, ns, Fraw, int64_t(NN), DWV, gpart, gpart_n, st);
      });
    }
    if (do_ind) {
      for (int s = 0; s < ns; ++s) {
        const int a = sd[s].a;
        if (a == 0) continue;
        const double* Ca = sd[s].Co + size_t(N) * sd[s].c;
        double* P1 = ctx->buf("wv_phi", size_t(N) * a * DWV);
        double* O2 = ctx->buf("wv_o2", size_t(a) * a * DWV);
        GemmDesc g;  // P1[mu, j, A] = sum_nu Gamma_WV[mu, nu, A] Ca[nu, j]
        g.M = N;
        g.N = a;
        g.K = N;
        g.batch = DWV;
        g.A = WV->gamma.p;
        g.Am = lin(1);
        g.Ak = lin(N);
        g.Ab = int64_t(NN);
        g.B = Ca;
        g.Bk = lin(1);
        g.Bn = lin(N);
        g.C = P1;
        g.Cm = lin(1);
        g.Cn = lin(N);
        g.Cb = int64_t(N) * a;
        run_gemm(ctx, g, "2e_phi", true);
        GemmDesc h;  // O2[i, (j,A)] = sum_mu Ca[mu,i] P1[mu,(j,A)]
        h.M = a;
        h.N = a * DWV;
        h.K = N;
        h.A = Ca;
        h.Am = lin(N);
        h.Ak = lin(1);
        h.B = P1;
        h.Bk = lin(1);
        h.Bn = lin(N);
        h.C = O2;
        h.Cm = lin(1);
        h.Cn = lin(a);
        run_gemm(ctx, h, "2e_o2", true);
        DotDesc dd;  // E -= 1/2 e k1 sum_A s_A |T_A|^2  (two_electron.cpp:112-114)
        dd.x = O2;
        dd.n = int64_t(a) * a * DWV;
        dd.col_w = WV->sign.p;
        dd.ld = int64_t(a) * a;
        dd.col_n = DWV;
        dot(dd, -0.5 * e * k1, Eacc + UW12G_E2_INDIRECT, 1, partial, st);
        if (calc_fock) {  // F -= k1 R diag(s) R^T  (two_electron.cpp:116-128)
          GemmDesc f;
          f.M = N;
          f.N = N;
          f.K = a * DWV;
          f.alpha = -k1;
          f.beta = 1.0;
          f.A = P1;
          f.Am = lin(1);
          f.Ak = lin(N);
          f.ks = WV->sign.p;
          f.Km = two(uint32_t(a), 0, 1);
          f.B = P1;
          f.Bk = lin(N);
          f.Bn = lin(1);
          f.C = Fraw + NN * s;
          f.Cm = lin(1);
          f.Cn = lin(N);
          run_gemm(ctx, f, "2e_ind_fock", true);
        }
      }
    }
  }

  // Psi[m, k, A] = sum_alpha (U^T Gamma)[m, alpha, A] Co[alpha, k] for one spin and operator
  auto psi_transform = [&](const SpinData& q, int op) {
    const uw12g_base_integrals* X = op == 0 ? W : V;
    GemmDesc gp;
    gp.M = Mr;
    gp.N = q.o;
    gp.K = N;
    gp.batch = X->D;
    gp.A = gtp[op];
    gp.Am = lin(1);
    gp.Ak = lin(Mr);
    gp.Ab = int64_t(Mr) * N;
    gp.B = q.Co;
    gp.Bk = lin(1);
    gp.Bn = lin(N);
    gp.C = op == 0 ? q.psiW : q.psiV;
    gp.Cm = lin(1);
    gp.Cn = lin(Mr);
    gp.Cb = int64_t(Mr) * q.o;
    run_gemm(ctx, gp, "psi_transform", true);
  };
  // A fused evaluation of two spins keeps one set of the large ABS-row buffers (Psi, PhiT_V)
  // and re-forms Psi per spin for the same-spin terms; forward / adjoint keep both (the tape).
  // The shared set is sized for the larger occupied / active count of the two spins up front:
  // growing it for the second spin would free the memory the first spin's pointers still name.
  const bool share = mode == kFused && ns == 2;
  int o_max = 0, a_max = 0;
  for (int s = 0; s < ns; ++s) {
    o_max = std::max(o_max, sd[s].o);
    a_max = std::max(a_max, sd[s].a);
  }

  // ================= W/V transforms: Phi, O2, t, PhiT_V, X =================
  for (int s = 0; s < ns; ++s) {
    SpinData& q = sd[s];
    const int o = q.o, a = q.a, c = q.c;
    const std::string sfx = std::to_string(s), psfx = share ? std::string("s") : sfx;
    q.phiW = ctx->buf("phiW" + sfx, size_t(N) * o * DW);
    q.phiV = ctx->buf("phiV" + sfx, size_t(N) * o * DV);
    const int o_buf = share ? o_max : o, a_buf = share ? a_max : a;
    q.psiW = ctx->buf("psiW" + psfx, size_t(Mr) * o_buf * DW);
    q.psiV = ctx->buf("psiV" + psfx, size_t(Mr) * o_buf * DV);
    q.o2W = ctx->buf("o2W" + sfx, size_t(a) * o * DW);
    q.o2V = ctx->buf("o2V" + sfx, size_t(a) * o * DV);
    q.t = ctx->buf("t" + sfx, size_t(DW) * DV);
    q.x = ctx->buf("x" + sfx, size_t(DW) * DV);
    q.phitV = ctx->buf("phitV" + psfx, size_t(Mr) * a_buf * DV);
    if (mode == kAdjoint) continue;  // kept from the forward call
    const double* Ca = q.Co + size_t(N) * c;
    for (int op = 0; op < 2; ++op) {
      const uw12g_base_integrals* X = op == 0 ? W : V;
      double* phi = op == 0 ? q.phiW : q.phiV;
      double* o2 = op == 0 ? q.o2W : q.o2V;
      GemmDesc g;  // Phi[mu, k, A] = sum_alpha Gamma[mu, alpha, A] Co[alpha, k] (AO rows mu)
      g.M = N;
      g.N = o;
      g.K = N;
      g.batch = X->D;
      g.A = X->gamma.p;
      g.Am = lin(1);
      g.Ak = lin(M);
      g.Ab = int64_t(M) * N;
      g.B = q.Co;
      g.Bk = lin(1);
      g.Bn = lin(N);
      g.C = phi;
      g.Cm = lin(1);
      g.Cn = lin(N);
      g.Cb = int64_t(N) * o;
      run_gemm(ctx, g, "phi_transform", true);
      psi_transform(q, op);
      GemmDesc h;  // O2[i, (k, A)] = sum_mu Ca[mu, i] Phi[mu, (k, A)]
      h.M = a;
      h.N = o * X->D;
      h.K = N;
      h.A = Ca;
      h.Am = lin(N);
      h.Ak = lin(1);
      h.B = phi;
      h.Bk = lin(1);
      h.Bn = lin(N);
      h.C = o2;
      h.Cm = lin(1);
      h.Cn = lin(a);
      run_gemm(ctx, h, "o2_transform", true);
    }
    GemmDesc t;  // t = O2_W^T O2_V  (four_electron_utils.cpp:33-49)
    t.M = DW;
    t.N = DV;
    t.K = a * o;
    t.A = q.o2W;
    t.Am = lin(int64_t(a) * o);
    t.Ak = lin(1);
    t.B = q.o2V;
    t.Bk = lin(1);
    t.Bn = lin(int64_t(a) * o);
    t.C = q.t;
    t.Cm = lin(1);
    t.Cn = lin(DW);
    run_gemm(ctx, t, "4e_tab", true);
    // S+ Phi_V[:, c+i, B] = U PhiT_V[:, i, B] with PhiT_V[m, i, B] = (1/s_m) Psi_V[m, c+i, B]
    timed(ctx, "3e_abs_apply", "scale_rows_kernel", double(Mr) * a * DV, 16.0 * Mr * a * DV, [&] {
      scale_rows(q.psiV + size_t(Mr) * c, Mr, a, DV, Mr, int64_t(Mr) * o, P->dinv.p, nullptr, q.phitV, st);
    });
    GemmDesc x;  // X[A, B] = sum_{m, i} Psi_W[m, c+i, A] PhiT_V[m, i, B]  (direct_utils.cpp:62-133)
    x.M = DW;
    x.N = DV;
    x.K = Mr * a;
    x.A = q.psiW + size_t(Mr) * c;
    x.Am = lin(int64_t(Mr) * o);
    x.Ak = lin(1);
    x.B = q.phitV;
    x.Bk = lin(1);
    x.Bn = lin(int64_t(Mr) * a);
    x.C = q.x;
    x.Cm = lin(1);
    x.Cn = lin(DW);
    run_gemm(ctx, x, "3e_xab", true);
  }
  progress(ctx, "orbital transforms, t and X");
  // direct energies (four_electron.cpp:118-119, direct_utils.cpp:136-157)
  for (int s = 0; s < ns; ++s)
    for (int u = 0; u < ns; ++u) {
      const double f = spin_factor(ns, s, u, k0, k1);
      if (f == 0) continue;
      DotDesc dd;
      dd.n = int64_t(DW) * DV;
      dd.row_w = W->sign.p;
      dd.col_w = V->sign.p;
      dd.ld = DW;
      dd.col_n = DV;
      dd.x = sd[u].t;
      dd.y = sd[s].t;
      dot(dd, 0.5 * f, Eacc + UW12G_E4_DIRECT, 1, partial, st);
      dd.x = sd[s].x;
      dd.y = sd[u].t;
      dot(dd, -f, Eacc + UW12G_E3_DIRECT, 1, partial, st);
    }

  // ================= direct Fock: 4e + 3e (X dt~, t~ dX) =================
  if (calc_fock) {
    double* Tb = ctx->buf("Tb", size_t(DW) * DV);
    double* Z = ctx->buf("Z", size_t(DW) * DV);
    for (int s = 0; s < ns; ++s) {
      SpinData& q = sd[s];
      const int o = q.o;
      const double fs = ns == 1 ? 0.5 : 1.0;
      const double c0 = fs * spin_factor(ns, s, 0, k0, k1);
      const double c1 = ns == 2 ? fs * spin_factor(ns, s, 1, k0, k1) : 0.0;
      combine_dd(DW, DV, sd[0].t, ns == 2 ? sd[1].t : nullptr, sd[0].x, ns == 2 ? sd[1].x : nullptr, c0, c1,
                 W->sign.p, V->sign.p, Tb, Z, st);
      double* G = ctx->buf("G", size_t(N) * o * DV);
      GemmDesc g;  // G[(beta,k), B] = sum_A W1[beta, k, A] Z[A, B]
      g.M = N * o;
      g.N = DV;
      g.K = DW;
      g.A = q.phiW;
      g.Am = lin(1);
      g.Ak = lin(int64_t(N) * o);
      g.B = Z;
      g.Bk = lin(1);
      g.Bn = lin(DW);
      g.C = G;
      g.Cm = lin(1);
      g.Cn = lin(int64_t(N) * o);
      run_gemm(ctx, g, "direct_fock_G");
      GemmDesc f;  // F += sum_{k,B} V1[alpha,k,B] w_k G[beta,k,B]  (four_electron.cpp:11-77, direct_utils.cpp:164-204)
      f.M = N;
      f.N = N;
      f.K = o * DV;
      f.beta = 1.0;
      f.A = q.phiV;
      f.Am = lin(1);
      f.Ak = lin(N);
      f.ks = q.wk;
      f.Km = two(uint32_t(std::max(o, 1)), 1, 0);
      f.B = G;
      f.Bk = lin(N);
      f.Bn = lin(1);
      f.C = Fraw + NN * s;
      f.Cm = lin(1);
      f.Cn = lin(N);
      run_gemm(ctx, f, "direct_fock_F", true);
      // t~ dX/dD (direct_utils.cpp:242-347): F -= sum_AB Tb_AB Gamma_W[:, a, A]^T S+ Gamma_V[:, b, B]
      // = sum_{m, A} Gt_W[m, a, A] (1/s_m) Xi[m, b, A], Xi = Gt_V Tb^T over B
      const size_t MrN = size_t(Mr) * N;
      // A blocks of Xi within half the block cap, of equal size rounded up to a
      // multiple of 64 (whole MFMA tiles along A)
      const int64_t nb_max = std::max<int64_t>(1, std::min<int64_t>(DW, int64_t(ctx->block_cap / 2) /
                                                                            std::max<int64_t>(int64_t(MrN), 1)));
      const int64_t nblk = (DW + nb_max - 1) / nb_max;
      int nb = int((DW + nblk - 1) / nblk);
      if (nb > 64) nb = int(std::min<int64_t>(std::min<int64_t>(DW, nb_max), (nb + 63) / 64 * 64));
      double* Xi = ctx->buf("blk3", MrN * std::max(nb, 1));
      for (int A0 = 0; Mr > 0 && A0 < DW; A0 += nb) {
        const int na = std::min(nb, DW - A0);
        GemmDesc a1;  // Xi[(m, beta), A] = sum_B Gt_V[(m, beta), B] Tb[A0 + A, B]
        a1.M = int(MrN);
        a1.N = na;
        a1.K = DV;
        a1.A = gtp[1];
        a1.Am = lin(1);
        a1.Ak = lin(int64_t(MrN));
        a1.B = Tb + A0;
        a1.Bk = lin(DW);
        a1.Bn = lin(1);
        a1.C = Xi;
        a1.Cm = lin(1);
        a1.Cn = lin(int64_t(MrN));
        // Gt_V is the A operand of every block and every evaluation: its planes (or row maxima) are kept
        use_kept_gt(ctx, V, a1, "3e_ttdx_xi");
        run_gemm(ctx, a1, "3e_ttdx_xi");
        GemmDesc a3;  // F -= sum_{m, A} Gt_W[m, alpha, A0+A] (1/s_m) Xi[m, beta, A]
        a3.M = N;
        a3.N = N;
        a3.K = Mr * na;
        a3.alpha = -1.0;
        a3.beta = 1.0;
        a3.A = gtp[0] + MrN * A0;
        a3.Am = lin(Mr);
        a3.Ak = two(uint32_t(Mr), 1, int64_t(MrN));
        a3.ks = P->dinv.p;
        a3.Km = two(uint32_t(Mr), 1, 0);
        a3.B = Xi;
        a3.Bk = two(uint32_t(Mr), 1, int64_t(MrN));
        a3.Bn = lin(Mr);
        a3.C = Fraw + NN * s;
        a3.Cm = lin(1);
        a3.Cn = lin(N);
        run_gemm(ctx, a3, "3e_ttdx_fock", true);
      }
    }
  }

  if (calc_fock) progress(ctx, "direct Fock");
  // ================= same-spin (indirect) terms, x k1 =================
  if (do_ind) {
    for (int s = 0; s < ns; ++s) {
      SpinData& q = sd[s];
      const int o = q.o, a = q.a, c = q.c;
      if (a == 0 || o == 0) continue;
      const int64_t ao = int64_t(a) * o;
      if (share) {  // the shared ABS-row buffers hold the other spin's transforms
        psi_transform(q, 0);
        psi_transform(q, 1);
        if (!calc_fock)
          scale_rows(q.psiV + size_t(Mr) * c, Mr, a, DV, Mr, int64_t(Mr) * o, P->dinv.p, nullptr, q.phitV, st);
      }
      // signed two-index transforms  W2s = sW o O2_W, V2s = sV o O2_V
      double* W2s = ctx->buf("W2s", size_t(ao) * DW);
      double* V2s = ctx->buf("V2s", size_t(ao) * DV);
      scale_cols(q.o2W, ao, DW, W->sign.p, W2s, st);
      scale_cols(q.o2V, ao, DV, V->sign.p, V2s, st);
      // ---- 4e indirect energy (four_electron.cpp:143-179, 274-278)
      if (mode != kAdjoint) {
        double* W4 = ctx->buf("blk1", size_t(ao) * ao);
        double* V4 = ctx->buf("blk2", size_t(ao) * ao);
        for (int op = 0; op < 2; ++op) {
          GemmDesc g;  // X4 = O2 diag(s) O2^T
          g.M = int(ao);
          g.N = int(ao);
          g.K = op == 0 ? DW : DV;
          g.A = op == 0 ? W2s : V2s;
          g.Am = lin(1);
          g.Ak = lin(ao);
          g.B = op == 0 ? q.o2W : q.o2V;
          g.Bk = lin(ao);
          g.Bn = lin(1);
          g.C = op == 0 ? W4 : V4;
          g.Cm = lin(1);
          g.Cn = lin(ao);
          run_gemm(ctx, g, "4e_ind_x4", true);
        }
        timed(ctx, "4e_ind_energy", "e4i_kernel", 2.0 * ao * ao, 16.0 * ao * ao,
              [&] { e4i_energy(W4, V4, a, o, -0.5 * e * k1, Eacc + UW12G_E4_INDIRECT, partial, st); });
      }
      // ---- 4e indirect Fock (four_electron.cpp:181-249), blocked over i (and l) so that
      // the X_i / Y_i blocks stay within the block cap
      if (calc_fock) {
        // V2sT[B + DV (i + a k)] = sV_B O2_V[i + a k, B]: the DF index contiguous, so the Y_i
        // contraction reads its B operand along unit stride
        double* V2sT = ctx->buf("V2sT", size_t(ao) * DV);
        timed(ctx, "4e_ind_fock_v2t", "transpose_scale_kernel", 0.0, 16.0 * ao * DV,
              [&] { transpose_scale(q.o2V, ao, DV, V->sign.p, V2sT, st); });
        const int64_t per_i = int64_t(N) * ao;
        const int ni = int(std::max<int64_t>(1, std::min<int64_t>(a, int64_t(ctx->block_cap) / per_i)));
        double* Xb = ctx->buf("blk1", size_t(per_i) * ni);
        double* Yb = ctx->buf("blk2", size_t(per_i) * ni);
        const double fz = c > 0 ? 0.5 : 1.0;  // frozen core: 1/2 (F_act + F_frozen)
        for (int i0 = 0; i0 < a; i0 += ni) {
          const int in = std::min(ni, a - i0);
          // X_i and Y_i of the block as single products: i joins X's rows (alpha, i) and Y's
          // columns (i, k), so neither launch pads the short N / o dimension per i
          GemmDesc gx;  // X_i[alpha, (j,k)] = sum_A Phi_W[alpha, c+i, A] sW_A O2_W[(j,k), A]
          gx.M = N * in;
          gx.N = int(ao);
          gx.K = DW;
          gx.A = q.phiW + size_t(N) * (c + i0);
          gx.Am = lin(1);
          gx.Ak = lin(int64_t(N) * o);
          gx.B = W2s;
          gx.Bk = lin(ao);
          gx.Bn = lin(1);
          gx.C = Xb;
          gx.Cm = two(uint32_t(N), 1, per_i);
          gx.Cn = lin(N);
          run_gemm(ctx, gx, "4e_ind_fock_x");
          GemmDesc gy;  // Y_i[(beta, j), k] = sum_B Phi_V[beta, c+j, B] sV_B O2_V[i + a k, B]
          gy.M = N * a;
          gy.N = o * in;
          gy.K = DV;
          gy.A = q.phiV + size_t(N) * c;
          gy.Am = lin(1);
          gy.Ak = lin(int64_t(N) * o);
          gy.B = V2sT + int64_t(DV) * i0;
          gy.Bk = lin(1);
          gy.Bn = two(uint32_t(in), DV, int64_t(DV) * a);
          gy.C = Yb;
          gy.Cm = lin(1);
          gy.Cn = two(uint32_t(in), per_i, int64_t(N) * a);
          run_gemm(ctx, gy, "4e_ind_fock_y");
          GemmDesc gf;  // F -= 2 k1 sum_{(j,k,i)} X[alpha, .] Y[beta, .]
          gf.M = N;
          gf.N = N;
          gf.K = int(ao * in);
          gf.alpha = -2.0 * k1 * fz;
          gf.beta = 1.0;
          gf.A = Xb;
          gf.Am = lin(1);
          gf.Ak = lin(N);
          gf.B = Yb;
          gf.Bk = lin(N);
          gf.Bn = lin(1);
          gf.C = Fraw + NN * s;
          gf.Cm = lin(1);
          gf.Cn = lin(N);
          run_gemm(ctx, gf, "4e_ind_fock_f", true);
        }
        if (c > 0) {
          const int nl = int(std::max<int64_t>(1, std::min<int64_t>(o, int64_t(ctx->block_cap) / per_i)));
          double* X2b = ctx->buf("blk3", size_t(per_i) * nl);
          double* Y2b = ctx->buf("blk4", size_t(per_i) * nl);
          for (int l0 = 0; l0 < o; l0 += nl) {
            const int ln = std::min(nl, o - l0);
            GemmDesc hx;  // X'_l[alpha, (j,k)] for every occupied l
            hx.M = N;
            hx.N = int(ao);
            hx.K = DW;
            hx.batch = ln;
            hx.A = q.phiW + size_t(N) * l0;
            hx.Am = lin(1);
            hx.Ak = lin(int64_t(N) * o);
            hx.Ab = N;
            hx.B = W2s;
            hx.Bk = lin(ao);
            hx.Bn = lin(1);
            hx.C = X2b;
            hx.Cm = lin(1);
            hx.Cn = lin(N);
            hx.Cb = per_i;
            run_gemm(ctx, hx, "4e_ind_fock_x");
            GemmDesc hy;  // Y'_l[(beta, k), j] = sum_B Phi_V[beta, k, B] sV_B O2_V[j + a l, B]
            hy.M = N * o;
            hy.N = a;
            hy.K = DV;
            hy.batch = ln;
            hy.A = q.phiV;
            hy.Am = lin(1);
            hy.Ak = lin(int64_t(N) * o);
            hy.B = V2s + int64_t(a) * l0;
            hy.Bk = lin(ao);
            hy.Bn = lin(1);
            hy.Bb = a;
            hy.C = Y2b;
            hy.Cm = two(uint32_t(N), 1, int64_t(N) * a);
            hy.Cn = lin(N);
            hy.Cb = per_i;
            run_gemm(ctx, hy, "4e_ind_fock_y");
            GemmDesc hf;
            hf.M = N;
            hf.N = N;
            hf.K = int(ao * ln);
            hf.alpha = -2.0 * k1 * fz;
            hf.beta = 1.0;
            hf.A = X2b;
            hf.Am = lin(1);
            hf.Ak = lin(N);
            hf.B = Y2b;
            hf.Bk = lin(N);
            hf.Bn = lin(1);
            hf.C = Fraw + NN * s;
            hf.Cm = lin(1);
            hf.Cn = lin(N);
            run_gemm(ctx, hf, "4e_ind_fock_f", true);
          }
        }
        progress(ctx, "4e indirect Fock, spin %d", s);
      }
      // ---- 3e indirect energy (indirect_utils.cpp:19-133), blocked over k.  With the
      // Fock it comes from the Fock's L intermediate instead (below).
      if (mode != kAdjoint && !calc_fock) {
        const int64_t Ma = int64_t(Mr) * a;
        const size_t cap = ctx->block_cap;
        int nk = int(std::max<int64_t>(1, std::min<int64_t>(o, int64_t(cap) / std::max<int64_t>(Ma * a, 1))));
        double* GWb = ctx->buf("blk1", size_t(Ma) * a * nk);
        double* GVb = ctx->buf("blk2", size_t(Ma) * a * nk);
        for (int k0 = 0; Mr > 0 && k0 < o; k0 += nk) {
          const int kn = std::min(nk, o - k0);
          GemmDesc g;  // GW[(m, i), (j, k)] = sum_A Psi_W[m, c+i, A] W2s[(j, k), A]
          g.M = int(Ma);
          g.N = a * kn;
          g.K = DW;
          g.A = q.psiW + size_t(Mr) * c;
          g.Am = lin(1);
          g.Ak = lin(int64_t(Mr) * o);
          g.B = W2s + int64_t(a) * k0;
          g.Bk = lin(ao);
          g.Bn = lin(1);
          g.C = GWb;
          g.Cm = lin(1);
          g.Cn = lin(Ma);
          run_gemm(ctx, g, "3e_ind_e_gw", true);
          GemmDesc h = g;  // GV[(m, j), (i, k)] = sum_B PhiT_V[m, j, B] V2s[(i, k), B]
          h.K = DV;
          h.A = q.phitV;
          h.Ak = lin(Ma);
          h.B = V2s + int64_t(a) * k0;
          h.C = GVb;
          run_gemm(ctx, h, "3e_ind_e_gv", true);
          timed(ctx, "3e_ind_energy", "e3i_kernel", 2.0 * Ma * a * kn, 16.0 * Ma * a * kn,
                [&] { e3i_energy(GWb, GVb, Mr, a, kn, k1 * e, Eacc + UW12G_E3_INDIRECT, partial, st); });
        }
      }
      // ---- 3e indirect Fock (indirect_utils.cpp:138-345)
      if (calc_fock && Mr > 0) {
        // the K~ operands PhiT_Y sY = diag(1/s) Psi_Y[:, act] diag(sY): the DF signs folded in,
        // so the K~ contraction needs no per-k scale
        double* phitWs = ctx->buf("phitWs", size_t(Mr) * a * DW);
        // a fused evaluation no longer needs the unsigned PhiT_V (X is formed): reuse its buffer
        double* phitVs = mode == kFused ? q.phitV : ctx->buf("phitVs", size_t(Mr) * a * DV);
        timed(ctx, "3e_abs_apply", "scale_rows_kernel", double(Mr) * a * DW, 16.0 * Mr * a * DW, [&] {
          scale_rows(q.psiW + size_t(Mr) * c, Mr, a, DW, Mr, int64_t(Mr) * o, P->dinv.p, W->sign.p, phitWs, st);
        });
        timed(ctx, "3e_abs_apply", "scale_rows_kernel", double(Mr) * a * DV, 16.0 * Mr * a * DV, [&] {
          scale_rows(q.psiV + size_t(Mr) * c, Mr, a, DV, Mr, int64_t(Mr) * o, P->dinv.p, V->sign.p, phitVs, st);
        });
        const size_t MrN = size_t(Mr) * N;
        const int64_t MNo = int64_t(MrN) * o;
        const size_t cap = ctx->block_cap;
        // L columns (DF index A) in blocks within twice the block cap; each block re-forms K~
        const int DXmax = std::max(DW, DV);
        const int64_t nA_max = std::max<int64_t>(1, std::min<int64_t>(DXmax, int64_t(2 * cap) / int64_t(MrN)));
        double* L = ctx->buf("blk2", MrN * size_t(nA_max));
        // j-block: K~ within the block cap and, when L runs in FP32, its k = o jn within one
        // FP32 range (no split-K pass over L); blocks of equal size
        int64_t nj_max = std::min<int64_t>(a, int64_t(cap) / std::max<int64_t>(MNo, 1));
        if (ctx->prec != 0) nj_max = std::min<int64_t>(nj_max, kF32Chunk / std::max(o, 1));
        nj_max = std::max<int64_t>(nj_max, 1);
        const int64_t n_blocks = (a + nj_max - 1) / nj_max;
        const int nj = int((a + n_blocks - 1) / n_blocks);
        // The second pass has no R, so its L runs over unordered active pairs.  O2_X[(j, c+k), A]
        // is symmetric in the active j, k (the AO block of Gamma is built from the packed
        // mu >= nu pairs of the input, so it is symmetric), hence
        //   sum_{j,k} K~[., c+k | j] X(j, k) = sum_{j <= k} S(j, k) X(j, k),
        //   S(j, k) = K~[., c+k | j] + K~[., c+j | k] (j < k),  S(j, j) = K~[., c+j | j].
        // The j run in blocks of jb; block [j0, j0 + nb) keeps the core k, its own active k
        // (plain K~) and the active k past it (S), nk = c + a - j0 columns per j, so L's k
        // runs over c a + sum nb (a - j0) pairs instead of o a (about 0.6 o a at jb = 16).
        struct PairBlock {
          int j0, nb, nk;
          int64_t p0;  // first pair of the block in the packed pair order
        };
        std::vector<PairBlock> pblk;
        int64_t n_pairs = 0;
        const int jb = int(std::min<int64_t>(kPairBlock, nj_max));
        for (int j0 = 0; j0 < a; j0 += jb) {
          const int nb = std::min(jb, a - j0), nk = c + a - j0;
          pblk.push_back({j0, nb, nk, n_pairs});
          n_pairs += int64_t(nb) * nk;
        }
        // groups of consecutive blocks whose sums S fit the block cap together (and, when L
        // runs in FP32, one FP32 range of its k): one L contraction per group
        auto pairs_fit = [&](int64_t np) {
          return int64_t(MrN) * np <= int64_t(cap) && (ctx->prec == 0 || np <= kF32Chunk);
        };
        std::vector<std::pair<size_t, size_t>> pgrp;  // [first, end) block indices
        int64_t np_max = 0;
        for (size_t i = 0; i < pblk.size();) {
          size_t e = i;
          int64_t np = 0;
          do {
            np += int64_t(pblk[e].nb) * pblk[e].nk;
            ++e;
          } while (e < pblk.size() && pairs_fit(np + int64_t(pblk[e].nb) * pblk[e].nk));
          pgrp.emplace_back(i, e);
          np_max = std::max(np_max, np);
          i = e;
        }
        // The split forms keep K~ and the pair sums S in FP32 storage: the K~ products sum in
        // FP32 (over D_Y k), so their values are FP32 numbers up to the combination of k chunks,
        // and L and R round them again when they split them into planes; this halves the bytes
        // K~ moves through HBM.  Only the plane-tensor path reads such storage.
        bool k32 = ctx->prec >= UW12G_MIXED_BF16X3 && g_bf16_planes && !ctx->arena;
        for (const std::string& x : ctx->exact_extra)
          if (x.rfind("3e_ind_fock", 0) == 0) k32 = false;
        const size_t kt_n = std::max(size_t(MNo) * nj, MrN * size_t(np_max));
        double* Kt = ctx->buf("blk1", k32 ? (kt_n + 1) / 2 : kt_n);
        auto kt_at = [&](size_t off) { return advance(Kt, int64_t(off), k32); };
        // With FP32 storage the K~, pair-sum and L products run transposed (K~: rows (beta, k),
        // columns (m, j); L: rows A, columns (m, beta)): their output columns are then the
        // contiguous ABS rows m, so each lane stores its four columns as one vector access (and
        // reads them so for beta).
        const bool kt_tr = k32 && Mr % 4 == 0;
        auto run_kt = [&](const GemmDesc& g, const char* stage) {
          run_gemm(ctx, kt_tr ? transposed(g) : g, stage);
        };
        // X2sP[A + DX p] = X2sT[A + DX (j + a k)] for pair p = (j, k) of the packed order
        double* X2sP = ctx->buf("x2sp", size_t(n_pairs) * DV);
        // X2sT[A + DX (j + a k)] = sX_A O2_X[(j, k), A]: the DF index contiguous, so the
        // L contraction below reads its B operand along unit stride
        double* X2sT = ctx->buf("blk3", size_t(ao) * DXmax);
        // fock_k's R needs Q~[m, i, alpha | j] = sum_B PhiT_V[m, j, B] sV_B Phi_V[alpha, c+i, B]
        // (indirect_utils.cpp:199-259), which is the first pass's K~[m, (alpha, c+i) | j]: R and
        // fock_k are formed from each K~ block of that pass before it is overwritten.  The ABS
        // index runs over the Mr rotated rows m throughout.
        double* R = ctx->buf("blk4", size_t(DW) * N * nj);
        double* Lc = ctx->buf("blk5", size_t(Mr) * a * std::min<int64_t>(DW, nA_max));
        // FP16 split form: the row scales of K~ / S (L's rows (m, beta), R's columns (alpha, j))
        // and of K~'s factors come from Cauchy-Schwarz bounds, |K~[m, beta, k | j]| <=
        // ||PhiT_Y[m, j, :]|| ||Phi_Y[beta, k, :]||, instead of passes over those tensors
        const bool cs = ctx->prec == UW12G_MIXED_FP16X3;
        double* nrmT = cs ? ctx->buf("cs_nrmT", size_t(Mr) * a) : nullptr;  // ||PhiT_Y[m, j, :]||
        double* nrmP = cs ? ctx->buf("cs_nrmP", size_t(N) * o) : nullptr;   // ||Phi_Y[beta, k, :]||
        double* csu = cs ? ctx->buf("cs_u", size_t(Mr)) : nullptr;          // max over j
        double* csw = cs ? ctx->buf("cs_w", size_t(a)) : nullptr;           // max over m
        double* csv = cs ? ctx->buf("cs_v", size_t(N)) : nullptr;           // max over k
        double* hintL = cs ? ctx->buf("cs_hintL", MrN) : nullptr;
        double* hintR = cs ? ctx->buf("cs_hintR", size_t(N) * nj) : nullptr;
        // fock_i (X = W, Y = V) then fock_j (X = V, Y = W)
        for (int pass = 0; pass < 2; ++pass) {
          const uw12g_base_integrals* X = pass == 0 ? W : V;
          const uw12g_base_integrals* Y = pass == 0 ? V : W;
          const double* phitY = pass == 0 ? phitVs : phitWs;
          const double* phiY = pass == 0 ? q.phiV : q.phiW;
          const int DX = X->D, DY = Y->D;
          if (cs) {
            row_norms(phitY, int64_t(Mr) * a, int64_t(Mr) * a, DY, nrmT, st);
            row_norms(phiY, int64_t(N) * o, int64_t(N) * o, DY, nrmP, st);
            max_over(nrmT, Mr, a, true, csu, st);
            max_over(nrmT, Mr, a, false, csw, st);
            max_over(nrmP, N, o, true, csv, st);
            // L's rows: K~ (first pass) or the pair sums S of two K~ elements (second pass)
            outer(csu, Mr, csv, N, pass == 1 ? 2.0 : 1.0, hintL, st);
          }
          timed(ctx, "3e_ind_fock_x2t", "transpose_scale_kernel", 0.0, 16.0 * ao * DX,
                [&] { transpose_scale(pass == 0 ? q.o2W : q.o2V, ao, DX, X->sign.p, X2sT, st); });
          const bool paired = pass == 1;
          if (paired)
            timed(ctx, "3e_ind_fock_x2t", "pair_gather_kernel", 0.0, 16.0 * n_pairs * DX, [&] {
              for (const PairBlock& pb : pblk)
                pair_gather(X2sT, DX, a, c, pb.j0, pb.nb, pb.nk, X2sP + int64_t(DX) * pb.p0, st);
            });
          const int64_t nAb = (DX + nA_max - 1) / nA_max;
          const int nA = int((DX + nAb - 1) / nAb);
          for (int A0 = 0; A0 < DX; A0 += nA) {
            const int An = std::min(nA, DX - A0);
            for (size_t gi = 0; paired && gi < pgrp.size(); ++gi) {
              const int64_t pbase = pblk[pgrp[gi].first].p0;
              for (size_t bi = pgrp[gi].first; bi < pgrp[gi].second; ++bi) {
                const PairBlock& pb = pblk[bi];
                double* S = kt_at(MrN * size_t(pb.p0 - pbase));  // S[(m, beta) + MrN (kk + nk jj)]
                const int nbeyond = a - pb.j0 - pb.nb;
                // S = K~[m, (beta, k) | j], j in the block, k core or active from j0 on; rows (m, j)
                GemmDesc g;
                g.K = DY;
                g.A = phitY + size_t(Mr) * pb.j0;
                g.Am = lin(1);
                g.Ak = lin(int64_t(Mr) * a);
                g.Bk = lin(int64_t(N) * o);
                g.Bn = lin(1);
                g.Cm = two(uint32_t(Mr), 1, int64_t(MrN) * pb.nk);
                g.Cn = lin(Mr);
                g.M = Mr * pb.nb;
                g.c32 = k32;
                if (cs) g.hintA = nrmT + size_t(Mr) * pb.j0;
                if (c > 0) {
                  g.N = N * c;
                  g.B = phiY;
                  g.C = S;
                  if (cs) g.hintB = nrmP;
                  run_kt(g, "3e_ind_fock_kt");
                }
                g.N = N * (a - pb.j0);
                g.B = phiY + size_t(N) * (c + pb.j0);
                g.C = advance(S, int64_t(MrN) * c, k32);
                if (cs) g.hintB = nrmP + size_t(N) * (c + pb.j0);
                run_kt(g, "3e_ind_fock_kt");
                if (nbeyond > 0) {
                  // S[., c + k - j0 | j] += K~[m, (beta, c+j) | k] for the active k past the block:
                  // rows (m, k), columns (beta, j), one launch
                  GemmDesc h;
                  h.M = Mr * nbeyond;
                  h.N = N * pb.nb;
                  h.K = DY;
                  h.beta = 1.0;
                  h.A = phitY + size_t(Mr) * (pb.j0 + pb.nb);
                  h.Am = lin(1);
                  h.Ak = lin(int64_t(Mr) * a);
                  h.B = phiY + size_t(N) * (c + pb.j0);
                  h.Bk = lin(int64_t(N) * o);
                  h.Bn = lin(1);
                  h.C = advance(S, int64_t(MrN) * (c + pb.nb), k32);
                  h.c32 = k32;
                  h.Cm = two(uint32_t(Mr), 1, int64_t(MrN));
                  h.Cn = two(uint32_t(N), Mr, int64_t(MrN) * pb.nk);
                  if (cs) {
                    h.hintA = nrmT + size_t(Mr) * (pb.j0 + pb.nb);
                    h.hintB = nrmP + size_t(N) * (c + pb.j0);
                  }
                  run_kt(h, "3e_ind_fock_kt_pair");
                }
              }
              const PairBlock& last = pblk[pgrp[gi].second - 1];
              GemmDesc h;  // L[(m, beta), A] += sum_p S[(m, beta), p] X2sP[A, p] over the group's pairs
              h.M = int(MrN);
              h.N = An;
              h.K = int(last.p0 + int64_t(last.nb) * last.nk - pbase);
              h.beta = gi == 0 ? 0.0 : 1.0;
              h.A = Kt;
              h.a32 = k32;
              h.Am = lin(1);
              h.Ak = lin(int64_t(MrN));
              h.B = X2sP + int64_t(DX) * pbase + A0;
              h.Bk = lin(DX);
              h.Bn = lin(1);
              h.C = L;
              h.Cm = lin(1);
              h.Cn = lin(int64_t(MrN));
              h.hintA = hintL;
              run_kt(h, "3e_ind_fock_l");
              progress(ctx, "same-spin 3e Fock, spin %d pass %d, A %d-%d of %d, pair group %d of %d (j %d-%d)", s,
                       pass, A0, A0 + An, DX, int(gi) + 1, int(pgrp.size()), pblk[pgrp[gi].first].j0,
                       last.j0 + last.nb);
            }
            for (int j0 = 0; !paired && j0 < a; j0 += nj) {
              const int jn = std::min(nj, a - j0);
              // K~[m, (beta, k) | j] = sum_B (PhiT_Y sY)[m, j, B] Phi_Y[beta, k, B], one product with
              // rows (m, j): the j of the block extend the ABS rows, so no tile pads Mr per j
              GemmDesc g;
              g.M = Mr * jn;
              g.N = N * o;
              g.K = DY;
              g.A = phitY + size_t(Mr) * j0;
              g.Am = lin(1);
              g.Ak = lin(int64_t(Mr) * a);
              g.B = phiY;
              g.Bk = lin(int64_t(N) * o);
              g.Bn = lin(1);
              g.C = Kt;
              g.c32 = k32;
              g.Cm = two(uint32_t(Mr), 1, MNo);
              g.Cn = lin(Mr);
              if (cs) {
                g.hintA = nrmT + size_t(Mr) * j0;
                g.hintB = nrmP;
              }
              run_kt(g, "3e_ind_fock_kt");
              GemmDesc h;  // L[(m, beta), A] += sum_{(k, j)} K~[(m, beta), k | j] X2sT[A, j + a k]
              h.M = int(MrN);
              h.N = An;
              h.K = o * jn;
              h.beta = j0 == 0 ? 0.0 : 1.0;
              h.A = Kt;
              h.a32 = k32;
              h.Am = lin(1);
              h.Ak = lin(int64_t(MrN));
              h.B = X2sT + int64_t(DX) * j0 + A0;
              h.Bk = two(uint32_t(o), int64_t(DX) * a, DX);
              h.Bn = lin(1);
              h.C = L;
              h.Cm = lin(1);
              h.Cn = lin(int64_t(MrN));
              h.hintA = hintL;
              run_kt(h, "3e_ind_fock_l");
              if (pass == 0 && A0 == 0) {
                // R[A, alpha | j] = sum_{(m, i)} Psi_W[m, c+i, A] K~[m, (alpha, c+i) | j], with jf
                // of the block's j as extra columns (alpha, j) of one product: wider than N alone,
                // while every column offset into the K~ block stays below 2^27 doubles
                int jf = 1;
                for (int d = jn; d > 1; --d)
                  if (jn % d == 0 && int64_t(Mr) * (N - 1) + MNo * (d - 1) < (int64_t(1) << 27)) {
                    jf = d;
                    break;
                  }
                GemmDesc r;
                r.M = DW;
                r.N = N * jf;
                r.K = Mr * a;
                r.batch = jn / jf;
                r.A = q.psiW + size_t(Mr) * c;
                r.Am = lin(int64_t(Mr) * o);
                r.Ak = lin(1);
                r.B = kt_at(MrN * size_t(c));
                r.b32 = k32;
                r.Bk = two(uint32_t(Mr), 1, int64_t(MrN));
                r.Bn = two(uint32_t(N), Mr, MNo);
                r.Bb = MNo * jf;
                r.C = R;
                r.Cm = lin(1);
                r.Cn = lin(DW);
                r.Cb = int64_t(DW) * N * jf;
                if (cs) {
                  // column (alpha, j) of the block: |K~[m, alpha, c+i | j]| <= w_j v_alpha
                  outer(csv, N, csw + j0, jn, 1.0, hintR, st);
                  r.hintB = hintR;
                  r.hBb = int64_t(N) * jf;
                }
                run_gemm(ctx, r, "3e_ind_fock_r");
                GemmDesc f;  // F += k1 sum_{(A, j)} sW_A Phi_W[beta, c+j, A] R[A, alpha | j]
                f.M = N;
                f.N = N;
                f.K = DW * jn;
                f.alpha = k1;
                f.beta = 1.0;
                f.A = q.phiW + size_t(N) * (c + j0);
                f.Am = lin(1);
                f.Ak = two(uint32_t(DW), int64_t(N) * o, N);
                f.ks = W->sign.p;
                f.Km = two(uint32_t(DW), 1, 0);
                f.B = R;
                f.Bk = two(uint32_t(DW), 1, int64_t(DW) * N);
                f.Bn = lin(DW);
                f.C = Fraw + NN * s;
                f.Cm = lin(1);
                f.Cn = lin(N);
                run_gemm(ctx, f, "3e_ind_fock_k", true);
              }
              progress(ctx, "same-spin 3e Fock, spin %d pass %d, A %d-%d of %d, j %d-%d of %d", s, pass, A0,
                       A0 + An, DX, j0, j0 + jn, a);
            }
            if (pass == 0 && mode != kAdjoint) {
              // E3_ind = e k1 sum_{m, i, A} Psi_W[m, c+i, A] sum_beta L[m, beta, A] Ca[beta, i]:
              // L holds sum_{j,k,B} PhiT_V[m, j, B] sV_B Phi_V[beta, k, B] W2s[(j, k), A], and Ca^T
              // over beta turns sV_B Phi_V[beta, k, B] into V2s[(i, k), B] (indirect_utils.cpp:19-133)
              GemmDesc g;
              g.M = Mr;
              g.N = a;
              g.K = N;
              g.batch = An;
              g.A = L;
              g.Am = lin(1);
              g.Ak = lin(Mr);
              g.Ab = int64_t(MrN);
              g.B = q.Co + size_t(N) * c;
              g.Bk = lin(1);
              g.Bn = lin(N);
              g.C = Lc;
              g.Cm = lin(1);
              g.Cn = lin(Mr);
              g.Cb = int64_t(Mr) * a;
              run_gemm(ctx, g, "3e_ind_e_lc", true);
              DotDesc dd;
              dd.x = q.psiW + size_t(Mr) * c + size_t(Mr) * o * A0;
              dd.y = Lc;
              dd.n = int64_t(Mr) * a * An;
              dd.ld = int64_t(Mr) * a;
              dd.xs = int64_t(Mr) * o;
              dd.ys = int64_t(Mr) * a;
              dot(dd, e * k1, Eacc + UW12G_E3_INDIRECT, 1, partial, st);
            }
            GemmDesc f;  // F += k1 sum_{m, A} Gt_X[m, alpha, A] L[m, beta, A]
            f.M = N;
            f.N = N;
            f.K = Mr * An;
            f.alpha = k1;
            f.beta = 1.0;
            f.A = gtp[pass] + MrN * A0;
            f.Am = lin(Mr);
            f.Ak = two(uint32_t(Mr), 1, int64_t(MrN));
            f.B = L;
            f.Bk = two(uint32_t(Mr), 1, int64_t(MrN));
            f.Bn = lin(Mr);
            f.C = Fraw + NN * s;
            f.Cm = lin(1);
            f.Cn = lin(N);
            run_gemm(ctx, f, "3e_ind_fock_f", true);
          }
        }
      }
    }
  }

  results();
}

}  // namespace

// ------------------------------------------------------------------ C ABI
extern "C" {

const char* uw12g_last_error(void) { return g_last_error.c_str(); }

int uw12g_context_create(int device, uw12g_context** out) {
  return guard([&] {
    require(out != nullptr, "null output handle");
    int count = 0;
    UW12G_HIP(hipGetDeviceCount(&count));
    require(device >= 0 && device < count, "invalid device index");
    std::unique_ptr<uw12g_context> c(new uw12g_context());
    c->device = device;
    UW12G_HIP(hipSetDevice(device));
    UW12G_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    size_t free_b = 0, total_b = 0;
    UW12G_HIP(hipMemGetInfo(&free_b, &total_b));
    c->block_cap = std::max<size_t>(size_t(1) << 24, std::min<size_t>(total_b / 16, size_t(16) << 30) / 8);
    c->block_cap_default = c->block_cap;
    c->tape_cap = c->tape_cap_default = total_b / 4 / sizeof(double);
    init_scratch(device);
    if (rocblas_create_handle(&c->rb) != rocblas_status_success) throw Error(-3, "rocblas_create_handle failed");
    rocblas_set_stream(c->rb, c->stream);
    *out = c.release();
  });
}

void uw12g_context_destroy(uw12g_context* ctx) {
  if (!ctx) return;
  if (ctx->live > 0) {  // deleted with its last integral / projector handle
    ctx->doomed = true;
    return;
  }
  (void)hipSetDevice(ctx->device);
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  delete ctx;
}

int uw12g_setup_base_integrals(uw12g_context* ctx, const double* X3, const double* X2, const double* X3_ri,
                               size_t n_ao, size_t n_df, size_t n_ri, uw12g_base_integrals** out) {
  return guard([&] {
    ThreeIndexSource src;
    src.X3 = X3;
    src.X3ri = X3_ri;
    require(X3 != nullptr, "null integral array");
    setup_base(ctx, src, X2, n_ao, n_df, n_ri, out);
  });
}

int uw12g_setup_base_integrals_device(uw12g_context* ctx, const double* X3_dev, const double* X2,
                                      const double* X3_ri_dev, size_t n_ao, size_t n_df, size_t n_ri,
                                      uw12g_base_integrals** out) {
  return guard([&] {
    ThreeIndexSource src;
    src.X3 = X3_dev;
    src.X3ri = X3_ri_dev;
    src.device = true;
    require(X3_dev != nullptr, "null integral array");
    setup_base(ctx, src, X2, n_ao, n_df, n_ri, out);
  });
}

int uw12g_setup_base_integrals_fn(uw12g_context* ctx, const double* X2, uw12g_three_index_fn three_index,
                                  uw12g_three_index_fn three_index_ri, void* user, const size_t* df_sizes,
                                  size_t n_shells, size_t n_ao, size_t n_df, size_t n_ri, uw12g_base_integrals** out) {
  return guard([&] {
    require(three_index != nullptr, "null three_index callback");
    require(df_sizes != nullptr && n_shells > 0, "no density-fitting shells");
    ThreeIndexSource src;
    src.fn = three_index;
    src.fn_ri = three_index_ri;
    src.user = user;
    src.shells.assign(df_sizes, df_sizes + n_shells);
    setup_base(ctx, src, X2, n_ao, n_df, n_ri, out);
  });
}

int uw12g_setup_base_integrals_gen(uw12g_context* ctx, const double* X2, uw12g_three_index_gen_fn gen, void* user,
                                   size_t n_ao, size_t n_df, size_t n_ri, int stored, uw12g_base_integrals** out) {
  return guard([&] {
    require(gen != nullptr, "null three-index generator");
    ThreeIndexSource src;
    src.gen = gen;
    src.user = user;
    src.stored = stored != 0;
    setup_base(ctx, src, X2, n_ao, n_df, n_ri, out);
  });
}

void uw12g_base_integrals_destroy(uw12g_base_integrals* ints) {
  if (!ints) return;
  uw12g_context* ctx = ints->ctx;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  if (ctx->tape.W == ints || ctx->tape.V == ints || ctx->tape.WV == ints) ctx->tape.id = 0;
  delete ints;
  release_child(ctx);
}

long uw12g_base_integrals_kept(const uw12g_base_integrals* ints) { return ints ? long(ints->D) : -1; }

int uw12g_setup_abs_projectors(uw12g_context* ctx, const double* S, size_t n_ao, size_t n_ri,
                               uw12g_abs_projectors** out) {
  return guard([&] {
    require(ctx && S && out, "null argument");
    require(n_ao > 0, "Number of ao orbitals is zero");
    const size_t Mz = n_ao + n_ri;
    require(Mz < 65536, "size out of range", -2);
    require(is_symmetric(S, Mz, 1e-10), "matrix must be symmetric");
    set_device(ctx);
    const int M = int(Mz);
    hipStream_t s = ctx->stream;
    std::unique_ptr<uw12g_abs_projectors> P(new uw12g_abs_projectors());
    P->ctx = ctx;
    P->N = int(n_ao);
    P->Nr = int(n_ri);
    P->M = M;
    DBuf Vd;
    Vd.get(Mz * Mz);
    UW12G_HIP(hipMemcpyAsync(Vd.p, S, sizeof(double) * Mz * Mz, hipMemcpyHostToDevice, s));
    std::vector<double> w = eig_device(ctx, Vd.p, M);
    double mx = 0.0;
    for (double x : w) mx = std::max(mx, std::fabs(x));
    const double thr = std::max(mx * 1e-8, 1e-6);  // ri_utils.hpp:49-50, linalg.hpp:1166-1169
    std::vector<int> keep;
    std::vector<double> ones, inv;
    for (int i = 0; i < M; ++i)
      if (std::fabs(w[i]) > thr) {
        keep.push_back(i);
        ones.push_back(1.0);
        inv.push_back(1.0 / w[i]);
      }
    const int K = int(keep.size());
    static std::atomic<uint64_t> next_id{0};
    P->id = ++next_id;
    P->r = K;
    P->Sp.get(Mz * Mz);
    P->U.get(Mz * std::max(K, 1));
    P->dinv.get(std::max(K, 1));
    if (K == 0) {
      fill(P->Sp.p, int64_t(Mz * Mz), 0.0, s);
    } else {
      int* keep_d = nullptr;
      UW12G_HIP(hipMalloc(&keep_d, sizeof(int) * K));
      DBuf on, Vs;
      on.get(K);
      Vs.get(Mz * K);
      UW12G_HIP(hipMemcpyAsync(keep_d, keep.data(), sizeof(int) * K, hipMemcpyHostToDevice, s));
      UW12G_HIP(hipMemcpyAsync(on.p, ones.data(), sizeof(double) * K, hipMemcpyHostToDevice, s));
      UW12G_HIP(hipMemcpyAsync(P->dinv.p, inv.data(), sizeof(double) * K, hipMemcpyHostToDevice, s));
      gather_scale_cols(Vd.p, M, keep_d, on.p, P->U.p, K, s);
      gather_scale_cols(Vd.p, M, keep_d, P->dinv.p, Vs.p, K, s);
      GemmDesc g;  // S+ = U diag(1/s) U^T
      g.M = M;
      g.N = M;
      g.K = K;
      g.A = Vs.p;
      g.Am = lin(1);
      g.Ak = lin(M);
      g.B = P->U.p;
      g.Bk = lin(M);
      g.Bn = lin(1);
      g.C = P->Sp.p;
      g.Cm = lin(1);
      g.Cn = lin(M);
      gemm_f64(g, s, &ctx->gemm_ws);
      UW12G_HIP(hipStreamSynchronize(s));
      (void)hipFree(keep_d);
    }
    UW12G_HIP(hipStreamSynchronize(s));
    adopt_child(ctx);
    *out = P.release();
  });
}

void uw12g_abs_projectors_destroy(uw12g_abs_projectors* abs) {
  if (!abs) return;
  uw12g_context* ctx = abs->ctx;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  if (ctx->tape.P == abs) ctx->tape.id = 0;
  delete abs;
  release_child(ctx);
}

int uw12g_abs_projectors_get(const uw12g_abs_projectors* abs, double* out) {
  return guard([&] {
    require(abs && out, "null argument");
    set_device(abs->ctx);
    UW12G_HIP(hipMemcpyAsync(out, abs->Sp.p, sizeof(double) * size_t(abs->M) * abs->M, hipMemcpyDeviceToHost,
                             abs->ctx->stream));
    UW12G_HIP(hipStreamSynchronize(abs->ctx->stream));
  });
}

namespace {
// Sets the contraction precision for one evaluation (the setup paths always run in FP64).
struct PrecScope {
  uw12g_context* c;
  PrecScope(uw12g_context* ctx, int prec) : c(ctx) { c->prec = prec; }
  ~PrecScope() { c->prec = 0; }
};
}  // namespace

int uw12g_form_fock(uw12g_context* ctx, const uw12g_base_integrals* W, const uw12g_base_integrals* V,
                    const uw12g_base_integrals* WV, const uw12g_abs_projectors* abs, const double* C, size_t n_ao,
                    size_t n_orb, size_t n_spin, const double* occ, const size_t* n_occ, const size_t* n_active,
                    int indirect_term, double scale_opp_spin, double scale_same_spin, uw12g_precision precision,
                    double* fock, double* terms_out, double* energy) {
  return guard([&] {
    require(precision >= UW12G_FP64 && precision <= UW12G_MIXED_FP16X3, "unsupported precision", -2);
    require(ctx != nullptr, "null context");
    double t6[6] = {0, 0, 0, 0, 0, 0};
    double* tp = terms_out ? terms_out : t6;
    PrecScope scope(ctx, int(precision));
    ++ctx->n_eval;
    evaluate(ctx, W, V, WV, abs, C, n_ao, n_orb, n_spin, occ, n_occ, n_active, indirect_term != 0,
             scale_opp_spin, scale_same_spin, fock, tp, energy);
  });
}

namespace {
double* carve_workspace(void* owner, size_t n) { return static_cast<uw12g_context*>(owner)->carve("gemm_ws", n); }

// The no-allocation mode for one call: named buffers and the GEMM workspace
// come from the device's static scratch array; profiling (events) is off.
struct ArenaScope {
  uw12g_context* c;
  bool prof_on;
  ArenaScope(uw12g_context* ctx, double* base) : c(ctx), prof_on(ctx->prof.on) {
    c->arena = base;
    c->arena_cap = UW12G_NOALLOC_SCRATCH_DOUBLES;
    c->arena_used = 0;
    c->arena_bufs.clear();
    c->arena_ws.carve = carve_workspace;
    c->arena_ws.owner = c;
    c->ws_active = &c->arena_ws;
    c->prof.on = false;
  }
  ~ArenaScope() {
    c->arena = nullptr;
    c->arena_bufs.clear();
    c->ws_active = &c->gemm_ws;
    c->prof.on = prof_on;
  }
};
}  // namespace

int uw12g_form_fock_noalloc(uw12g_context* ctx, const uw12g_base_integrals* W, const uw12g_base_integrals* V,
                            const uw12g_base_integrals* WV, const uw12g_abs_projectors* abs, const double* C,
                            size_t n_ao, size_t n_orb, size_t n_spin, const double* occ, const size_t* n_occ,
                            const size_t* n_active, int indirect_term, double scale_opp_spin, double scale_same_spin,
                            uw12g_precision precision, double* fock, double* terms_out, double* energy) {
  return guard([&] {
    require(precision >= UW12G_FP64 && precision <= UW12G_MIXED_FP16X3, "unsupported precision", -2);
    require(ctx != nullptr, "null context");
    double* base = scratch_address(ctx->device);
    require(base != nullptr, "no scratch array on this device", -3);
    double t6[6] = {0, 0, 0, 0, 0, 0};
    double* tp = terms_out ? terms_out : t6;
    PrecScope scope(ctx, int(precision));
    ArenaScope arena(ctx, base);
    // the context's named buffers (and so a forward call's tape) are not touched
    evaluate(ctx, W, V, WV, abs, C, n_ao, n_orb, n_spin, occ, n_occ, n_active, indirect_term != 0,
             scale_opp_spin, scale_same_spin, fock, tp, energy);
  });
}

int uw12g_forward(uw12g_context* ctx, const uw12g_base_integrals* W, const uw12g_base_integrals* V,
                  const uw12g_base_integrals* WV, const uw12g_abs_projectors* abs, const double* C, size_t n_ao,
                  size_t n_orb, size_t n_spin, const double* occ, const size_t* n_occ, const size_t* n_active,
                  int indirect_term, double scale_opp_spin, double scale_same_spin, uw12g_precision precision,
                  double* terms_out, double* energy, unsigned long long* tape) {
  return guard([&] {
    require(precision >= UW12G_FP64 && precision <= UW12G_MIXED_FP16X3, "unsupported precision", -2);
    require(ctx != nullptr && tape != nullptr, "null argument");
    require(C != nullptr && occ != nullptr && n_occ != nullptr && n_active != nullptr, "null input array");
    require(n_spin == 1 || n_spin == 2, "Number of spin channels must be 1 or 2");
    double t6[6] = {0, 0, 0, 0, 0, 0};
    double* tp = terms_out ? terms_out : t6;
    require(W && V && WV && abs, "null handle");
    PrecScope scope(ctx, int(precision));
    ++ctx->n_eval;
    ctx->tape.id = 0;  // no valid tape while the forward runs
    // The intermediates the adjoint would read: per spin Phi (AO rows), Psi (rotated ABS rows) and
    // (ik|A) of W and V, t and X, and diag(1/s) Psi_V.  Past the tape cap, or over non-stored
    // integrals, the forward keeps nothing and the adjoint recomputes them.
    const double DW = W->D, DV = V->D, Nao = double(n_ao), Mr = abs->r;
    double kept = 0.0;
    for (size_t s = 0; s < n_spin; ++s) {
      const double o = double(n_occ[s]), a = double(n_active[s]);
      kept += (Nao + Mr + a) * o * (DW + DV) + 2.0 * DW * DV + Mr * a * DV;
    }
    const bool recompute = !W->stored || !V->stored || kept > double(ctx->tape_cap);
    evaluate(ctx, W, V, WV, abs, C, n_ao, n_orb, n_spin, occ, n_occ, n_active, indirect_term != 0,
             scale_opp_spin, scale_same_spin, nullptr, tp, energy, recompute ? kFused : kForward);
    uw12g_context::Tape& t = ctx->tape;
    t.recompute = recompute;
    size_t n_occ_tot = 0;
    for (size_t s = 0; s < n_spin; ++s) n_occ_tot += n_occ[s];
    t.W = W;
    t.V = V;
    t.WV = WV;
    t.P = abs;
    t.C.assign(C, C + n_ao * n_orb * n_spin);
    t.occ.assign(occ, occ + n_occ_tot);
    t.n_occ.assign(n_occ, n_occ + n_spin);
    t.n_active.assign(n_active, n_active + n_spin);
    t.n_ao = n_ao;
    t.n_orb = n_orb;
    t.n_spin = n_spin;
    t.indirect = indirect_term != 0;
    t.k0 = scale_opp_spin;
    t.k1 = scale_same_spin;
    t.prec = ctx->prec;
    t.n_eval = ctx->n_eval;
    t.id = ++ctx->n_tapes;
    *tape = t.id;
  });
}

int uw12g_adjoint(uw12g_context* ctx, unsigned long long tape, double* fock) {
  return guard([&] {
    require(ctx != nullptr && fock != nullptr, "null argument");
    const uw12g_context::Tape& t = ctx->tape;
    require(tape != 0 && tape == t.id, "unknown or superseded forward tape", -4);
    require(t.recompute || t.n_eval == ctx->n_eval,
            "the forward intermediates were overwritten by a later evaluation", -4);
    PrecScope scope(ctx, t.prec);
    double t6[6];
    if (t.recompute) {
      // nothing kept: recompute every intermediate, i.e. the fused energy + Fock evaluation
      ++ctx->n_eval;
      evaluate(ctx, t.W, t.V, t.WV, t.P, t.C.data(), t.n_ao, t.n_orb, t.n_spin, t.occ.data(), t.n_occ.data(),
               t.n_active.data(), t.indirect, t.k0, t.k1, fock, t6, nullptr, kFused);
      return;
    }
    evaluate(ctx, t.W, t.V, t.WV, t.P, t.C.data(), t.n_ao, t.n_orb, t.n_spin, t.occ.data(), t.n_occ.data(),
             t.n_active.data(), t.indirect, t.k0, t.k1, fock, t6, nullptr, kAdjoint);
  });
}

static bool nearly_zero(double v) { return std::fabs(v) <= 4 * DBL_EPSILON; }  // utils.hpp:186-189

int uw12g_energy(uw12g_context* ctx, const uw12g_base_integrals* W, const uw12g_base_integrals* V,
                 const uw12g_base_integrals* WV, const uw12g_abs_projectors* abs, const double* C, size_t n_ao,
                 size_t n_orb, size_t n_spin, const double* occ, const size_t* n_occ, const size_t* n_active,
                 double scale_opp_spin, double scale_same_spin, double* energy) {
  return uw12g_form_fock(ctx, W, V, WV, abs, C, n_ao, n_orb, n_spin, occ, n_occ, n_active,
                         !nearly_zero(scale_same_spin), scale_opp_spin, scale_same_spin, UW12G_FP64, nullptr,
                         nullptr, energy);
}

int uw12g_fock(uw12g_context* ctx, double* fock, const uw12g_base_integrals* W, const uw12g_base_integrals* V,
               const uw12g_base_integrals* WV, const uw12g_abs_projectors* abs, const double* C, size_t n_ao,
               size_t n_orb, size_t n_spin, const double* occ, const size_t* n_occ, const size_t* n_active,
               double scale_opp_spin, double scale_same_spin, double* energy) {
  if (!fock) {
    g_last_error = "null fock buffer";
    return -1;
  }
  return uw12g_form_fock(ctx, W, V, WV, abs, C, n_ao, n_orb, n_spin, occ, n_occ, n_active,
                         !nearly_zero(scale_same_spin), scale_opp_spin, scale_same_spin, UW12G_FP64, fock, nullptr,
                         energy);
}

int uw12g_test_set_stream_cap(uw12g_context* ctx, size_t doubles) {
  return guard([&] {
    require(ctx != nullptr && doubles > 0, "invalid argument");
    ctx->stream_cap = doubles;
  });
}

int uw12g_test_set_tape_cap(uw12g_context* ctx, size_t doubles) {
  return guard([&] {
    require(ctx != nullptr, "null context");
    ctx->tape_cap = doubles > 0 ? doubles : ctx->tape_cap_default;
  });
}

int uw12g_test_set_block_cap(uw12g_context* ctx, size_t doubles) {
  return guard([&] {
    require(ctx != nullptr, "null context");
    ctx->block_cap = doubles > 0 ? doubles : ctx->block_cap_default;
  });
}

int uw12g_test_set_exact_stages(uw12g_context* ctx, const char* prefixes) {
  return guard([&] {
    require(ctx != nullptr, "null context");
    ctx->exact_extra.clear();
    std::string cur;
    for (const char* c = prefixes ? prefixes : ""; ; ++c) {
      if (*c == ',' || *c == '\0') {
        if (!cur.empty()) ctx->exact_extra.push_back(cur);
        cur.clear();
        if (*c == '\0') break;
      } else {
        cur += *c;
      }
    }
  });
}

int uw12g_context_set_verbosity(uw12g_context* ctx, int level) {
  return guard([&] {
    require(ctx != nullptr, "null context");
    ctx->verbosity = level;
  });
}

int uw12g_context_set_profiling(uw12g_context* ctx, int on) {
  return guard([&] {
    require(ctx != nullptr, "null context");
    ctx->prof.on = on != 0;
  });
}

int uw12g_context_profile(const uw12g_context* ctx, const char** names, double* ms, double* flops, double* bytes,
                          int cap) {
  if (!ctx) return -1;
  const int n = int(ctx->prof.recs.size());
  for (int i = 0; i < n && i < cap; ++i) {
    const ProfileRec& r = ctx->prof.recs[i];
    float t = 0.f;
    if (hipEventElapsedTime(&t, r.start, r.stop) != hipSuccess) t = -1.f;
    if (names) names[i] = r.name.c_str();
    if (ms) ms[i] = t;
    if (flops) flops[i] = r.flops;
    if (bytes) bytes[i] = r.bytes;
  }
  return n;
}

}  // extern "C"
