// Internal declarations shared by the HIP translation units of libuw12g.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

namespace uw12g {

// Exact unsigned 32-bit division by a run-time constant d >= 1
// (round-up multiplier method, Granlund & Montgomery, PLDI 1994):
//   l = ceil(log2 d), m = floor(2^32 (2^l - d) / d) + 1,
//   x / d = (umulhi(x, m) + x) >> l   for all x < 2^32.
struct FastDiv {
  uint32_t d = 1, m = 0, s = 0;
};

inline FastDiv make_fastdiv(uint32_t d) {
  FastDiv f;
  f.d = d;
  uint32_t l = 0;
  while ((uint64_t(1) << l) < d) ++l;
  f.s = l;
  f.m = uint32_t(((uint64_t(1) << 32) * ((uint64_t(1) << l) - d)) / d + 1);
  return f;
}

__host__ __device__ inline uint32_t fd_div(uint32_t x, const FastDiv& f) {
#ifdef __HIP_DEVICE_COMPILE__
  const uint32_t hi = __umulhi(x, f.m);
#else
  const uint32_t hi = uint32_t((uint64_t(x) * f.m) >> 32);
#endif
  return uint32_t((uint64_t(hi) + x) >> f.s);
}

// Two-level index map: off(i) = (i mod n0) s0 + (i div n0) s1.  A one-level
// map off(i) = i s0 uses n0 = 2^32 - 1 (the quotient is 0 for every valid i),
// so evaluation is branch-free on the device.
struct Map2 {
  int64_t s0 = 0, s1 = 0;
  uint32_t n0 = 0xffffffffu;
  FastDiv fd = {0xffffffffu, 2u, 32u};
  __host__ __device__ inline int64_t off(uint32_t i) const {
    const uint32_t q = fd_div(i, fd);
    return int64_t(i - q * n0) * s0 + int64_t(q) * s1;
  }
};

inline Map2 lin(int64_t s) {
  Map2 m;
  m.s0 = s;
  return m;
}

inline Map2 two(uint32_t n0, int64_t s0, int64_t s1) {
  Map2 m;
  m.s0 = s0;
  m.s1 = s1;
  m.n0 = n0;
  m.fd = make_fastdiv(n0);
  return m;
}

// C(m,n) <- alpha sum_k A(m,k) ks(k) B(k,n) + beta C(m,n), for every batch b.
// A(m,k) = A[Ab*b + Am.off(m) + Ak.off(k)], and likewise for B, C, ks.
// Longest k range a reduced-precision (prec >= 1) contraction accumulates in FP32;
// longer sums are split into such ranges and the partial results added in FP64.
constexpr int kF32Chunk = 4096;

struct GemmDesc {
  int M = 0, N = 0, K = 0, batch = 1;
  double alpha = 1.0, beta = 0.0;
  const double* A = nullptr;
  Map2 Am, Ak;
  int64_t Ab = 0;
  const double* B = nullptr;
  Map2 Bk, Bn;
  int64_t Bb = 0;
  double* C = nullptr;
  Map2 Cm, Cn;
  int64_t Cb = 0;
  const double* ks = nullptr;
  Map2 Km;
  int64_t Kb = 0;
  // matrix-core arithmetic: 0 FP64, 1 FP32 (operands and chunk sums in FP32, FP64 storage),
  // 2 / 3 BF16 split products with two / three planes per operand, 4 FP16 split products
  // with two half planes of row-scaled operands (FP32 chunk sums)
  int prec = 0;
  // FP16 split form only (optional): upper bounds on max_k |A(m, k) ks(k)| at hintA[b hAb + m]
  // and on max_k |B(k, n)| at hintB[b hBb + n]; the row scales then come from them instead of
  // a pass over the operand
  const double* hintA = nullptr;
  int64_t hAb = 0;
  const double* hintB = nullptr;
  int64_t hBb = 0;
  // split forms over plane tensors only: A / B / C hold FP32 values (float storage, maps in
  // elements); the other paths refuse such descriptors
  int a32 = 0, b32 = 0, c32 = 0;
  // split forms over plane tensors only (optional): A already split by split_a_planes for a
  // descriptor with the same A, maps, M, K, k scale, hint and prec; its split is then skipped
  const void* planesA = nullptr;
};

// Allocation registry: every device buffer the library allocates is recorded
// so that each launch can check the extent of its operands on the host
// before it reaches the GPU.
void register_alloc(const void* p, size_t bytes);
void unregister_alloc(const void* p);
// Throws Error(-2) unless [base, base + (max_off + 1) doubles) lies inside one
// registered allocation; unregistered bases (caller-owned memory) pass.
void check_extent(const void* base, int64_t max_off, const char* what, int elem = 8);
// Pointer `off` elements of 4 (f32) or 8 bytes past p (FP32 storage behind a double* field).
inline const double* advance(const double* p, int64_t off, bool f32) {
  return reinterpret_cast<const double*>(reinterpret_cast<const char*>(p) + off * (f32 ? 4 : 8));
}
inline double* advance(double* p, int64_t off, bool f32) {
  return reinterpret_cast<double*>(reinterpret_cast<char*>(p) + off * (f32 ? 4 : 8));
}
// Largest offset a map produces over indices [0, extent).
int64_t map_max(const Map2& m, int64_t extent);

// Per-launch timing with HIP events on the launching stream (opt-in).
struct ProfileRec {
  std::string name;  // "stage|kernel"
  hipEvent_t start = nullptr, stop = nullptr;
  double flops = 0.0, bytes = 0.0;
};
struct Profiler {
  bool on = false;
  std::string stage;
  std::vector<ProfileRec> recs;
  void begin(ProfileRec& r, hipStream_t s);
  void end(ProfileRec& r, hipStream_t s, const std::string& kernel, double flops, double bytes);
  void clear();
  ~Profiler() { clear(); }
};

struct Workspace;
void gemm_f64(const GemmDesc& d, hipStream_t stream, Workspace* ws, Profiler* prof = nullptr);
// Operand A of a split-form product split once and kept (its planes, and for the FP16 form its
// row scales), for a product that recurs with the same A: plane_cache_bytes is the size of that
// buffer (0 when the product does not run over plane tensors), split_a_planes fills it, and the
// descriptor's planesA then points at it.
size_t plane_cache_bytes(const GemmDesc& d);
void split_a_planes(const GemmDesc& d, void* out, hipStream_t stream, Profiler* prof = nullptr);
// Test hook: a nonzero {BM, BN} forces that tile shape for every launch in the process.
extern int g_force_tile[2];
// Test hooks: g_bf16_planes 0 makes the BF16 split forms split their operands
// inside the GEMM kernel (no plane tensors) for every launch in the process;
// g_plane_bytes_max > 0 caps the plane workspace of one product (bytes).
extern int g_bf16_planes;
extern int64_t g_plane_bytes_max;
// Test hook: what the split forms keep of the rotated V integrals for the t~ dX product
// (1 planes or row maxima, 2 row maxima only, 0 nothing).
extern int g_kept_gt;
void gemm_f64_naive(const GemmDesc& d, hipStream_t stream);

// Device scratch owned by a context: grown on demand, reused across calls.
// With `carve` set the workspace never allocates: get() returns carve(n), a
// region of memory the caller owns (the no-allocation entry point's arena).
struct Workspace {
  void* ptr = nullptr;
  size_t bytes = 0;
  double* (*carve)(void* owner, size_t n_doubles) = nullptr;
  void* owner = nullptr;
  double* get(size_t n_doubles, hipStream_t stream);  // drains `stream` before growing
  ~Workspace();
};

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define UW12G_HIP(call)                                                                    \
  do {                                                                                     \
    hipError_t e_ = (call);                                                                \
    if (e_ != hipSuccess)                                                                  \
      throw ::uw12g::Error(-3, std::string(#call " failed: ") + hipGetErrorString(e_));   \
  } while (0)

}  // namespace uw12g
