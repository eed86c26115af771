// Test hooks (include/uw12g_testing.h).
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/uw12g_testing.h"
#include "uw12g_internal.h"

namespace uw12g {
void set_last_error(const std::string& msg);
}

using namespace uw12g;

namespace {
Map2 to_map(const uw12g_map2& m) { return m.n0 == 0 ? lin(m.s0) : two(m.n0, m.s0, m.s1); }  // n0 == 0: one level

struct Dev {
  double* p = nullptr;
  ~Dev() {
    if (p) {
      unregister_alloc(p);
      (void)hipFree(p);
    }
  }
  void up(const double* h, size_t n) {
    UW12G_HIP(hipMalloc(&p, sizeof(double) * (n ? n : 1)));
    register_alloc(p, sizeof(double) * (n ? n : 1));
    if (h && n) UW12G_HIP(hipMemcpy(p, h, sizeof(double) * n, hipMemcpyHostToDevice));
  }
  // FP32 storage: the host values rounded to float, in a buffer of n floats
  void up32(const double* h, size_t n) {
    std::vector<float> f(n ? n : 1, 0.0f);
    for (size_t i = 0; i < n; ++i) f[i] = float(h[i]);
    UW12G_HIP(hipMalloc(&p, sizeof(float) * f.size()));
    register_alloc(p, sizeof(float) * f.size());
    UW12G_HIP(hipMemcpy(p, f.data(), sizeof(float) * f.size(), hipMemcpyHostToDevice));
  }
  void down32(double* h, size_t n) const {
    std::vector<float> f(n ? n : 1);
    UW12G_HIP(hipMemcpy(f.data(), p, sizeof(float) * n, hipMemcpyDeviceToHost));
    for (size_t i = 0; i < n; ++i) h[i] = double(f[i]);
  }
};
}  // namespace

extern "C" int uw12g_test_set_bf16_planes(int on, long long max_bytes) {
  g_bf16_planes = on ? 1 : 0;
  g_plane_bytes_max = max_bytes > 0 ? int64_t(max_bytes) : 0;
  return 0;
}

extern "C" int uw12g_test_set_kept_gt(int mode) {
  if (mode < 0 || mode > 2) {
    set_last_error("mode must be 0, 1 or 2");
    return -1;
  }
  g_kept_gt = mode;
  return 0;
}

extern "C" long long uw12g_test_map2(uw12g_map2 m, unsigned i) { return (long long)to_map(m).off(i); }

extern "C" int uw12g_test_force_tile(int bm, int bn) {
  const bool ok = (bm == 0 && bn == 0) || (bm == 256 && bn == 256) || ((bm == 64 || bm == 128) && (bn == 64 || bn == 128));
  if (!ok) {
    set_last_error("tile must be 0 x 0 (automatic), 256 x 256 or 64 / 128 in each dimension");
    return -1;
  }
  g_force_tile[0] = bm;
  g_force_tile[1] = bn;
  return 0;
}

extern "C" int uw12g_test_gemm(int device, int M, int N, int K, int batch, double alpha, double beta,
                               const double* A, size_t nA, uw12g_map2 Am, uw12g_map2 Ak, long long Ab,
                               const double* B, size_t nB, uw12g_map2 Bk, uw12g_map2 Bn, long long Bb, double* C,
                               size_t nC, uw12g_map2 Cm, uw12g_map2 Cn, long long Cb, const double* ks, size_t nks,
                               uw12g_map2 Km, int naive) {
  try {
    // mode bits 3..5: A, B, C held in FP32 storage on the device (split forms only); bit 6:
    // A split once into a kept plane buffer (split_a_planes) that the product then reads
    const int store = (naive >> 3) & 7;
    const bool keep = (naive >> 6) & 1;
    naive &= 7;
    if (naive < 0 || naive > 5)
      throw Error(-1, "mode must be 0 .. 5 (plus 8 / 16 / 32 for FP32 storage, 64 for kept A planes)");
    if (store && naive == 1) throw Error(-2, "the one-thread-per-output kernel reads FP64 storage only");
    if (keep && naive < 3) throw Error(-2, "kept A planes: split forms only");
    UW12G_HIP(hipSetDevice(device));
    Dev a, b, c, k, kept;  // kept: A's plane buffer (mode bit 6), freed after the synchronisation below
    if (store & 1) a.up32(A, nA);
    else a.up(A, nA);
    if (store & 2) b.up32(B, nB);
    else b.up(B, nB);
    if (store & 4) c.up32(C, nC);
    else c.up(C, nC);
    if (ks) k.up(ks, nks);
    GemmDesc d;
    d.M = M;
    d.N = N;
    d.K = K;
    d.batch = batch;
    d.alpha = alpha;
    d.beta = beta;
    d.A = a.p;
    d.Am = to_map(Am);
    d.Ak = to_map(Ak);
    d.Ab = Ab;
    d.B = b.p;
    d.Bk = to_map(Bk);
    d.Bn = to_map(Bn);
    d.Bb = Bb;
    d.C = c.p;
    d.Cm = to_map(Cm);
    d.Cn = to_map(Cn);
    d.Cb = Cb;
    d.ks = ks ? k.p : nullptr;
    d.Km = to_map(Km);
    Workspace ws;
    if (naive == 1) {
      gemm_f64_naive(d, nullptr);
    } else {
      d.prec = naive >= 2 ? naive - 1 : 0;  // 2: FP32, 3: BF16x3, 4: BF16x6, 5: FP16x3
      d.a32 = store & 1;
      d.b32 = (store >> 1) & 1;
      d.c32 = (store >> 2) & 1;
      if (keep) {
        const size_t bytes = plane_cache_bytes(d);
        if (bytes == 0) throw Error(-2, "kept A planes: the product does not run over plane tensors");
        kept.up(nullptr, (bytes + 7) / 8);
        split_a_planes(d, kept.p, nullptr);
        d.planesA = kept.p;
      }
      gemm_f64(d, nullptr, &ws);
    }
    UW12G_HIP(hipDeviceSynchronize());
    if (store & 4) c.down32(C, nC);
    else UW12G_HIP(hipMemcpy(C, c.p, sizeof(double) * nC, hipMemcpyDeviceToHost));
    return 0;
  } catch (const Error& e) {
    set_last_error(e.what());
    return e.code;
  } catch (const std::exception& e) {
    set_last_error(e.what());
    return -1;
  }
}

extern "C" int uw12g_time_gemm(int device, int M, int N, int K, int batch, double alpha, double beta,
                               const double* A, size_t nA, uw12g_map2 Am, uw12g_map2 Ak, long long Ab,
                               const double* B, size_t nB, uw12g_map2 Bk, uw12g_map2 Bn, long long Bb, double* C,
                               size_t nC, uw12g_map2 Cm, uw12g_map2 Cn, long long Cb, const double* ks, size_t nks,
                               uw12g_map2 Km, int prec, int reps, double* ms) {
  try {
    // bits 3..5: A, B, C hold FP32 values (split forms only); the counts are then of floats
    const int store = prec >> 3;
    prec &= 7;
    if (prec < 0 || prec > 4)
      throw Error(-1, "prec must be 0 (FP64), 1 (FP32), 2 (BF16x3), 3 (BF16x6) or 4 (FP16x3), plus 8 / 16 / 32");
    if (M <= 0 || N <= 0 || K <= 0 || batch <= 0) throw Error(-1, "M, N, K and batch must be positive");
    // every element the descriptor addresses lies inside the caller's arrays (the
    // maps are checked non-negative, so the largest offset is the sum of the maxima)
    auto within = [](const uw12g_map2& m0, const uw12g_map2& m1, int e0, int e1, long long bs, int nb, size_t n,
                     const char* what) {
      const Map2 a = to_map(m0), b = to_map(m1);
      if (a.s0 < 0 || a.s1 < 0 || b.s0 < 0 || b.s1 < 0 || bs < 0) throw Error(-2, std::string(what) + ": negative stride");
      const int64_t last = int64_t(bs) * (nb - 1) + map_max(a, e0) + map_max(b, e1);
      if (uint64_t(last) >= uint64_t(n)) throw Error(-2, std::string(what) + ": descriptor reaches past the array");
    };
    within(Am, Ak, M, K, Ab, batch, nA, "time_gemm A");
    within(Bk, Bn, K, N, Bb, batch, nB, "time_gemm B");
    within(Cm, Cn, M, N, Cb, batch, nC, "time_gemm C");
    if (ks) within(Km, uw12g_map2{0, 0, 0}, K, 1, 0, 1, nks, "time_gemm ks");
    UW12G_HIP(hipSetDevice(device));
    GemmDesc d;
    d.prec = prec;
    d.M = M;
    d.N = N;
    d.K = K;
    d.batch = batch;
    d.alpha = alpha;
    d.beta = beta;
    d.A = A;
    d.Am = to_map(Am);
    d.Ak = to_map(Ak);
    d.Ab = Ab;
    d.B = B;
    d.Bk = to_map(Bk);
    d.Bn = to_map(Bn);
    d.Bb = Bb;
    d.C = C;
    d.Cm = to_map(Cm);
    d.Cn = to_map(Cn);
    d.Cb = Cb;
    d.ks = ks;
    d.Km = to_map(Km);
    d.a32 = store & 1;
    d.b32 = (store >> 1) & 1;
    d.c32 = (store >> 2) & 1;
    struct Stream {
      hipStream_t s = nullptr;
      ~Stream() {
        if (s) {
          (void)hipStreamSynchronize(s);
          (void)hipStreamDestroy(s);
        }
      }
    } guard;
    UW12G_HIP(hipStreamCreateWithFlags(&guard.s, hipStreamNonBlocking));
    Workspace ws;  // destroyed before the stream guard: the guard is declared first
    hipStream_t s = guard.s;
    gemm_f64(d, s, &ws);  // warm-up (also sizes the split-K workspace)
    hipEvent_t e0, e1;
    UW12G_HIP(hipEventCreate(&e0));
    UW12G_HIP(hipEventCreate(&e1));
    UW12G_HIP(hipEventRecord(e0, s));
    for (int r = 0; r < reps; ++r) gemm_f64(d, s, &ws);
    UW12G_HIP(hipEventRecord(e1, s));
    UW12G_HIP(hipEventSynchronize(e1));
    float t = 0.0f;
    UW12G_HIP(hipEventElapsedTime(&t, e0, e1));
    *ms = double(t) / (reps > 0 ? reps : 1);
    UW12G_HIP(hipEventDestroy(e0));
    UW12G_HIP(hipEventDestroy(e1));
    UW12G_HIP(hipStreamSynchronize(s));
    return 0;
  } catch (const Error& e) {
    set_last_error(e.what());
    return e.code;
  } catch (const std::exception& e) {
    set_last_error(e.what());
    return -1;
  }
}
