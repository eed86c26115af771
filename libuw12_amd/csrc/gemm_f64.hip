// FP64 strided tensor-contraction GEMM on CDNA4 matrix cores.
//
// One kernel serves every contraction of the UW12 evaluation: each operand
// index (rows, cols, batch) is addressed through a two-level map
// off(i) = (i mod n0) s0 + (i div n0) s1, so sums over compound indices such
// as (mu', A) or (j, k) run as a single K loop without materialising permuted
// copies.  Tiles are staged through LDS; the inner product uses
// v_mfma_f64_16x16x4_f64 with the B tile as the MFMA "A" operand so that lane
// index maps to the output row m (column-major outputs store coalesced).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <string>

#include "uw12g_internal.h"

namespace uw12g {

typedef double d4 __attribute__((ext_vector_type(4)));

namespace {

constexpr int BK = 16;
constexpr int NT = 256;

struct Params {
  GemmDesc d;
  int tiles_m, tiles_n, splits, kchunk;
  double* out;  // split-K partials (dense, M x N per split and batch) or nullptr
  int remap;    // 1: spread consecutive tiles over XCDs in contiguous chunks
};

// One 16-deep K slab per iteration, double-buffered in LDS (one barrier per
// slab).  Global loads for slab t+1 are issued before the MFMAs of slab t and
// consumed only when written to LDS after them, so their latency overlaps the
// matrix work: every load address is clamped to an in-bounds element (offset 0)
// when out of range and the predicate is applied at the LDS write.
#ifndef UW12G_LDS_PAD
#define UW12G_LDS_PAD 4
#endif
template <int BM, int BN, bool AKC, bool BKC, bool KS>
__global__ __launch_bounds__(NT) void gemm_f64_kernel(Params p) {
  constexpr int PAD = UW12G_LDS_PAD;
  constexpr int WM = BM / 2, WN = BN / 2;
  constexpr int MB = WM / 16, NB = WN / 16;
  constexpr int NA = BM * BK / NT;  // A elements loaded per thread and slab
  constexpr int NBL = BN * BK / NT;
  __shared__ double As[2][BK][BM + PAD];
  __shared__ double Bs[2][BK][BN + PAD];

  const GemmDesc& d = p.d;
  const int t = threadIdx.x;
  const int lane = t & 63, wave = t >> 6;
  int tile = blockIdx.x;
  if (p.remap) {  // blockIdx -> XCD is round-robin over 8 XCDs; give each XCD a contiguous range
    const int per = (p.tiles_m * p.tiles_n) >> 3;
    tile = (blockIdx.x & 7) * per + (blockIdx.x >> 3);
  }
  const int tm = tile % p.tiles_m, tn = tile / p.tiles_m;
  const int m0 = tm * BM, n0 = tn * BN;
  const int b = blockIdx.z;
  const int k_begin = blockIdx.y * p.kchunk;
  const int k_end = min(d.K, k_begin + p.kchunk);

  const double* __restrict__ Ap = d.A + d.Ab * b;
  const double* __restrict__ Bp = d.B + d.Bb * b;
  const double* __restrict__ Kp = KS ? d.ks + d.Kb * b : nullptr;

  // ---- per-thread load assignment (fixed over the K loop)
  // A: AKC -> k = t % BK, rows t/BK + (NT/BK) i ; else row t % BM, k = t/BM + (NT/BM) i.
  // In the second form every k a wave loads is wave-uniform (BM, BN >= 64), so
  // its offset, range test and scale factor are computed once per wave on the
  // scalar unit and the vector work per element is one address add.
  constexpr int NAM = AKC ? NA : 1;
  int64_t a_moff[NAM];
  uint32_t a_mok = 0;
  const int a_kk0 = AKC ? t % BK : __builtin_amdgcn_readfirstlane(t / BM);
#pragma unroll
  for (int i = 0; i < NAM; ++i) {
    const int m = m0 + (AKC ? t / BK + (NT / BK) * i : t % BM);
    const bool ok = m < d.M;
    a_mok |= uint32_t(ok) << i;
    a_moff[i] = ok ? d.Am.off(m) : 0;
  }
  constexpr int NBN = BKC ? NBL : 1;
  int64_t b_noff[NBN];
  uint32_t b_nok = 0;
  const int b_kk0 = BKC ? t % BK : __builtin_amdgcn_readfirstlane(t / BN);
#pragma unroll
  for (int i = 0; i < NBN; ++i) {
    const int n = n0 + (BKC ? t / BK + (NT / BK) * i : t % BN);
    const bool ok = n < d.N;
    b_nok |= uint32_t(ok) << i;
    b_noff[i] = ok ? d.Bn.off(n) : 0;
  }

  double ra[NA], rb[NBL], rs[AKC ? 1 : NA];
  uint32_t a_ok = 0, b_ok = 0;
  auto load = [&](int k0) {
    a_ok = 0;
    b_ok = 0;
    if constexpr (AKC) {
      const int k = k0 + a_kk0;
      const bool kok = k < k_end;
      const int64_t ko = kok ? d.Ak.off(k) : 0;
      if constexpr (KS) rs[0] = Kp[kok ? d.Km.off(k) : 0];
#pragma unroll
      for (int i = 0; i < NA; ++i) {
        const bool ok = kok && ((a_mok >> i) & 1u);
        a_ok |= uint32_t(ok) << i;
        ra[i] = Ap[ok ? a_moff[i] + ko : 0];
      }
    } else {
      // row m = 0 and column k = 0 exist, so a clamped offset stays in bounds
#pragma unroll
      for (int i = 0; i < NA; ++i) {
        const int k = k0 + a_kk0 + (NT / BM) * i;  // wave-uniform
        const bool kok = k < k_end;
        a_ok |= uint32_t(kok) << i;
        ra[i] = Ap[a_moff[0] + (kok ? d.Ak.off(k) : 0)];
        if constexpr (KS) rs[i] = Kp[kok ? d.Km.off(k) : 0];
      }
    }
    if constexpr (BKC) {
      const int k = k0 + b_kk0;
      const bool kok = k < k_end;
      const int64_t ko = kok ? d.Bk.off(k) : 0;
#pragma unroll
      for (int i = 0; i < NBL; ++i) {
        const bool ok = kok && ((b_nok >> i) & 1u);
        b_ok |= uint32_t(ok) << i;
        rb[i] = Bp[ok ? b_noff[i] + ko : 0];
      }
    } else {
#pragma unroll
      for (int i = 0; i < NBL; ++i) {
        const int k = k0 + b_kk0 + (NT / BN) * i;  // wave-uniform
        const bool kok = k < k_end;
        b_ok |= uint32_t(kok) << i;
        rb[i] = Bp[b_noff[0] + (kok ? d.Bk.off(k) : 0)];
      }
    }
  };
  // non-k-contiguous forms: a_ok / b_ok hold the uniform k tests, a_mok / b_nok the lane's row test
  const uint32_t a_lane = AKC ? ~0u : (a_mok & 1u ? ~0u : 0u);
  const uint32_t b_lane = BKC ? ~0u : (b_nok & 1u ? ~0u : 0u);
  auto store = [&](int buf) {
    const uint32_t am = a_ok & a_lane, bm = b_ok & b_lane;
#pragma unroll
    for (int i = 0; i < NA; ++i) {
      double v = ((am >> i) & 1u) ? ra[i] : 0.0;
      if constexpr (KS) v *= rs[AKC ? 0 : i];
      if constexpr (AKC) As[buf][a_kk0][t / BK + (NT / BK) * i] = v;
      else As[buf][a_kk0 + (NT / BM) * i][t % BM] = v;
    }
#pragma unroll
    for (int i = 0; i < NBL; ++i) {
      const double v = ((bm >> i) & 1u) ? rb[i] : 0.0;
      if constexpr (BKC) Bs[buf][b_kk0][t / BK + (NT / BK) * i] = v;
      else Bs[buf][b_kk0 + (NT / BN) * i][t % BN] = v;
    }
  };

  d4 acc[MB][NB];
#pragma unroll
  for (int i = 0; i < MB; ++i)
#pragma unroll
    for (int j = 0; j < NB; ++j) acc[i][j] = d4{0.0, 0.0, 0.0, 0.0};

  const int wm0 = (wave & 1) * WM, wn0 = (wave >> 1) * WN;
  const int lr = lane & 15, lk = lane >> 4;

  if (k_begin < k_end) {
    load(k_begin);
    int buf = 0;
    for (int k0 = k_begin; k0 < k_end; k0 += BK) {
      store(buf);
      __syncthreads();
      if (k0 + BK < k_end) load(k0 + BK);
#pragma unroll
      for (int ks = 0; ks < BK; ks += 4) {
        double af[MB], bf[NB];
#pragma unroll
        for (int i = 0; i < MB; ++i) af[i] = As[buf][ks + lk][wm0 + 16 * i + lr];
#pragma unroll
        for (int j = 0; j < NB; ++j) bf[j] = Bs[buf][ks + lk][wn0 + 16 * j + lr];
#pragma unroll
        for (int i = 0; i < MB; ++i)
#pragma unroll
          for (int j = 0; j < NB; ++j)
            acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(bf[j], af[i], acc[i][j], 0, 0, 0);
      }
      buf ^= 1;
    }
  }

  // ---- epilogue.  v_mfma_f64_16x16x4f64 leaves D[i][j] of its 16x16 block in
  // register i / 4 of lane j + 16 (i % 4) (one-row groups, four groups per
  // block), so lane (lr, lk) register r holds C[m = .. + lr][n = .. + lk + 4 r].
  if (p.out) {
    double* W = p.out + (int64_t(blockIdx.y) * d.batch + b) * int64_t(d.M) * d.N;
#pragma unroll
    for (int i = 0; i < MB; ++i) {
      const int m = m0 + wm0 + 16 * i + lr;
      if (m >= d.M) continue;
#pragma unroll
      for (int j = 0; j < NB; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int n = n0 + wn0 + 16 * j + lk + 4 * r;
          if (n < d.N) W[m + int64_t(d.M) * n] = d.alpha * acc[i][j][r];
        }
    }
    return;
  }
  double* Cp = d.C + d.Cb * b;
#pragma unroll
  for (int i = 0; i < MB; ++i) {
    const int m = m0 + wm0 + 16 * i + lr;
    if (m >= d.M) continue;
    const int64_t mo = d.Cm.off(m);
#pragma unroll
    for (int j = 0; j < NB; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int n = n0 + wn0 + 16 * j + lk + 4 * r;
        if (n >= d.N) continue;
        double* c = Cp + mo + d.Cn.off(n);
        const double v = d.alpha * acc[i][j][r];
        *c = d.beta == 0.0 ? v : v + d.beta * *c;
      }
  }
}

// Sum split-K partials and apply beta / the output map (fixed order: deterministic).
__global__ void splitk_reduce_kernel(GemmDesc d, const double* W, int splits) {
  const int64_t MN = int64_t(d.M) * d.N;
  const int b = blockIdx.y;
  for (int64_t e = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; e < MN; e += int64_t(gridDim.x) * blockDim.x) {
    double s = 0.0;
    for (int q = 0; q < splits; ++q) s += W[(int64_t(q) * d.batch + b) * MN + e];
    const int m = int(e % d.M), n = int(e / d.M);
    double* c = d.C + d.Cb * b + d.Cm.off(m) + d.Cn.off(n);
    *c = d.beta == 0.0 ? s : s + d.beta * *c;
  }
}

// Plain one-thread-per-output reference kernel (checks the MFMA kernel).
__global__ void gemm_naive_kernel(GemmDesc d) {
  const int64_t MN = int64_t(d.M) * d.N;
  const int b = blockIdx.y;
  for (int64_t e = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; e < MN; e += int64_t(gridDim.x) * blockDim.x) {
    const int m = int(e % d.M), n = int(e / d.M);
    double s = 0.0;
    for (int k = 0; k < d.K; ++k) {
      double a = d.A[d.Ab * b + d.Am.off(m) + d.Ak.off(k)];
      if (d.ks) a *= d.ks[d.Kb * b + d.Km.off(k)];
      s += a * d.B[d.Bb * b + d.Bk.off(k) + d.Bn.off(n)];
    }
    double* c = d.C + d.Cb * b + d.Cm.off(m) + d.Cn.off(n);
    *c = d.beta == 0.0 ? d.alpha * s : d.alpha * s + d.beta * *c;
  }
}

template <int BM, int BN>
std::string cfg_name(bool akc, bool bkc, bool ks) {
  return "gemm_f64_kernel<" + std::to_string(BM) + "," + std::to_string(BN) + "," + (akc ? "1" : "0") + "," +
         (bkc ? "1" : "0") + "," + (ks ? "1" : "0") + ">";
}

template <int BM, int BN, bool KS>
void launch_ks(const Params& p, bool akc, bool bkc, dim3 grid, hipStream_t s) {
  if (akc && bkc) hipLaunchKernelGGL((gemm_f64_kernel<BM, BN, true, true, KS>), grid, dim3(NT), 0, s, p);
  else if (akc) hipLaunchKernelGGL((gemm_f64_kernel<BM, BN, true, false, KS>), grid, dim3(NT), 0, s, p);
  else if (bkc) hipLaunchKernelGGL((gemm_f64_kernel<BM, BN, false, true, KS>), grid, dim3(NT), 0, s, p);
  else hipLaunchKernelGGL((gemm_f64_kernel<BM, BN, false, false, KS>), grid, dim3(NT), 0, s, p);
}

template <int BM, int BN>
void launch_cfg(const Params& p, bool akc, bool bkc, dim3 grid, hipStream_t s) {
  if (p.d.ks) launch_ks<BM, BN, true>(p, akc, bkc, grid, s);
  else launch_ks<BM, BN, false>(p, akc, bkc, grid, s);
}

// Unit stride on the low level of a map (contiguous runs along that index).
bool unit(const Map2& m) { return m.s0 == 1 && m.n0 >= 2; }

}  // namespace

int64_t map_max(const Map2& m, int64_t extent) {
  if (extent <= 0) return 0;
  const int64_t last = extent - 1;
  const int64_t q = last / m.n0, r = last % m.n0;
  int64_t best = r * m.s0 + q * m.s1;
  if (q >= 1) best = std::max(best, int64_t(m.n0 - 1) * m.s0 + (q - 1) * m.s1);
  return best;
}

static void check_desc(const GemmDesc& d) {
  if (d.K > 0) {
    check_extent(d.A, d.Ab * (d.batch - 1) + map_max(d.Am, d.M) + map_max(d.Ak, d.K), "gemm A");
    check_extent(d.B, d.Bb * (d.batch - 1) + map_max(d.Bk, d.K) + map_max(d.Bn, d.N), "gemm B");
    if (d.ks) check_extent(d.ks, d.Kb * (d.batch - 1) + map_max(d.Km, d.K), "gemm kscale");
  }
  check_extent(d.C, d.Cb * (d.batch - 1) + map_max(d.Cm, d.M) + map_max(d.Cn, d.N), "gemm C");
}

void gemm_f64(const GemmDesc& d, hipStream_t stream, Workspace* ws, Profiler* prof) {
  if (d.M <= 0 || d.N <= 0 || d.batch <= 0) return;
  check_desc(d);
  if (d.batch > 65535) throw Error(-2, "gemm batch exceeds 65535");
  if (d.K <= 0) {  // C <- beta C
    GemmDesc z = d;
    z.K = 0;
    gemm_f64_naive(z, stream);
    return;
  }
  // Tile shape: large tiles for large outputs, small for thin ones.
  int BM = d.M >= 96 ? 128 : 64;
  int BN = d.N >= 96 ? 128 : 64;
  // coalesce along k only when that index has the unit stride and the other does not
  const bool akc = unit(d.Ak) && !unit(d.Am);
  const bool bkc = unit(d.Bk) && !unit(d.Bn);
  Params p;
  p.d = d;
  p.tiles_m = (d.M + BM - 1) / BM;
  p.tiles_n = (d.N + BN - 1) / BN;
  const int64_t tiles = int64_t(p.tiles_m) * p.tiles_n;
  if (tiles > 0x7fffffff) throw Error(-2, "gemm grid too large");
  // Split K when the grid would leave most of the 256 CUs idle.
  int splits = 1;
  const int64_t wgs = tiles * d.batch;
  const int kt = (d.K + BK - 1) / BK;
  if (wgs < 512 && kt >= 16 && ws) {
    splits = int(std::min<int64_t>((1024 + wgs - 1) / wgs, kt / 8));
    splits = std::max(1, std::min(splits, 64));
    while (splits > 1 && int64_t(splits) * d.batch > 65535) --splits;
  }
  p.kchunk = ((kt + splits - 1) / splits) * BK;
  splits = (d.K + p.kchunk - 1) / p.kchunk;
  p.splits = splits;
  p.out = nullptr;
  if (splits > 1) p.out = ws->get(size_t(splits) * d.batch * size_t(d.M) * d.N, stream);
  p.remap = (tiles % 8 == 0 && tiles >= 64) ? 1 : 0;
  dim3 grid(unsigned(tiles), unsigned(splits), unsigned(d.batch));
  const bool timing = prof && prof->on;
  ProfileRec rec;
  if (timing) prof->begin(rec, stream);
  std::string kname;
  if (BM == 128 && BN == 128) {
    launch_cfg<128, 128>(p, akc, bkc, grid, stream);
    if (timing) kname = cfg_name<128, 128>(akc, bkc, d.ks != nullptr);
  } else if (BM == 128) {
    launch_cfg<128, 64>(p, akc, bkc, grid, stream);
    if (timing) kname = cfg_name<128, 64>(akc, bkc, d.ks != nullptr);
  } else if (BN == 128) {
    launch_cfg<64, 128>(p, akc, bkc, grid, stream);
    if (timing) kname = cfg_name<64, 128>(akc, bkc, d.ks != nullptr);
  } else {
    launch_cfg<64, 64>(p, akc, bkc, grid, stream);
    if (timing) kname = cfg_name<64, 64>(akc, bkc, d.ks != nullptr);
  }
  UW12G_HIP(hipGetLastError());
  if (timing) prof->end(rec, stream, kname, 2.0 * d.M * double(d.N) * d.K * d.batch, 0.0);
  if (splits > 1) {
    const int64_t MN = int64_t(d.M) * d.N;
    dim3 g(unsigned(std::min<int64_t>((MN + 255) / 256, 4096)), unsigned(d.batch));
    ProfileRec r2;
    if (timing) prof->begin(r2, stream);
    hipLaunchKernelGGL(splitk_reduce_kernel, g, dim3(256), 0, stream, d, p.out, splits);
    UW12G_HIP(hipGetLastError());
    if (timing) prof->end(r2, stream, "splitk_reduce_kernel", double(splits) * d.batch * MN, 8.0 * (splits + 1) * d.batch * MN);
  }
}

void gemm_f64_naive(const GemmDesc& d, hipStream_t stream) {
  if (d.M <= 0 || d.N <= 0 || d.batch <= 0) return;
  check_desc(d);
  const int64_t MN = int64_t(d.M) * d.N;
  dim3 g(unsigned(std::min<int64_t>((MN + 255) / 256, 4096)), unsigned(d.batch));
  hipLaunchKernelGGL(gemm_naive_kernel, g, dim3(256), 0, stream, d);
  UW12G_HIP(hipGetLastError());
}

}  // namespace uw12g
