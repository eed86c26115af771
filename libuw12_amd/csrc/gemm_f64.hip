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
#include <map>
#include <mutex>
#include <numeric>
#include <string>
#include <type_traits>

#include "gemm_planes.h"
#include "kernels.h"
#include "uw12g_internal.h"

namespace uw12g {

typedef double d4 __attribute__((ext_vector_type(4)));

int g_force_tile[2] = {0, 0};
int g_bf16_planes = 1;
int64_t g_plane_bytes_max = 0;

namespace {

constexpr int BK = 16;

// LDS row padding (elements) of a tile loaded with k along the lanes: one
// wave writes 16 k-rows of one or two tile columns at once.  FP64: an odd row
// stride puts the 16 k-rows in 16 different bank pairs.  FP32: a stride of
// 18 mod 32 words spreads the 32 writes of a half-wave over all 32 banks and
// costs the MFMA reads (two k-rows of 16 consecutive columns) two collisions.
// Measured on TN 4096^3: FP64 64.2 -> 67.5 TF/s, FP32 109 -> 120 TF/s, and
// the SQ_LDS_BANK_CONFLICT count of these kernels was 4-8 per LDS instruction
// with the NN padding (4 doubles / 16 floats) used for the other tiles.
constexpr int kPadK64 = 1, kPadK32 = 18;

// Tile rows per group of the grouped tile order (Params::group).  On the K~
// product of n_ao = 320, n_df = 960, 80 occupied (425 x 200 tiles of 128 x 128,
// K = 960) the HBM fetch per launch was 80 GiB with the row-fastest order, for
// 0.6 GB of operands; grouped, the K~ products of an evaluation run at 124
// TF/s instead of 117.  Groups of 4, 8 and 16 rows were within 1% there; 4 was
// 5% faster on the pair sums S (680 x 40 tiles).
constexpr int kTileGroup = 4;

// Buffer-resource loads.  The resource base is a wave-uniform address (scalar
// registers); the per-lane part of an element address is a 32-bit byte offset
// fixed for the whole K loop and the per-element part a scalar byte offset, so
// a load costs no vector address arithmetic.
// A resource with num_records = 0 reads zero without touching memory.
__device__ inline __amdgpu_buffer_rsrc_t rsrc(const double* base, bool valid = true) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<double*>(base), (short)0, valid ? 0x7fffffff : 0, 0x00020000);
}
__device__ inline double ld_b(__amdgpu_buffer_rsrc_t r, uint32_t lane_off, int32_t uniform_off) {
  return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, lane_off, uniform_off, 0));
}

// Instruction order inside one slab of the main path, as scheduling groups:
// the first step's LDS reads, then every MFMA followed by its share of the
// slab's LDS writes, global loads, LDS reads and scalar work, spread evenly,
// so that each wave issues memory work in the shadow of its own MFMAs.
template <int NM, int NW, int NG, int NR0, int NR, int NS, int NV>
__device__ inline void slab_schedule() {
  __builtin_amdgcn_sched_group_barrier(0x100, NR0, 0);  // LDS reads of the first step
#pragma unroll
  for (int j = 0; j < NM; ++j) {
    __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);  // MFMA
    if ((j * NW) / NM != ((j + 1) * NW) / NM) __builtin_amdgcn_sched_group_barrier(0x200, 1, 0);  // LDS write
    if ((j * NG) / NM != ((j + 1) * NG) / NM) __builtin_amdgcn_sched_group_barrier(0x020, 1, 0);  // global load
    if ((j * NR) / NM != ((j + 1) * NR) / NM) __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);  // LDS read
    if ((j * NV) / NM != ((j + 1) * NV) / NM) __builtin_amdgcn_sched_group_barrier(0x002, 1, 0);  // VALU
    if ((j * NS) / NM != ((j + 1) * NS) / NM) __builtin_amdgcn_sched_group_barrier(0x004, 1, 0);  // SALU
  }
}

// K runs in 16-deep slabs, double-buffered in LDS with one barrier per slab.
// Each slab is split into 4 load groups, one per 4-deep MFMA step.  FP32:
// during step q of slab t a wave writes group q of slab t+1 (loaded one slab
// earlier) into the other LDS buffer and issues the group-q loads of slab t+2,
// so global loads, LDS writes and MFMAs interleave.  FP64 writes all groups of
// slab t+1 and issues all loads of slab t+2 at the top of slab t.  Measured on
// the UW12 shapes: the FP64 order is 3-12% faster for FP64 and 4-9% slower for
// FP32, which runs 3-4 waves per SIMD against FP64's 2.  On CDNA4 the FP64 MFMA shares issue
// with vector ALU work, so the main path (p.fast: affine k maps, 32-bit
// offsets) loads through buffer resources with no vector predicates: rows and
// columns past the matrix read row / column 0 (they only feed outputs that are
// not stored); k past the end reads zero through an empty resource where k is
// wave-uniform, and where k runs along the lanes every slab is whole (the host
// runs the k remainder separately) and the look-ahead past the last slab
// re-reads it.  Other descriptors take a general path with 64-bit offsets and
// clamped, zeroed loads.
// PREC selects the matrix-core arithmetic: 0 = FP64 (v_mfma_f64_16x16x4f64),
// 1 = FP32 (v_mfma_f32_16x16x4f32: operands rounded to FP32 at the LDS write,
// FP32 accumulation inside one launch chunk of at most kF32Chunk k values,
// chunks combined in FP64 by the split-K reduction).  Storage stays FP64.
template <int BM, int BN, bool AKC, bool BKC, bool KS, int PREC>
__global__ __launch_bounds__(NT) void gemm_f64_kernel(Params p) {
  using T = std::conditional_t<PREC == 0, double, float>;
  using AccT = std::conditional_t<PREC == 0, d4, f4>;
  // LDS tiles are [k][row].  Rows loaded along the lanes (NN) pad to 4 doubles /
  // 16 floats, so the MFMA reads of two k-rows fall in different bank halves;
  // rows loaded with k along the lanes (AKC / BKC) use kPadK64 / kPadK32.
  constexpr int PADA = PREC == 0 ? (AKC ? kPadK64 : 4) : (AKC ? kPadK32 : 16);
  constexpr int PADB = PREC == 0 ? (BKC ? kPadK64 : 4) : (BKC ? kPadK32 : 16);
  constexpr int WM = BM / 2, WN = BN / 2;
  constexpr int MB = WM / 16, NB = WN / 16;
  constexpr int NA = BM * BK / NT;  // A elements loaded per thread and slab
  constexpr int NBL = BN * BK / NT;
  constexpr int G = 4;              // load groups per slab
  constexpr int GA = NA / G, GB = NBL / G;
  __shared__ T As[2][BK][BM + PADA];
  __shared__ T Bs[2][BK][BN + PADB];
  using I0 = std::integral_constant<int, 0>;
  using I1 = std::integral_constant<int, 1>;
  using I2 = std::integral_constant<int, 2>;
  using I3 = std::integral_constant<int, 3>;
  const GemmDesc& d = p.d;
  const int t = threadIdx.x;
  const int lane = t & 63, wave = t >> 6;
  // Work item w = tile + tiles (split + splits b).  Workgroups go to the 8 XCDs
  // round-robin by blockIdx; giving each XCD a contiguous range of w puts the
  // tiles of one (k chunk, batch) -- which share its A and B panels -- on one
  // XCD, so each panel is read from HBM into that XCD's L2 once.
  int64_t w = blockIdx.x;
  if (p.per > 0) {
    w = int64_t(blockIdx.x & 7) * p.per + (blockIdx.x >> 3);
    if (w >= p.work) return;  // tail of the last XCD's range (whole workgroup, before any barrier)
  }
  const int tiles = p.tiles_m * p.tiles_n;
  const int tile = int(w % tiles);
  const int64_t q = w / tiles;
  const int split = int(q % p.splits);
  const int b = int(q / p.splits);
  // Tile order.  Tall outputs with few tile columns: column index fastest, so
  // the tiles that share an A panel run together and the panel is read from
  // HBM once.  Many tile rows and columns: groups of p.group tile rows (every
  // tile column), row index fastest inside a group, so the workgroups an XCD
  // runs at once cover a compact block of tiles and share its A and B panels
  // in L2 (row-fastest over all rows brings every A panel from HBM once per
  // tile column).  Otherwise row index fastest.
  int tm, tn;
  if (p.nfast) {
    tm = tile / p.tiles_n;
    tn = tile % p.tiles_n;
  } else if (p.group > 1) {
    const int per_group = p.group * p.tiles_n;
    const int g = tile / per_group, r = tile - g * per_group;
    const int first = g * p.group;
    const int gs = min(p.tiles_m - first, p.group);
    tm = first + r % gs;
    tn = r / gs;
  } else {
    tm = tile % p.tiles_m;
    tn = tile / p.tiles_m;
  }
  const int m0 = tm * BM, n0 = tn * BN;
  const int k_begin = p.kbase + split * p.kchunk;
  const int k_end = min(d.K, k_begin + p.kchunk);

  const double* __restrict__ Ap = d.A + d.Ab * b;
  const double* __restrict__ Bp = d.B + d.Bb * b;
  const double* __restrict__ Kp = KS ? d.ks + d.Kb * b : nullptr;

  // ---- per-thread load assignment (fixed over the K loop)
  // A: AKC -> k = t % BK, rows t/BK + (NT/BK) i ; else row t % BM, k = t/BM + (NT/BM) i.
  // In the second form every k a wave loads is wave-uniform (BM, BN >= 64).
  // Rows past M use row 0's offset.
  constexpr int NAM = AKC ? NA : 1;
  constexpr int NBN = BKC ? NBL : 1;
  const int a_kk0 = AKC ? t % BK : __builtin_amdgcn_readfirstlane(t / BM);
  const int b_kk0 = BKC ? t % BK : __builtin_amdgcn_readfirstlane(t / BN);
  auto a_row = [&](int i) { return m0 + (AKC ? t / BK + (NT / BK) * i : t % BM); };
  auto b_col = [&](int i) { return n0 + (BKC ? t / BK + (NT / BK) * i : t % BN); };

  double ra[NA], rb[NBL], rs[AKC ? G : NA];

  auto store = [&](auto qc, int buf) {
    constexpr int q = decltype(qc)::value;
#pragma unroll
    for (int i = q * GA; i < (q + 1) * GA; ++i) {
      double v = ra[i];
      if constexpr (KS) v *= rs[AKC ? q : i];
      if constexpr (AKC) As[buf][a_kk0][t / BK + (NT / BK) * i] = T(v);
      else As[buf][a_kk0 + (NT / BM) * i][t % BM] = T(v);
    }
#pragma unroll
    for (int i = q * GB; i < (q + 1) * GB; ++i) {
      if constexpr (BKC) Bs[buf][b_kk0][t / BK + (NT / BK) * i] = T(rb[i]);
      else Bs[buf][b_kk0 + (NT / BN) * i][t % BN] = T(rb[i]);
    }
  };

  AccT acc[MB][NB];
#pragma unroll
  for (int i = 0; i < MB; ++i)
#pragma unroll
    for (int j = 0; j < NB; ++j) acc[i][j] = AccT{0, 0, 0, 0};

  const int wm0 = (wave & 1) * WM, wn0 = (wave >> 1) * WN;
  const int lr = lane & 15, lk = lane >> 4;

  auto kstep = [&](auto qc, int buf) {
    constexpr int q = decltype(qc)::value;
    T af[MB], bf[NB];
#pragma unroll
    for (int i = 0; i < MB; ++i) af[i] = As[buf][4 * q + lk][wm0 + 16 * i + lr];
#pragma unroll
    for (int j = 0; j < NB; ++j) bf[j] = Bs[buf][4 * q + lk][wn0 + 16 * j + lr];
#pragma unroll
    for (int i = 0; i < MB; ++i)
#pragma unroll
      for (int j = 0; j < NB; ++j) {
        if constexpr (PREC == 0)
          acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(bf[j], af[i], acc[i][j], 0, 0, 0);
        else
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(bf[j], af[i], acc[i][j], 0, 0, 0);
      }
  };
  // Slab pipeline over [k_begin, k_end) with a group loader ld(q, k0); k0 of
  // a look-ahead slab may lie past the end (the loader clamps or masks it).
  auto pipeline = [&](auto fastc, auto&& ld) {
    ld(I0{}, k_begin);
    ld(I1{}, k_begin);
    ld(I2{}, k_begin);
    ld(I3{}, k_begin);
    store(I0{}, 0);
    store(I1{}, 0);
    store(I2{}, 0);
    store(I3{}, 0);
    __syncthreads();
    ld(I0{}, k_begin + BK);
    ld(I1{}, k_begin + BK);
    ld(I2{}, k_begin + BK);
    ld(I3{}, k_begin + BK);
    const int nslab = (k_end - k_begin + BK - 1) / BK;
    for (int s = 0; s < nslab; ++s) {
      // MFMAs from buffer cur; the other buffer receives slab s+1 (loaded one
      // slab earlier) and the registers slab s+2
      const int cur = s & 1, nx = cur ^ 1;
      const int k2 = k_begin + (s + 2) * BK;
      if constexpr (PREC == 0) {
        // FP64 (2 waves per SIMD): all LDS writes and loads first, so a load
        // has a whole slab of MFMAs to land before its LDS write; then the
        // four steps, each step's LDS reads inside the previous step's MFMAs
        store(I0{}, nx);
        store(I1{}, nx);
        store(I2{}, nx);
        store(I3{}, nx);
        ld(I0{}, k2);
        ld(I1{}, k2);
        ld(I2{}, k2);
        ld(I3{}, k2);
        __builtin_amdgcn_sched_barrier(0);
        // the MFMA steps at raised wave priority, so the SIMD's other wave issues
        // its loads and LDS writes in their shadow (FP64: +1-5% on the UW12 shapes)
        __builtin_amdgcn_s_setprio(1);
        kstep(I0{}, cur);
        kstep(I1{}, cur);
        kstep(I2{}, cur);
        kstep(I3{}, cur);
        __builtin_amdgcn_sched_group_barrier(0x100, (MB + NB) / 2, 0);
#pragma unroll
        for (int qq = 0; qq < 4; ++qq) {
          __builtin_amdgcn_sched_group_barrier(0x008, MB * NB / 2, 0);
          if (qq < 3) __builtin_amdgcn_sched_group_barrier(0x100, (MB + NB) / 2, 0);
          __builtin_amdgcn_sched_group_barrier(0x008, MB * NB / 2, 0);
        }
        __builtin_amdgcn_sched_barrier(0);
        __builtin_amdgcn_s_setprio(0);
        __syncthreads();
        continue;
      }
      // FP32 (3-4 waves per SIMD): loads, LDS writes and MFMAs interleaved
      // through the slab
      kstep(I0{}, cur);
      store(I0{}, nx);
      ld(I0{}, k2);
      kstep(I1{}, cur);
      store(I1{}, nx);
      ld(I1{}, k2);
      kstep(I2{}, cur);
      store(I2{}, nx);
      ld(I2{}, k2);
      kstep(I3{}, cur);
      store(I3{}, nx);
      ld(I3{}, k2);
      if constexpr (decltype(fastc)::value) {
        // per slab: 4 MB NB MFMAs, NA + NBL writes and loads, 2 (MB + NB) LDS
        // reads (as read2), about 4 scalar ops per MFMA
        slab_schedule<4 * MB * NB, NA + NBL, NA + NBL + (KS ? (AKC ? G : NA) : 0), (MB + NB) / 2,
                      3 * (MB + NB) / 2, 4 * MB * NB, KS ? NA : 0>();
      }
      __builtin_amdgcn_sched_barrier(0);
      __syncthreads();
    }
  };

  if (k_begin < k_end) {
    if (p.fast) {
      uint32_t a_lo[NAM], b_lo[NBN];  // lane byte offsets from the slab's base
#pragma unroll
      for (int i = 0; i < NAM; ++i) {
        const int m = a_row(i);
        a_lo[i] = uint32_t(8 * ((m < d.M ? d.Am.off(m) : 0) + (AKC ? int64_t(a_kk0) * d.Ak.s0 : 0)));
      }
#pragma unroll
      for (int i = 0; i < NBN; ++i) {
        const int n = b_col(i);
        b_lo[i] = uint32_t(8 * ((n < d.N ? d.Bn.off(n) : 0) + (BKC ? int64_t(b_kk0) * d.Bk.s0 : 0)));
      }
      const uint32_t k_lo = uint32_t(8 * (AKC ? int64_t(a_kk0) * d.Km.s0 : 0));
      const int32_t a_ks = int32_t(8 * (NT / BM) * d.Ak.s0), b_ks = int32_t(8 * (NT / BN) * d.Bk.s0);
      const int64_t k_kstep = int64_t(NT / BM) * d.Km.s0;
      // slab k0 is affine (off(k0 + dk) = off(k0) + dk s0), so one offset per slab
      const int k_last = k_end - BK;
      pipeline(std::true_type{}, [&](auto qc, int k0) {
        constexpr int q = decltype(qc)::value;
        if constexpr (AKC) {
          const int kc = min(k0, k_last);  // look-ahead past the end re-reads the last slab
          const auto r = rsrc(Ap + d.Ak.off(kc));
          if constexpr (KS) rs[q] = ld_b(rsrc(Kp + d.Km.off(kc)), k_lo, 0);
#pragma unroll
          for (int i = q * GA; i < (q + 1) * GA; ++i) ra[i] = ld_b(r, a_lo[i], 0);
        } else {
          const int kb = k0 + a_kk0;  // wave-uniform
          const int64_t kbo = KS ? d.Km.off(kb) : 0;
#pragma unroll
          for (int i = q * GA; i < (q + 1) * GA; ++i) {
            const bool kok = kb + (NT / BM) * i < k_end;
            ra[i] = ld_b(rsrc(Ap + d.Ak.off(kb), kok), a_lo[0], i * a_ks);
            if constexpr (KS) rs[i] = Kp[kok ? kbo + i * k_kstep : 0];
          }
        }
        if constexpr (BKC) {
          const auto r = rsrc(Bp + d.Bk.off(min(k0, k_last)));
#pragma unroll
          for (int i = q * GB; i < (q + 1) * GB; ++i) rb[i] = ld_b(r, b_lo[i], 0);
        } else {
          const int kb = k0 + b_kk0;
#pragma unroll
          for (int i = q * GB; i < (q + 1) * GB; ++i) {
            const bool kok = kb + (NT / BN) * i < k_end;
            rb[i] = ld_b(rsrc(Bp + d.Bk.off(kb), kok), b_lo[0], i * b_ks);
          }
        }
      });
    } else {
      // general path: 64-bit offsets, clamped in-bounds addresses (row / column
      // 0, k = 0) and zeroed values for k past the end
      int64_t a_moff[NAM], b_noff[NBN];
#pragma unroll
      for (int i = 0; i < NAM; ++i) {
        const int m = a_row(i);
        a_moff[i] = m < d.M ? d.Am.off(m) : 0;
      }
#pragma unroll
      for (int i = 0; i < NBN; ++i) {
        const int n = b_col(i);
        b_noff[i] = n < d.N ? d.Bn.off(n) : 0;
      }
      pipeline(std::false_type{}, [&](auto qc, int k0) {
        constexpr int q = decltype(qc)::value;
        if constexpr (AKC) {
          const int k = k0 + a_kk0;
          const bool kok = k < k_end;
          const int64_t ko = kok ? d.Ak.off(k) : 0;
          if constexpr (KS) rs[q] = Kp[kok ? d.Km.off(k) : 0];
#pragma unroll
          for (int i = q * GA; i < (q + 1) * GA; ++i) {
            const double v = Ap[a_moff[i] + ko];
            ra[i] = kok ? v : 0.0;
          }
        } else {
#pragma unroll
          for (int i = q * GA; i < (q + 1) * GA; ++i) {
            const int k = k0 + a_kk0 + (NT / BM) * i;  // wave-uniform
            const bool kok = k < k_end;
            const double v = Ap[a_moff[0] + (kok ? d.Ak.off(k) : 0)];
            ra[i] = kok ? v : 0.0;
            if constexpr (KS) rs[i] = Kp[kok ? d.Km.off(k) : 0];
          }
        }
        if constexpr (BKC) {
          const int k = k0 + b_kk0;
          const bool kok = k < k_end;
          const int64_t ko = kok ? d.Bk.off(k) : 0;
#pragma unroll
          for (int i = q * GB; i < (q + 1) * GB; ++i) {
            const double v = Bp[b_noff[i] + ko];
            rb[i] = kok ? v : 0.0;
          }
        } else {
#pragma unroll
          for (int i = q * GB; i < (q + 1) * GB; ++i) {
            const int k = k0 + b_kk0 + (NT / BN) * i;  // wave-uniform
            const bool kok = k < k_end;
            const double v = Bp[b_noff[0] + (kok ? d.Bk.off(k) : 0)];
            rb[i] = kok ? v : 0.0;
          }
        }
      });
    }
  }

  // ---- epilogue.  v_mfma_f64_16x16x4f64 leaves D[i][j] of its 16x16 block in
  // register i / 4 of lane j + 16 (i % 4) (one-row groups, four groups per
  // block), so lane (lr, lk) register r holds C[m = .. + lr][n = .. + lk + 4 r];
  // v_mfma_f32_16x16x4f32 uses four-row groups: register i % 4 of lane
  // j + 16 (i / 4), i.e. C[m = .. + lr][n = .. + 4 lk + r].
  auto ncol = [&](int j, int r) { return n0 + wn0 + 16 * j + (PREC == 0 ? lk + 4 * r : 4 * lk + r); };
  if (p.out) {
    double* W = p.out + (int64_t(split) * d.batch + b) * int64_t(d.M) * d.N;
#pragma unroll
    for (int i = 0; i < MB; ++i) {
      const int m = m0 + wm0 + 16 * i + lr;
      if (m >= d.M) continue;
#pragma unroll
      for (int j = 0; j < NB; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int n = ncol(j, r);
          if (n < d.N) W[m + int64_t(d.M) * n] = d.alpha * double(acc[i][j][r]);
        }
    }
    return;
  }
  double* Cp = d.C + d.Cb * b;
  if constexpr (PREC == 0) {
    if (d.beta != 0.0) {
      // read-modify-write: the old values of a row group are all loaded before
      // any of its stores, so a wave waits on memory once per row group
#pragma unroll
      for (int i = 0; i < MB; ++i) {
        const int m = m0 + wm0 + 16 * i + lr;
        if (m >= d.M) continue;
        const int64_t mo = d.Cm.off(m);
        int64_t off[NB][4];
        double old[NB][4];
#pragma unroll
        for (int j = 0; j < NB; ++j)
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int n = ncol(j, r);
            off[j][r] = n < d.N ? mo + d.Cn.off(n) : -1;
            old[j][r] = off[j][r] >= 0 ? Cp[off[j][r]] : 0.0;
          }
#pragma unroll
        for (int j = 0; j < NB; ++j)
#pragma unroll
          for (int r = 0; r < 4; ++r)
            if (off[j][r] >= 0) Cp[off[j][r]] = d.alpha * double(acc[i][j][r]) + d.beta * old[j][r];
      }
      return;
    }
  }
#pragma unroll
  for (int i = 0; i < MB; ++i) {
    const int m = m0 + wm0 + 16 * i + lr;
    if (m >= d.M) continue;
    const int64_t mo = d.Cm.off(m);
#pragma unroll
    for (int j = 0; j < NB; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int n = ncol(j, r);
        if (n >= d.N) continue;
        double* c = Cp + mo + d.Cn.off(n);
        const double v = d.alpha * double(acc[i][j][r]);
        *c = d.beta == 0.0 ? v : v + d.beta * *c;
      }
  }
}

// ---------------------------------------------------------------------------
// BF16 split-product forms (prec 2 and 3).  Each FP64 operand is split into
// bfloat16 planes -- prec 2: a ~ h + l, prec 3: a ~ h + m + l, each plane the
// round-to-nearest bfloat16 of what the planes before it leave (Henry, Tang &
// Heinecke, "Leveraging the bfloat16 artificial intelligence datatype for
// higher-precision computations", ARITH 2019) -- and the product runs as the
// plane products of order at most one (prec 2, three products: hh + hl + lh)
// or two (prec 3, six products: + hm + mh + mm, with h, m, l in place of h, l)
// on v_mfma_f32_16x16x32_bf16, accumulated in FP32 inside one launch chunk and
// combined in FP64 across chunks like the FP32 form.  Per-product relative
// error: about 3 * 2^-16 (prec 2), about 2^-22 (prec 3).
//
// K runs in 32-deep slabs (one MFMA k step).  Each thread owns octets of 8
// consecutive k of one row (column): one 16-byte LDS write per plane and
// octet, and one 16-byte LDS read fills an MFMA operand.  The LDS tile is
// single-buffered: while the MFMAs of slab s run, the wave splits the FP64
// values of slab s + 1 (loaded during slab s - 1) into packed plane words in
// registers; after a barrier the words go to LDS and the loads of slab s + 2
// are issued.
constexpr int kRowWords = BKB / 2 + 4;  // plane row stride in 32-bit words: 80 bytes, 16-byte aligned

typedef float fl2 __attribute__((ext_vector_type(2)));
typedef __bf16 bf2 __attribute__((ext_vector_type(2)));

// Round-to-nearest bfloat16 of two FP32 values, packed (first in the low half).
__device__ inline uint32_t pack_bf16(float a, float b) {
  return __builtin_bit_cast(uint32_t, __builtin_convertvector(fl2{a, b}, bf2));
}
__device__ inline float lo_f(uint32_t w) { return __builtin_bit_cast(float, w << 16); }
__device__ inline float hi_f(uint32_t w) { return __builtin_bit_cast(float, w & 0xffff0000u); }

// Plane words of two consecutive k values v0, v1.  Two planes: h from the FP32
// value, l from the exact FP32 residual (the FP32 rounding of v, 2^-24, is far
// below the form's 2^-16).  Three planes: the first residual is formed in FP64
// (v - h is exact there) and rounded to FP32 (2^-24 of a residual of at most
// 2^-8 |v|), the second residual exactly in FP32.
template <int NP>
__device__ inline void split_pair(double v0, double v1, uint32_t (&w)[NP]) {
  const float f0 = float(v0), f1 = float(v1);
  const uint32_t h = pack_bf16(f0, f1);
  w[0] = h;
  if constexpr (NP == 2) {
    w[1] = pack_bf16(f0 - lo_f(h), f1 - hi_f(h));
  } else {
    const float r0 = float(v0 - double(lo_f(h))), r1 = float(v1 - double(hi_f(h)));
    const uint32_t m = pack_bf16(r0, r1);
    w[1] = m;
    w[2] = pack_bf16(r0 - lo_f(m), r1 - hi_f(m));
  }
}

// FP16 split form (prec 4): each operand row is scaled by a power of two that
// puts its largest magnitude in [2^14, 2^15) (row_scale_kernel), then split
// into two IEEE half planes a s ~ h + l (h the round-to-nearest half of a s,
// l that of the residual, formed exactly in FP64).  With 11 significant bits
// per plane the three products hh + hl + lh on v_mfma_f32_16x16x32_f16 carry
// a relative error of about 3 * 2^-22 per product (against 3 * 2^-16 for two
// bfloat16 planes) at the same matrix-core cost; the output is scaled back
// by 1 / (s_row s_col) in the epilogue (exact).  Elements more than 2^18 below
// their row's maximum lose low bits of l to the half subnormal range: their
// absolute error stays below 2^-40 of the row maximum.
typedef _Float16 h2 __attribute__((ext_vector_type(2)));

__device__ inline uint32_t pack_f16(float a, float b) {
  return __builtin_bit_cast(uint32_t, __builtin_convertvector(fl2{a, b}, h2));
}
__device__ inline double f16_lo(uint32_t w) { return double(__builtin_bit_cast(_Float16, uint16_t(w & 0xffffu))); }
__device__ inline double f16_hi(uint32_t w) { return double(__builtin_bit_cast(_Float16, uint16_t(w >> 16))); }

// Plane words of two consecutive (already scaled) k values: h = half(v) (via
// FP32: the double rounding leaves |v - h| within one half ulp of h, and
// v - h is exact in FP64), l = half(v - h).
__device__ inline void split_pair_f16(double v0, double v1, uint32_t (&w)[2]) {
  const uint32_t h = pack_f16(float(v0), float(v1));
  w[0] = h;
  w[1] = pack_f16(float(v0 - f16_lo(h)), float(v1 - f16_hi(h)));
}

// Power-of-two scale s of a row whose largest magnitude is m: m s in
// [2^14, 2^15) (an all-zero row keeps s = 1); inv = 1 / s, exact.
__device__ inline void pow2_scale(double m, double& s, double& inv) {
  if (!(m > 0.0) || !(m < 1e300)) {
    s = 1.0;
    inv = 1.0;
    return;
  }
  int E = 0;
  (void)frexp(m, &E);  // m = f 2^E, f in [0.5, 1)
  const int e = min(max(15 - E, -1000), 1000);
  s = ldexp(1.0, e);
  inv = ldexp(1.0, -e);
}

template <int BM, int BN, bool AKC, bool BKC, bool KS, int NP>
__global__ __launch_bounds__(NT) void gemm_bf16x_kernel(Params p) {
  constexpr int WM = BM / 2, WN = BN / 2;
  constexpr int MB = WM / 16, NB = WN / 16;
  constexpr int OA = BM / 64, OB = BN / 64;  // octets per thread and slab
  constexpr int OT = OA + OB;
  __shared__ __attribute__((aligned(16))) uint32_t As[NP][BM][kRowWords];
  __shared__ __attribute__((aligned(16))) uint32_t Bs[NP][BN][kRowWords];
  const GemmDesc& d = p.d;
  const int t = threadIdx.x;
  const int lane = t & 63, wave = t >> 6;
  int64_t w = blockIdx.x;
  if (p.per > 0) {
    w = int64_t(blockIdx.x & 7) * p.per + (blockIdx.x >> 3);
    if (w >= p.work) return;
  }
  const int tiles = p.tiles_m * p.tiles_n;
  const int tile = int(w % tiles);
  const int64_t q = w / tiles;
  const int split = int(q % p.splits);
  const int b = int(q / p.splits);
  int tm, tn;
  if (p.nfast) {
    tm = tile / p.tiles_n;
    tn = tile % p.tiles_n;
  } else if (p.group > 1) {
    const int per_group = p.group * p.tiles_n;
    const int g = tile / per_group, r = tile - g * per_group;
    const int first = g * p.group;
    const int gs = min(p.tiles_m - first, p.group);
    tm = first + r % gs;
    tn = r / gs;
  } else {
    tm = tile % p.tiles_m;
    tn = tile / p.tiles_m;
  }
  const int m0 = tm * BM, n0 = tn * BN;
  const int k_begin = p.kbase + split * p.kchunk;
  const int k_end = min(d.K, k_begin + p.kchunk);

  const double* __restrict__ Ap = d.A + d.Ab * b;
  const double* __restrict__ Bp = d.B + d.Bb * b;
  const double* __restrict__ Kp = KS ? d.ks + d.Kb * b : nullptr;

  // Octet assignment.  k along the lanes (AKC / BKC): octet t % 4 of the slab,
  // rows t / 4 + 64 i.  Rows along the lanes: row t % BM, octets t / BM +
  // (NT / BM) i (wave-uniform k).
  auto a_row = [&](int i) { return AKC ? t / 4 + 64 * i : t % BM; };
  auto b_col = [&](int i) { return BKC ? t / 4 + 64 * i : t % BN; };
  auto a_oct = [&](int i) { return AKC ? t % 4 : t / BM + (NT / BM) * i; };
  auto b_oct = [&](int i) { return BKC ? t % 4 : t / BN + (NT / BN) * i; };
  // first k (slab-relative) of this thread's octets; wave-uniform with rows along the lanes
  const int a_k0 = AKC ? 8 * (t % 4) : __builtin_amdgcn_readfirstlane(8 * (t / BM));
  const int b_k0 = BKC ? 8 * (t % 4) : __builtin_amdgcn_readfirstlane(8 * (t / BN));
  constexpr int DA = AKC ? 0 : 8 * (NT / BM), DB = BKC ? 0 : 8 * (NT / BN);  // k step between octets

  double ra[OA][8], rb[OB][8], rs[AKC ? 1 : OA][8];
  uint32_t pa[OA][NP][4], pb[OB][NP][4];

  // split octet o (A octets first, then B) of the loaded slab into plane words
  auto convert = [&](int o) {
    if (o < OA) {
#pragma unroll
      for (int e = 0; e < 8; e += 2) {
        double v0 = ra[o][e], v1 = ra[o][e + 1];
        if constexpr (KS) {
          v0 *= rs[AKC ? 0 : o][e];
          v1 *= rs[AKC ? 0 : o][e + 1];
        }
        uint32_t wd[NP];
        split_pair<NP>(v0, v1, wd);
#pragma unroll
        for (int pl = 0; pl < NP; ++pl) pa[o][pl][e / 2] = wd[pl];
      }
    } else {
      const int ob = o - OA;
#pragma unroll
      for (int e = 0; e < 8; e += 2) {
        uint32_t wd[NP];
        split_pair<NP>(rb[ob][e], rb[ob][e + 1], wd);
#pragma unroll
        for (int pl = 0; pl < NP; ++pl) pb[ob][pl][e / 2] = wd[pl];
      }
    }
  };
  auto write_lds = [&]() {
#pragma unroll
    for (int i = 0; i < OA; ++i)
#pragma unroll
      for (int pl = 0; pl < NP; ++pl)
        *reinterpret_cast<uint4*>(&As[pl][a_row(i)][4 * a_oct(i)]) =
            uint4{pa[i][pl][0], pa[i][pl][1], pa[i][pl][2], pa[i][pl][3]};
#pragma unroll
    for (int i = 0; i < OB; ++i)
#pragma unroll
      for (int pl = 0; pl < NP; ++pl)
        *reinterpret_cast<uint4*>(&Bs[pl][b_col(i)][4 * b_oct(i)]) =
            uint4{pb[i][pl][0], pb[i][pl][1], pb[i][pl][2], pb[i][pl][3]};
  };

  f4 acc[MB][NB];
#pragma unroll
  for (int i = 0; i < MB; ++i)
#pragma unroll
    for (int j = 0; j < NB; ++j) acc[i][j] = f4{0, 0, 0, 0};

  const int wm0 = (wave & 1) * WM, wn0 = (wave >> 1) * WN;
  const int lr = lane & 15, lk = lane >> 4;

  // The MFMAs of one slab from LDS, with (convc true) the conversion of the
  // next slab's octets spread over the row blocks.  Lane (lr, lk) holds k =
  // 8 lk .. 8 lk + 7 of row / column lr of each 16-wide block, the same k order
  // in both operands.  Smallest plane products first.
  constexpr int NPROD = NP == 3 ? 6 : 3;
  auto compute = [&](auto convc) {
    bf8 bfr[NB][NP];
#pragma unroll
    for (int j = 0; j < NB; ++j)
#pragma unroll
      for (int pl = 0; pl < NP; ++pl)
        bfr[j][pl] = __builtin_bit_cast(bf8, *reinterpret_cast<const uint4*>(&Bs[pl][wn0 + 16 * j + lr][4 * lk]));
#pragma unroll
    for (int i = 0; i < MB; ++i) {
      bf8 afr[NP];
#pragma unroll
      for (int pl = 0; pl < NP; ++pl)
        afr[pl] = __builtin_bit_cast(bf8, *reinterpret_cast<const uint4*>(&As[pl][wm0 + 16 * i + lr][4 * lk]));
#pragma unroll
      for (int j = 0; j < NB; ++j) {
        f4 c = acc[i][j];
        if constexpr (NP == 3) {
          c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bfr[j][2], afr[0], c, 0, 0, 0);
          c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bfr[j][0], afr[2], c, 0, 0, 0);
          c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bfr[j][1], afr[1], c, 0, 0, 0);
        }
        c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bfr[j][1], afr[0], c, 0, 0, 0);
        c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bfr[j][0], afr[1], c, 0, 0, 0);
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bfr[j][0], afr[0], c, 0, 0, 0);
      }
      if constexpr (decltype(convc)::value) {
#pragma unroll
        for (int o = (i * OT) / MB; o < ((i + 1) * OT) / MB; ++o) convert(o);
      }
    }
    if constexpr (decltype(convc)::value) {
      // one slab as scheduling groups: the B operand reads, then per row block its
      // A reads and its MFMAs, each MFMA followed by a share of the conversion
      // (about 8 vector instructions per converted element pair and plane)
      constexpr int kValuPerMfma = (OT * 8 * (NP == 3 ? 9 : 5) / 2 + MB * NB * NPROD - 1) / (MB * NB * NPROD);
      __builtin_amdgcn_sched_group_barrier(0x100, NB * NP, 0);
#pragma unroll
      for (int i = 0; i < MB; ++i) {
        __builtin_amdgcn_sched_group_barrier(0x100, NP, 0);
#pragma unroll
        for (int m = 0; m < NB * NPROD; ++m) {
          __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
          __builtin_amdgcn_sched_group_barrier(0x002, kValuPerMfma, 0);
        }
      }
    }
  };

  auto pipeline = [&](auto&& ld) {
    ld(k_begin);
#pragma unroll
    for (int o = 0; o < OT; ++o) convert(o);
    write_lds();
    __syncthreads();
    ld(k_begin + BKB);
    const int nslab = (k_end - k_begin + BKB - 1) / BKB;
    for (int s = 0; s + 1 < nslab; ++s) {
      compute(std::true_type{});
      __syncthreads();
      write_lds();
      ld(k_begin + (s + 2) * BKB);
      __syncthreads();
    }
    compute(std::false_type{});
  };

  if (k_begin < k_end) {
    if (p.fast) {
      uint32_t a_lo[AKC ? OA : 1], b_lo[BKC ? OB : 1];
#pragma unroll
      for (int i = 0; i < (AKC ? OA : 1); ++i) {
        const int m = m0 + a_row(i);
        a_lo[i] = uint32_t(8 * ((m < d.M ? d.Am.off(m) : 0) + (AKC ? int64_t(a_k0) * d.Ak.s0 : 0)));
      }
#pragma unroll
      for (int i = 0; i < (BKC ? OB : 1); ++i) {
        const int n = n0 + b_col(i);
        b_lo[i] = uint32_t(8 * ((n < d.N ? d.Bn.off(n) : 0) + (BKC ? int64_t(b_k0) * d.Bk.s0 : 0)));
      }
      const uint32_t k_lo = uint32_t(8 * (AKC ? int64_t(a_k0) * d.Km.s0 : 0));
      const int k_last = k_end - BKB;
      pipeline([&](int k0) {
        if constexpr (AKC) {
          const int kc = min(k0, k_last);  // look-ahead past the end re-reads the last slab
          const auto r = rsrc(Ap + d.Ak.off(kc));
          if constexpr (KS) {
            const auto rk = rsrc(Kp + d.Km.off(kc));
#pragma unroll
            for (int e = 0; e < 8; ++e) rs[0][e] = ld_b(rk, k_lo, int32_t(8 * e * d.Km.s0));
          }
#pragma unroll
          for (int i = 0; i < OA; ++i)
#pragma unroll
            for (int e = 0; e < 8; ++e) ra[i][e] = ld_b(r, a_lo[i], int32_t(8 * e * d.Ak.s0));
        } else {
          const int kb = k0 + a_k0;  // wave-uniform
          const int64_t kbo = KS ? d.Km.off(kb) : 0;
          const double* a0 = Ap + d.Ak.off(kb);
          if (kb + DA * (OA - 1) + 7 < k_end) {  // every k of the slab in range (wave-uniform)
            const auto r0 = rsrc(a0);
#pragma unroll
            for (int i = 0; i < OA; ++i)
#pragma unroll
              for (int e = 0; e < 8; ++e) {
                const int dk = DA * i + e;
                ra[i][e] = ld_b(r0, a_lo[0], int32_t(8 * dk * d.Ak.s0));
                if constexpr (KS) rs[i][e] = Kp[kbo + dk * d.Km.s0];
              }
          } else {
#pragma unroll
            for (int i = 0; i < OA; ++i)
#pragma unroll
              for (int e = 0; e < 8; ++e) {
                const int dk = DA * i + e;
                const bool kok = kb + dk < k_end;
                ra[i][e] = ld_b(rsrc(a0, kok), a_lo[0], int32_t(8 * dk * d.Ak.s0));
                if constexpr (KS) rs[i][e] = Kp[kok ? kbo + dk * d.Km.s0 : 0];
              }
          }
        }
        if constexpr (BKC) {
          const auto r = rsrc(Bp + d.Bk.off(min(k0, k_last)));
#pragma unroll
          for (int i = 0; i < OB; ++i)
#pragma unroll
            for (int e = 0; e < 8; ++e) rb[i][e] = ld_b(r, b_lo[i], int32_t(8 * e * d.Bk.s0));
        } else {
          const int kb = k0 + b_k0;
          const double* b0 = Bp + d.Bk.off(kb);
          if (kb + DB * (OB - 1) + 7 < k_end) {
            const auto r0 = rsrc(b0);
#pragma unroll
            for (int i = 0; i < OB; ++i)
#pragma unroll
              for (int e = 0; e < 8; ++e) rb[i][e] = ld_b(r0, b_lo[0], int32_t(8 * (DB * i + e) * d.Bk.s0));
          } else {
#pragma unroll
            for (int i = 0; i < OB; ++i)
#pragma unroll
              for (int e = 0; e < 8; ++e) {
                const int dk = DB * i + e;
                const bool kok = kb + dk < k_end;
                rb[i][e] = ld_b(rsrc(b0, kok), b_lo[0], int32_t(8 * dk * d.Bk.s0));
              }
          }
        }
      });
    } else {
      // general path: 64-bit offsets, clamped in-bounds addresses and zeroed
      // values for k past the end
      int64_t a_moff[AKC ? OA : 1], b_noff[BKC ? OB : 1];
#pragma unroll
      for (int i = 0; i < (AKC ? OA : 1); ++i) {
        const int m = m0 + a_row(i);
        a_moff[i] = m < d.M ? d.Am.off(m) : 0;
      }
#pragma unroll
      for (int i = 0; i < (BKC ? OB : 1); ++i) {
        const int n = n0 + b_col(i);
        b_noff[i] = n < d.N ? d.Bn.off(n) : 0;
      }
      pipeline([&](int k0) {
#pragma unroll
        for (int i = 0; i < OA; ++i)
#pragma unroll
          for (int e = 0; e < 8; ++e) {
            const int k = k0 + a_k0 + DA * i + e;
            const bool kok = k < k_end;
            const double v = Ap[a_moff[AKC ? i : 0] + (kok ? d.Ak.off(k) : 0)];
            ra[i][e] = kok ? v : 0.0;
            if constexpr (KS) {
              if (!AKC || i == 0) rs[AKC ? 0 : i][e] = Kp[kok ? d.Km.off(k) : 0];
            }
          }
#pragma unroll
        for (int i = 0; i < OB; ++i)
#pragma unroll
          for (int e = 0; e < 8; ++e) {
            const int k = k0 + b_k0 + DB * i + e;
            const bool kok = k < k_end;
            const double v = Bp[b_noff[BKC ? i : 0] + (kok ? d.Bk.off(k) : 0)];
            rb[i][e] = kok ? v : 0.0;
          }
      });
    }
  }

  // epilogue: the FP32 16x16 output layout, C[m = .. + lr][n = .. + 4 lk + r]
  auto ncol = [&](int j, int r) { return n0 + wn0 + 16 * j + 4 * lk + r; };
  if (p.out) {
    double* W = p.out + (int64_t(split) * d.batch + b) * int64_t(d.M) * d.N;
#pragma unroll
    for (int i = 0; i < MB; ++i) {
      const int m = m0 + wm0 + 16 * i + lr;
      if (m >= d.M) continue;
#pragma unroll
      for (int j = 0; j < NB; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int n = ncol(j, r);
          if (n < d.N) W[m + int64_t(d.M) * n] = d.alpha * double(acc[i][j][r]);
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
        const int n = ncol(j, r);
        if (n >= d.N) continue;
        double* c = Cp + mo + d.Cn.off(n);
        const double v = d.alpha * double(acc[i][j][r]);
        *c = d.beta == 0.0 ? v : v + d.beta * *c;
      }
  }
}

// ---------------------------------------------------------------------------
// BF16 plane tensors.  For the split forms the operands are split once per
// product into packed bfloat16 planes (split_planes_kernel), so that the GEMM
// (gemm_planes_kernel) reads 2 bytes per element and plane instead of the 8 of
// the FP64 value, with no conversion in its main loop.  Layout of an operand
// viewed as X(r, k) (r: the output row for A, the output column for B) over
// nb batches: 16-byte units holding k = 8 q .. 8 q + 7 of one r in one plane,
// unit ((plane nb + b) KQ + q) Rp + r, with Rp the row count rounded up to
// 128 and KQ = 4 ceil(K / 32), padding zero.  A 32-deep slab of a 128-row
// tile is then 4 runs of 128 consecutive units (2 KiB) per plane.
struct PlaneDesc {
  const double* X = nullptr;
  Map2 Xr, Xk;
  int64_t Xb = 0;
  const double* ks = nullptr;  // optional k scale (A operand)
  Map2 Km;
  int64_t Kb = 0;
  int R = 0, K = 0, nb = 1, KQ = 0, Rp = 0;
  u32x4* out = nullptr;
  const double* scale = nullptr;  // FP16 form: power-of-two scale of row (b, r) at b Rp + r
  int x32 = 0;                     // X holds FP32 values
  __device__ double at(int64_t off) const {
    return x32 ? double(reinterpret_cast<const float*>(X)[off]) : X[off];
  }
};

// A workgroup per 128 rows (a lane per row, so the unit writes of a wave are
// contiguous), qc consecutive units (k runs of 8) and one batch, two units at a
// time; the row offset and scale are taken once.  Inside one period of the k
// map the 8 values of a unit are evenly spaced (one offset per unit); a unit
// that crosses a period or K, or has a k scale, takes one offset per k.
// F16: two half planes of the scaled values (NP = 2), for FP32 values scaled
// and split in FP32 (ldexp is exact, so the words equal those of the FP64
// steps); else NP bfloat16 planes.
template <int NP, bool F16>
__global__ __launch_bounds__(256) void split_planes_kernel(PlaneDesc s, int qc) {
  const int r = int(blockIdx.x) * 128 + int(threadIdx.x & 127);
  const int b = int(blockIdx.z);
  const int q1 = min(s.KQ, int(blockIdx.y + 1) * qc);
  const bool rok = r < s.R;
  const int64_t base = s.Xb * b + (rok ? s.Xr.off(uint32_t(r)) : 0);
  const int64_t plane = int64_t(s.nb) * s.KQ * s.Rp;
  const double sc = F16 ? s.scale[int64_t(b) * s.Rp + r] : 1.0;
  int ex = 0;
  if constexpr (F16) {
    (void)frexp(sc, &ex);  // sc = 2^(ex - 1)
    --ex;
  }
  const bool f32 = F16 && s.x32 && !s.ks;
  const float* xf = reinterpret_cast<const float*>(s.X);
  for (int q = int(blockIdx.y) * qc + int(threadIdx.x >> 7); q < q1; q += 2) {
    const int k0 = 8 * q;
    const uint32_t pq = fd_div(uint32_t(k0), s.Xk.fd);
    const uint32_t i0 = uint32_t(k0) - pq * s.Xk.n0;
    const bool even = rok && !s.ks && k0 + 8 <= s.K && i0 + 8 <= s.Xk.n0;
    const int64_t kb = base + int64_t(i0) * s.Xk.s0 + int64_t(pq) * s.Xk.s1;
    uint32_t w[4][NP];
    if (f32) {
      float v[8];
      if (even) {
#pragma unroll
        for (int e = 0; e < 8; ++e) v[e] = xf[kb + e * s.Xk.s0];
      } else {
#pragma unroll
        for (int e = 0; e < 8; ++e) {
          const int k = k0 + e;
          v[e] = rok && k < s.K ? xf[base + s.Xk.off(uint32_t(k))] : 0.0f;
        }
      }
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const float y0 = __builtin_amdgcn_ldexpf(v[2 * i], ex), y1 = __builtin_amdgcn_ldexpf(v[2 * i + 1], ex);
        const uint32_t h = pack_f16(y0, y1);
        w[i][0] = h;
        w[i][NP - 1] = pack_f16(y0 - float(f16_lo(h)), y1 - float(f16_hi(h)));
      }
    } else {
      double v[8];
      if (even) {
#pragma unroll
        for (int e = 0; e < 8; ++e) v[e] = s.at(kb + e * s.Xk.s0);
      } else {
#pragma unroll
        for (int e = 0; e < 8; ++e) {
          const int k = k0 + e;
          const bool ok = rok && k < s.K;
          double x = s.at(ok ? base + s.Xk.off(uint32_t(k)) : 0);
          if (s.ks) x *= s.ks[ok ? s.Kb * b + s.Km.off(uint32_t(k)) : 0];
          v[e] = ok ? x : 0.0;
        }
      }
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        if constexpr (F16) {
          static_assert(NP == 2, "the FP16 form has two planes");
          uint32_t t2[2];
          split_pair_f16(v[2 * i] * sc, v[2 * i + 1] * sc, t2);
          w[i][0] = t2[0];
          w[i][NP - 1] = t2[1];
        } else {
          split_pair<NP>(v[2 * i], v[2 * i + 1], w[i]);
        }
      }
    }
    const int64_t u = (int64_t(b) * s.KQ + q) * s.Rp + r;
#pragma unroll
    for (int pl = 0; pl < NP; ++pl) s.out[pl * plane + u] = u32x4{w[0][pl], w[1][pl], w[2][pl], w[3][pl]};
  }
}

// The same units for operands with k contiguous in memory and the rows strided
// (the kernel above would read 64 scattered words per wave and load): a
// workgroup per 64 rows, kc consecutive 32-k tiles and one batch stages each
// tile through LDS, loading along k (32 consecutive k of one row per 32 lanes)
// and writing the units along the rows (64 consecutive units per wave and
// plane); row offsets and scales are taken once.
template <int NP, bool F16>
__global__ __launch_bounds__(256) void split_planes_kc_kernel(PlaneDesc s, int kc) {
  __shared__ double T[64][33];
  const int64_t plane = int64_t(s.nb) * s.KQ * s.Rp;
  const int r0 = int(blockIdx.x) * 64, b = int(blockIdx.z);
  const int lr = int(threadIdx.x >> 5), kk = int(threadIdx.x & 31);
  int64_t ro[8];
  double sc[8];
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    const int r = r0 + lr + 8 * i;
    ro[i] = r < s.R ? s.Xb * b + s.Xr.off(uint32_t(r)) : -1;
    sc[i] = F16 ? s.scale[int64_t(b) * s.Rp + r] : 1.0;
  }
  const int kt1 = min(s.KQ / 4, int(blockIdx.y + 1) * kc);  // KQ = 4 ceil(K / 32)
  const int row = int(threadIdx.x & 63), qq = int(threadIdx.x >> 6);
  for (int kt = int(blockIdx.y) * kc; kt < kt1; ++kt) {
    const int k = 32 * kt + kk;
    const bool kok = k < s.K;
    const int64_t ko = kok ? s.Xk.off(uint32_t(k)) : 0;
    const double ksv = s.ks && kok ? s.ks[s.Kb * b + s.Km.off(uint32_t(k))] : 1.0;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      double x = 0.0;
      if (kok && ro[i] >= 0) {
        x = s.at(ro[i] + ko);
        if (s.ks) x *= ksv;
        if constexpr (F16) x *= sc[i];
      }
      T[lr + 8 * i][kk] = x;
    }
    __syncthreads();
    double v[8];
#pragma unroll
    for (int e = 0; e < 8; ++e) v[e] = T[row][8 * qq + e];
    uint32_t w[4][NP];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      if constexpr (F16) {
        uint32_t t2[2];
        split_pair_f16(v[2 * i], v[2 * i + 1], t2);
        w[i][0] = t2[0];
        w[i][NP - 1] = t2[1];
      } else {
        split_pair<NP>(v[2 * i], v[2 * i + 1], w[i]);
      }
    }
    const int64_t u = (int64_t(b) * s.KQ + 4 * kt + qq) * s.Rp + r0 + row;
#pragma unroll
    for (int pl = 0; pl < NP; ++pl) s.out[pl * plane + u] = u32x4{w[0][pl], w[1][pl], w[2][pl], w[3][pl]};
    __syncthreads();
  }
}

// Row scales of the FP16 form: for every row (b, r), r < Rp, of an operand the
// power of two of pow2_scale from max_k |X(r, k) ks(k)| (padding rows: 1), in
// two steps.  row_max_kernel: the k range is cut into kparts chunks so that
// operands with few rows still fill the chip; each thread (KC: each wave, with
// the lanes along k, when k is contiguous in memory) takes the maximum of one
// row over one chunk and folds it into mx[row] with an atomic maximum on the
// bit pattern (ordered like the value for non-negative doubles; mx starts at
// zero).  scale_from_max_kernel then turns mx[row] into the scale and its
// inverse, in place.
template <bool KC>
__global__ __launch_bounds__(256) void row_max_kernel(PlaneDesc s, int kparts, int kchunk, unsigned long long* mx) {
  const int64_t rows = int64_t(s.nb) * s.Rp;
  const int64_t gt = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  const int64_t w = KC ? gt / 64 : gt;
  const int lane = KC ? int(threadIdx.x & 63) : 0;
  if (w >= rows * kparts) return;  // wave-uniform when KC (the grid covers whole waves)
  const int64_t u = w % rows;
  const int part = int(w / rows);
  const int r = int(u % s.Rp), b = int(u / s.Rp);
  if (r >= s.R) return;
  const int k0 = part * kchunk, k1 = min(s.K, k0 + kchunk);
  const int64_t base = s.Xb * b + s.Xr.off(r);
  double m = 0.0;
  for (int k = k0 + lane; k < k1; k += (KC ? 64 : 1)) {
    double x = s.at(base + s.Xk.off(k));
    if (s.ks) x *= s.ks[s.Kb * b + s.Km.off(k)];
    m = fmax(m, fabs(x));
  }
  if constexpr (KC) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off));
    if (lane != 0) return;
  }
  if (m > 0.0) atomicMax(mx + u, static_cast<unsigned long long>(__double_as_longlong(m)));
}

__global__ __launch_bounds__(256) void scale_from_max_kernel(int64_t rows, double* scale, double* inv) {
  const int64_t u = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (u >= rows) return;
  double sc, iv;
  pow2_scale(scale[u], sc, iv);  // scale[u] holds the row maximum (the bits of mx)
  scale[u] = sc;
  inv[u] = iv;
}

// Row scales of the FP16 form from caller-supplied bounds on the row maxima:
// row (b, r) gets pow2_scale(hint[b hb + r]) (padding rows: 1).
__global__ __launch_bounds__(256) void hint_scale_kernel(const double* hint, int64_t hb, int R, int Rp, int nb,
                                                         double* scale, double* inv) {
  const int64_t u = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (u >= int64_t(nb) * Rp) return;
  const int r = int(u % Rp), b = int(u / Rp);
  double sc = 1.0, iv = 1.0;
  if (r < R) pow2_scale(hint[hb * b + r], sc, iv);
  scale[u] = sc;
  inv[u] = iv;
}

// GEMM over plane tensors: the tile order, split-K partials and epilogue of
// gemm_bf16x_kernel; k slabs of 32 loaded as 16-byte units (4 BM / 256 per
// thread and plane) into registers two slabs ahead, written to the same LDS
// layout, the same MFMA plane products.  F16: the planes are IEEE halves
// (v_mfma_f32_16x16x32_f16) and the epilogue undoes the row scales.
template <int BM, int BN, int NP, bool F16>
__global__ __launch_bounds__(NT, 2) void gemm_planes_kernel(Params p, PlaneArgs pa) {
  constexpr int WM = BM / 2, WN = BN / 2;
  constexpr int MB = WM / 16, NB = WN / 16;
  constexpr int UA = 4 * BM / NT, UB = 4 * BN / NT;  // units per thread, plane and slab
  // two planes: two LDS buffers for A (40 KiB at 128 x 128), and the B
  // fragments loaded by each wave straight from the plane tensor (whose units
  // are the fragments' 16-byte lane words) one slab ahead, so that only A goes
  // through LDS; three planes: A and B through one LDS buffer (their registers
  // allow one set of staged loads)
  constexpr int NBUF = NP == 2 ? 2 : 1;
  constexpr bool BG = NP == 2;
  __shared__ __attribute__((aligned(16))) uint32_t As[NBUF][NP][BM][kRowWords];
  __shared__ __attribute__((aligned(16))) uint32_t Bs[BG ? 1 : NBUF][NP][BG ? 1 : BN][kRowWords];
  const GemmDesc& d = p.d;
  const int t = threadIdx.x;
  const int lane = t & 63, wave = t >> 6;
  int tm, tn, split, b;
  if (!plane_tile(p, tm, tn, split, b)) return;
  const int m0 = tm * BM, n0 = tn * BN;
  const int k_begin = p.kbase + split * p.kchunk;
  const int k_end = min(d.K, k_begin + p.kchunk);

  const int64_t planeA = int64_t(pa.nbA) * pa.KQ * pa.RpA, planeB = int64_t(pa.nbB) * pa.KQ * pa.RpB;
  const u32x4* __restrict__ Ab = pa.A + int64_t(pa.nbA > 1 ? b : 0) * pa.KQ * pa.RpA + m0;
  const u32x4* __restrict__ Bb = pa.B + int64_t(pa.nbB > 1 ? b : 0) * pa.KQ * pa.RpB + n0;
  const int q0 = k_begin / 8;  // k_begin is a multiple of 32
  const int nslab = (k_end - k_begin + BKB - 1) / BKB;

  u32x4 xa[NP][UA], xb[NP][UB], ya[NP][UA], yb[NP][UB];  // ya, yb unused with three planes
  // unit i of this thread: row (t + NT i) % BM, slab-relative q (t + NT i) / BM
  auto ld = [&](u32x4(&ra)[NP][UA], u32x4(&rb)[NP][UB], int s) {
    const int qs = q0 + 4 * min(s, nslab - 1);
#pragma unroll
    for (int pl = 0; pl < NP; ++pl) {
#pragma unroll
      for (int i = 0; i < UA; ++i) {
        const int u = t + NT * i;
        ra[pl][i] = Ab[pl * planeA + int64_t(qs + u / BM) * pa.RpA + u % BM];
      }
      if constexpr (!BG) {
#pragma unroll
        for (int i = 0; i < UB; ++i) {
          const int u = t + NT * i;
          rb[pl][i] = Bb[pl * planeB + int64_t(qs + u / BN) * pa.RpB + u % BN];
        }
      }
    }
  };
  auto wr = [&](const u32x4(&ra)[NP][UA], const u32x4(&rb)[NP][UB], auto buf) {
#pragma unroll
    for (int pl = 0; pl < NP; ++pl) {
#pragma unroll
      for (int i = 0; i < UA; ++i) {
        const int u = t + NT * i;
        *reinterpret_cast<u32x4*>(&As[buf()][pl][u % BM][4 * (u / BM)]) = ra[pl][i];
      }
      if constexpr (!BG) {
#pragma unroll
        for (int i = 0; i < UB; ++i) {
          const int u = t + NT * i;
          *reinterpret_cast<u32x4*>(&Bs[buf()][pl][u % BN][4 * (u / BN)]) = rb[pl][i];
        }
      }
    }
  };

  f4 acc[MB][NB];
#pragma unroll
  for (int i = 0; i < MB; ++i)
#pragma unroll
    for (int j = 0; j < NB; ++j) acc[i][j] = f4{0, 0, 0, 0};
  const int wm0 = (wave & 1) * WM, wn0 = (wave >> 1) * WN;
  const int lr = lane & 15, lk = lane >> 4;
  // BG: this wave's B fragments of slab s (lane (lr, lk): unit lk of column wn0 + 16 j + lr)
  auto ldB = [&](u32x4(&f)[NB][NP], int s) {
    const int qs = q0 + 4 * min(s, nslab - 1) + lk;
#pragma unroll
    for (int pl = 0; pl < NP; ++pl)
#pragma unroll
      for (int j = 0; j < NB; ++j) f[j][pl] = Bb[pl * planeB + int64_t(qs) * pa.RpB + wn0 + 16 * j + lr];
  };
  // plane products (bfloat16 or half operands), smallest first: 3 products with
  // two planes, 6 with three; product p multiplies B plane kB[p] by A plane kA[p]
  constexpr int NPROD = NP == 3 ? 6 : 3;
  auto mma = [](const u32x4& b, const u32x4& a, f4 c) {
    if constexpr (F16)
      return __builtin_amdgcn_mfma_f32_16x16x32_f16(__builtin_bit_cast(h8, b), __builtin_bit_cast(h8, a), c, 0, 0, 0);
    else
      return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf8, b), __builtin_bit_cast(bf8, a), c, 0, 0, 0);
  };
  // One slab from LDS buffer `buf`: all fragments read first (those of the first
  // product first), then the products outermost, so that consecutive MFMAs go to
  // different accumulators (16 apart for 4 x 4 fragments) and never wait on each
  // other; each accumulator still sums its products in the same order.
  auto compute = [&](auto buf, const u32x4(&bg)[NB][NP]) {
    constexpr int kB[6] = {NP == 3 ? 2 : 1, 0, 1, 1, 0, 0};
    constexpr int kA[6] = {0, NP == 3 ? 2 : 1, 1, 0, 1, 0};
    constexpr int first = NP == 3 ? 0 : 3;  // two planes: the last three entries
    u32x4 afr[MB][NP], bfr[NB][NP];
#pragma unroll
    for (int i = 0; i < MB; ++i)
      afr[i][kA[first]] = *reinterpret_cast<const u32x4*>(&As[buf()][kA[first]][wm0 + 16 * i + lr][4 * lk]);
#pragma unroll
    for (int j = 0; j < NB; ++j)
      bfr[j][kB[first]] = BG ? bg[j][kB[first]]
                             : *reinterpret_cast<const u32x4*>(&Bs[buf()][kB[first]][wn0 + 16 * j + lr][4 * lk]);
#pragma unroll
    for (int pl = 0; pl < NP; ++pl) {
      if (pl != kA[first]) {
#pragma unroll
        for (int i = 0; i < MB; ++i)
          afr[i][pl] = *reinterpret_cast<const u32x4*>(&As[buf()][pl][wm0 + 16 * i + lr][4 * lk]);
      }
      if (pl != kB[first]) {
#pragma unroll
        for (int j = 0; j < NB; ++j)
          bfr[j][pl] = BG ? bg[j][pl] : *reinterpret_cast<const u32x4*>(&Bs[buf()][pl][wn0 + 16 * j + lr][4 * lk]);
      }
    }
#pragma unroll
    for (int pr = first; pr < 6; ++pr)
#pragma unroll
      for (int i = 0; i < MB; ++i)
#pragma unroll
        for (int j = 0; j < NB; ++j) acc[i][j] = mma(bfr[j][kB[pr]], afr[i][kA[pr]], acc[i][j]);
  };
  // Instruction order of one slab of the two-plane loop (compute, then the LDS
  // writes of the next slab and the loads of the one after it): the fragment
  // reads of the first product, then each MFMA paired with one of the other
  // fragment reads, LDS writes and global loads, so that they issue in the
  // shadow of the MFMAs.
  // (BG: the B fragment loads first, LDS reads and writes of A only)
  constexpr int NRD0 = BG ? MB : MB + NB, NRD1 = (BG ? MB : MB + NB) * (NP - 1), NMF = MB * NB * NPROD;
  constexpr int NWR = (UA + (BG ? 0 : UB)) * NP, NLD = (UA + (BG ? 0 : UB)) * NP, NBL = BG ? NB * NP : 0;
  auto schedule = [&]() {
    if constexpr (NBL > 0) __builtin_amdgcn_sched_group_barrier(0x020, NBL, 0);
    __builtin_amdgcn_sched_group_barrier(0x100, NRD0, 0);
#pragma unroll
    for (int i = 0; i < NRD1; ++i) {
      __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
      __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);
    }
#pragma unroll
    for (int i = 0; i < NWR; ++i) {
      __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
      __builtin_amdgcn_sched_group_barrier(0x200, 1, 0);
    }
#pragma unroll
    for (int i = 0; i < NLD; ++i) {
      __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
      __builtin_amdgcn_sched_group_barrier(0x020, 1, 0);
    }
    __builtin_amdgcn_sched_group_barrier(0x008, NMF - NRD1 - NWR - NLD, 0);
  };
  static_assert(NP == 3 || NMF >= NRD1 + NWR + NLD, "slab schedule");

  using B0 = std::integral_constant<int, 0>;
  if constexpr (NP == 3) {
    if (nslab > 0) {
      // three planes: one register set (two would leave one wave per SIMD);
      // LDS holds slab s while the loads of slab s + 1 are in flight
      ld(xa, xb, 0);
      wr(xa, xb, B0{});
      __syncthreads();
      ld(xa, xb, 1);
      u32x4 none[NB][NP];
      for (int s = 0;; ++s) {
        compute(B0{}, none);
        if (s + 1 >= nslab) break;
        __syncthreads();
        wr(xa, xb, B0{});
        __syncthreads();
        ld(xa, xb, s + 2);
      }
    }
  } else if (nslab > 0) {
    // slab s is read from LDS buffer s % 2 while the registers of slab s + 1
    // are written into the other buffer (whose last readers, slab s - 1, all
    // passed the previous barrier) and then reloaded with slab s + 3; the
    // other register set holds the loads of slab s + 2 in flight.  One
    // barrier per slab.  Past the last slab the writes and loads repeat it
    // (ld clamps the slab), unread.
    // The B fragments of slab s + 1 are loaded while slab s runs (bc, bn in turn).
    using B1 = std::integral_constant<int, NBUF - 1>;
    u32x4 bc[NB][NP], bn[NB][NP];
    ld(xa, xb, 0);
    ldB(bc, 0);
    ld(ya, yb, 1);
    wr(xa, xb, B0{});
    __syncthreads();
    ld(xa, xb, 2);
    for (int s = 0;; s += 2) {
      ldB(bn, s + 1);
      compute(B0{}, bc);
      wr(ya, yb, B1{});
      ld(ya, yb, s + 3);
      schedule();
      if (s + 1 >= nslab) break;
      __syncthreads();
      ldB(bc, s + 2);
      compute(B1{}, bn);
      wr(xa, xb, B0{});
      ld(xa, xb, s + 4);
      schedule();
      if (s + 2 >= nslab) break;
      __syncthreads();
    }
  }

  planes_store<MB, NB, F16>(p, pa, acc, m0 + wm0, n0 + wn0, lr, lk, split, b);
}

// Sum split-K partials and apply beta / the output map (fixed order: deterministic).
__global__ void splitk_reduce_kernel(GemmDesc d, const double* W, int splits) {
  const int64_t MN = int64_t(d.M) * d.N;
  const int b = blockIdx.y;
  for (int64_t e = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; e < MN; e += int64_t(gridDim.x) * blockDim.x) {
    double s = 0.0;
    for (int q = 0; q < splits; ++q) s += W[(int64_t(q) * d.batch + b) * MN + e];
    const int m = int(e % d.M), n = int(e / d.M);
    const int64_t off = d.Cb * b + d.Cm.off(m) + d.Cn.off(n);
    if (d.c32) {
      float* c = reinterpret_cast<float*>(d.C) + off;
      *c = float(d.beta == 0.0 ? s : s + d.beta * double(*c));
    } else {
      double* c = d.C + off;
      *c = d.beta == 0.0 ? s : s + d.beta * *c;
    }
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

// Profiling name, matching the template arguments of the instantiation
// (gemm_f64_kernel<BM, BN, AKC, BKC, KS, PREC>).
// (gemm_bf16x_kernel<BM, BN, AKC, BKC, KS, planes> for the BF16 split forms).
template <int BM, int BN>
std::string cfg_name(bool akc, bool bkc, bool ks, int prec) {
  const std::string tail = std::to_string(BM) + "," + std::to_string(BN) + "," + (akc ? "1" : "0") + "," +
                           (bkc ? "1" : "0") + "," + (ks ? "1" : "0") + ",";
  if (prec >= 2) return "gemm_bf16x_kernel<" + tail + std::to_string(prec) + ">";
  return "gemm_f64_kernel<" + tail + std::to_string(prec) + ">";
}

template <int BM, int BN, bool KS, int PREC>
void launch_ks(const Params& p, bool akc, bool bkc, dim3 grid, hipStream_t s) {
  if (akc && bkc) hipLaunchKernelGGL((gemm_f64_kernel<BM, BN, true, true, KS, PREC>), grid, dim3(NT), 0, s, p);
  else if (akc) hipLaunchKernelGGL((gemm_f64_kernel<BM, BN, true, false, KS, PREC>), grid, dim3(NT), 0, s, p);
  else if (bkc) hipLaunchKernelGGL((gemm_f64_kernel<BM, BN, false, true, KS, PREC>), grid, dim3(NT), 0, s, p);
  else hipLaunchKernelGGL((gemm_f64_kernel<BM, BN, false, false, KS, PREC>), grid, dim3(NT), 0, s, p);
}

template <int BM, int BN, bool KS, int NP>
void launch_bf(const Params& p, bool akc, bool bkc, dim3 grid, hipStream_t s) {
  if (akc && bkc) hipLaunchKernelGGL((gemm_bf16x_kernel<BM, BN, true, true, KS, NP>), grid, dim3(NT), 0, s, p);
  else if (akc) hipLaunchKernelGGL((gemm_bf16x_kernel<BM, BN, true, false, KS, NP>), grid, dim3(NT), 0, s, p);
  else if (bkc) hipLaunchKernelGGL((gemm_bf16x_kernel<BM, BN, false, true, KS, NP>), grid, dim3(NT), 0, s, p);
  else hipLaunchKernelGGL((gemm_bf16x_kernel<BM, BN, false, false, KS, NP>), grid, dim3(NT), 0, s, p);
}

template <int BM, int BN>
void launch_cfg(const Params& p, bool akc, bool bkc, dim3 grid, hipStream_t s) {
  if (p.d.prec >= 2) {
    const bool ks = p.d.ks != nullptr;
    if (p.d.prec == 2) ks ? launch_bf<BM, BN, true, 2>(p, akc, bkc, grid, s) : launch_bf<BM, BN, false, 2>(p, akc, bkc, grid, s);
    else ks ? launch_bf<BM, BN, true, 3>(p, akc, bkc, grid, s) : launch_bf<BM, BN, false, 3>(p, akc, bkc, grid, s);
  } else if (p.d.prec == 1) {
    if (p.d.ks) launch_ks<BM, BN, true, 1>(p, akc, bkc, grid, s);
    else launch_ks<BM, BN, false, 1>(p, akc, bkc, grid, s);
  } else {
    if (p.d.ks) launch_ks<BM, BN, true, 0>(p, akc, bkc, grid, s);
    else launch_ks<BM, BN, false, 0>(p, akc, bkc, grid, s);
  }
}

template <int BM, int BN, bool KS, int NP>
const void* kernel_ptr_bf(bool akc, bool bkc) {
  if (akc && bkc) return reinterpret_cast<const void*>(&gemm_bf16x_kernel<BM, BN, true, true, KS, NP>);
  if (akc) return reinterpret_cast<const void*>(&gemm_bf16x_kernel<BM, BN, true, false, KS, NP>);
  if (bkc) return reinterpret_cast<const void*>(&gemm_bf16x_kernel<BM, BN, false, true, KS, NP>);
  return reinterpret_cast<const void*>(&gemm_bf16x_kernel<BM, BN, false, false, KS, NP>);
}

template <int BM, int BN, bool KS, int PREC>
const void* kernel_ptr_ks(bool akc, bool bkc) {
  if (akc && bkc) return reinterpret_cast<const void*>(&gemm_f64_kernel<BM, BN, true, true, KS, PREC>);
  if (akc) return reinterpret_cast<const void*>(&gemm_f64_kernel<BM, BN, true, false, KS, PREC>);
  if (bkc) return reinterpret_cast<const void*>(&gemm_f64_kernel<BM, BN, false, true, KS, PREC>);
  return reinterpret_cast<const void*>(&gemm_f64_kernel<BM, BN, false, false, KS, PREC>);
}

// Resident workgroups per CU of one instantiation (cached) times the CU count.
template <int BM, int BN>
int64_t gemm_slots(bool akc, bool bkc, bool ks, int prec) {
  static std::mutex mu;
  static std::map<int, int64_t> cache;
  const int key = (akc ? 1 : 0) | (bkc ? 2 : 0) | (ks ? 4 : 0) | (prec << 3);
  std::lock_guard<std::mutex> lock(mu);
  auto it = cache.find(key);
  if (it != cache.end()) return it->second;
  const void* f = prec == 3   ? (ks ? kernel_ptr_bf<BM, BN, true, 3>(akc, bkc) : kernel_ptr_bf<BM, BN, false, 3>(akc, bkc))
                  : prec == 2 ? (ks ? kernel_ptr_bf<BM, BN, true, 2>(akc, bkc) : kernel_ptr_bf<BM, BN, false, 2>(akc, bkc))
                  : prec      ? (ks ? kernel_ptr_ks<BM, BN, true, 1>(akc, bkc) : kernel_ptr_ks<BM, BN, false, 1>(akc, bkc))
                              : (ks ? kernel_ptr_ks<BM, BN, true, 0>(akc, bkc) : kernel_ptr_ks<BM, BN, false, 0>(akc, bkc));
  int per_cu = 0, dev = 0, cus = 0;
  UW12G_HIP(hipGetDevice(&dev));
  UW12G_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  UW12G_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, f, NT, 0));
  const int64_t slots = int64_t(std::max(per_cu, 1)) * std::max(cus, 1);
  cache[key] = slots;
  return slots;
}

int64_t slots_for(int BM, int BN, bool akc, bool bkc, bool ks, int prec) {
  if (BM == 128 && BN == 128) return gemm_slots<128, 128>(akc, bkc, ks, prec);
  if (BM == 128) return gemm_slots<128, 64>(akc, bkc, ks, prec);
  if (BN == 128) return gemm_slots<64, 128>(akc, bkc, ks, prec);
  return gemm_slots<64, 64>(akc, bkc, ks, prec);
}

// Split of the k range: the number of chunks that maximises the useful
// fraction of the workgroup waves (a grid of w workgroups on S resident slots
// runs ceil(w / S) waves), discounted by the partial-sum reduction, whose cost
// relative to the GEMM is about 60 (2 s + 1) / K (8 bytes per partial element
// moved at ~1/60 of the FP64 matrix rate).
int choose_splits(int64_t wgs, int kt, int K, int batch, int64_t slots) {
  auto useful = [&](int s) {
    const int64_t units = wgs * s;
    const int64_t waves = (units + slots - 1) / slots;
    return double(units) / double(waves * slots);
  };
  int best_s = 1;
  double best = useful(1);
  for (int s = 2; s <= 64; ++s) {
    if (kt / s < 8 || int64_t(s) * batch > 65535) break;
    const double e = useful(s) / (1.0 + 60.0 * (2 * s + 1) / double(std::max(K, 1)));
    if (e > best + 0.01) {
      best = e;
      best_s = s;
    }
  }
  return best_s;
}

// Whether the two-plane product of an M x N output (batch products, kt k
// slabs of 32) runs faster on 256 x 256 tiles than on 128 x 128 tiles: each
// choice's share of useful output (padding) times its occupancy of the
// device over whole rounds of workgroups with the split-K count it would get,
// the 256 x 256 tiles weighted by their speed on full rounds (1.1: 1.07 to
// 1.14 measured on FP16x3 products from 25600 x 1280 x 960 to 8192^3).
bool wide_pays(int M, int N, int K, int kt, int batch, int64_t slots128, int64_t slots256) {
  auto score = [&](int t, int64_t slots) {
    const int64_t tiles = int64_t((M + t - 1) / t) * ((N + t - 1) / t) * batch;
    const int s = kt >= 16 ? choose_splits(tiles, kt, K, batch, slots) : 1;
    const int64_t units = tiles * s;
    const int64_t rounds = (units + slots - 1) / slots;
    const double pad = double(M) * N / (double((M + t - 1) / t * t) * double((N + t - 1) / t * t));
    return pad * double(units) / double(rounds * slots) / (1.0 + 60.0 * (2 * s + 1) / double(std::max(K, 1)));
  };
  return 1.1 * score(kWideTile, slots256) > score(128, slots128);
}

// Unit stride on the low level of a map (contiguous runs along that index).
bool unit(const Map2& m) { return m.s0 == 1 && m.n0 >= 2; }

// k depth of one slab: 16 for the FP64 and FP32 forms, 32 for the BF16 split forms.
int slab_depth(int prec) { return prec >= 2 ? BKB : BK; }

// A k map is affine inside every slab when its inner period is a multiple of
// the slab depth bk (slabs start at multiples of it) or it has a single level.
bool slab_affine(const Map2& m, int bk) { return m.n0 == 0xffffffffu || m.n0 % bk == 0; }
bool nonneg(const Map2& m) { return m.s0 >= 0 && m.s1 >= 0; }
}  // namespace
int64_t map_max(const Map2& m, int64_t extent);
namespace {

// Conditions for the unpredicated interior path of gemm_f64_kernel: affine k
// maps and every lane byte offset (row or column offset, plus the lane's k
// offset inside a slab in the k-contiguous forms) below 2^32.
// Largest lane offset (doubles) of the buffer-load path: lane plus uniform byte
// offset of every load below num_records = 2^31 - 1.
constexpr int64_t kLaneLim = (int64_t(1) << 28) - 1;

bool fast_maps_ok(const GemmDesc& d) {
  const int bk = slab_depth(d.prec);
  if (!slab_affine(d.Ak, bk) || !slab_affine(d.Bk, bk) || (d.ks && !slab_affine(d.Km, bk))) return false;
  return nonneg(d.Am) && nonneg(d.Ak) && nonneg(d.Bk) && nonneg(d.Bn) && (!d.ks || nonneg(d.Km));
}

bool fast_ok(const GemmDesc& d, bool akc, bool /*bkc*/) {
  if (!fast_maps_ok(d)) return false;
  const int bk = slab_depth(d.prec);
  const int64_t a = map_max(d.Am, d.M) + (bk - 1) * d.Ak.s0;
  const int64_t b = map_max(d.Bn, d.N) + (bk - 1) * d.Bk.s0;
  const int64_t k = (d.ks && akc) ? (bk - 1) * d.Km.s0 : 0;
  return a < kLaneLim && b < kLaneLim && k < kLaneLim;
}

// Inner period of a map over [0, extent): n0 for a two-level map that wraps
// inside the extent, else 1 (the map shifts at any start).
int64_t period(const Map2& m, int64_t extent) {
  return (m.n0 == 0xffffffffu || int64_t(m.n0) >= extent) ? 1 : int64_t(m.n0);
}

// Longest leading range [0, len) of the operand map `a` (len <= extent) whose
// offsets stay within `budget`, with len a multiple of the periods of `a` and
// of the output map `c`, so that both restart shifted after it:
// off(start + i) = off(start) + off(i) for start a multiple of len.  Ranges
// shorter than the extent are rounded to whole 128-wide tiles where the
// periods allow.  0: no such range.
int64_t chunk_len(const Map2& a, const Map2& c, int64_t extent, int64_t budget) {
  if (budget < 0) return 0;
  if (map_max(a, extent) <= budget) return extent;
  const int64_t pa = period(a, extent), pc = period(c, extent);
  const int64_t al = pa / std::gcd(pa, pc) * pc;
  int64_t lo = 0, hi = extent / al;  // map_max is nondecreasing in the length (nonnegative strides)
  while (lo < hi) {
    const int64_t mid = (lo + hi + 1) / 2;
    if (map_max(a, mid * al) <= budget) lo = mid;
    else hi = mid - 1;
  }
  int64_t len = lo * al;
  if (128 % al == 0 && len >= 128) len = len / 128 * 128;
  return len;
}

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
    check_extent(d.A, d.Ab * (d.batch - 1) + map_max(d.Am, d.M) + map_max(d.Ak, d.K), "gemm A", d.a32 ? 4 : 8);
    check_extent(d.B, d.Bb * (d.batch - 1) + map_max(d.Bk, d.K) + map_max(d.Bn, d.N), "gemm B", d.b32 ? 4 : 8);
    if (d.ks) check_extent(d.ks, d.Kb * (d.batch - 1) + map_max(d.Km, d.K), "gemm kscale");
  }
  check_extent(d.C, d.Cb * (d.batch - 1) + map_max(d.Cm, d.M) + map_max(d.Cn, d.N), "gemm C", d.c32 ? 4 : 8);
}

namespace {

// Largest split-K partial-sum buffer (doubles) a launch may ask for; past it
// the FP32 k chunks run as consecutive accumulating launches instead.
constexpr int64_t kSplitWsMax = int64_t(1) << 28;

// One launch over k in [kbase, d.K): split over K when the grid would leave
// most of the 256 CUs idle (kbase == 0 only).
void launch_gemm(const GemmDesc& d, int kbase, bool fast, bool akc, bool bkc, hipStream_t stream, Workspace* ws,
                 Profiler* prof) {
  // Tile shape: the best product of padding efficiency (useful fraction of
  // the padded output) and the relative speed of the tile shape.  Small
  // outputs (at most 64 tiles of 64 x 64, e.g. the N x N Fock sums over long
  // k, which run split-K) lose little with small tiles: measured relative
  // speeds there are ~0.96 / 0.86 for the half / quarter tiles, against ~0.9 /
  // 0.7 on large outputs.
  const bool small = int64_t((d.M + 63) / 64) * ((d.N + 63) / 64) * d.batch <= 64;
  int BM = 128, BN = 128;
  double best = -1.0;
  for (int bm : {128, 64})
    for (int bn : {128, 64}) {
      if (g_force_tile[0] && g_force_tile[0] != kWideTile && (bm != g_force_tile[0] || bn != g_force_tile[1]))
        continue;
      const double padded = double((d.M + bm - 1) / bm * bm) * double((d.N + bn - 1) / bn * bn);
      const double speed = (bm == 128 && bn == 128) ? 1.0 : (bm == 64 && bn == 64) ? (small ? 0.9 : 0.7)
                                                                                    : (small ? 0.95 : 0.9);
      const double score = speed * double(d.M) * double(d.N) / padded;
      if (score > best + 1e-12) {
        best = score;
        BM = bm;
        BN = bn;
      }
    }
  Params p;
  p.d = d;
  p.tiles_m = (d.M + BM - 1) / BM;
  p.tiles_n = (d.N + BN - 1) / BN;
  const int64_t tiles = int64_t(p.tiles_m) * p.tiles_n;
  if (tiles > 0x7fffffff) throw Error(-2, "gemm grid too large");
  const int klen = d.K - kbase;
  int splits = 1;
  const int64_t wgs = tiles * d.batch;
  const int bk = slab_depth(d.prec);
  const int kt = (klen + bk - 1) / bk;
  if (kbase == 0 && kt >= 16 && ws)
    splits = choose_splits(wgs, kt, klen, d.batch, slots_for(BM, BN, akc, bkc, d.ks != nullptr, d.prec));
  if (d.prec >= 1 && klen > kF32Chunk) {
    // FP32 accumulation runs over at most kF32Chunk k values; chunks add up in FP64
    const int nchunk = (klen + kF32Chunk - 1) / kF32Chunk;
    if (kbase != 0 || !ws || int64_t(std::max(splits, nchunk)) * d.batch * d.M * d.N > kSplitWsMax) {
      // large outputs (or a k remainder launch): the chunks run one after another,
      // each adding its FP32 sum onto the FP64 output (beta = 1 after the first),
      // so no partial-sum buffer of the output's size per chunk is needed
      for (int c = 0; c < nchunk; ++c) {
        GemmDesc e = d;
        e.K = std::min(d.K, kbase + (c + 1) * kF32Chunk);
        if (c > 0) e.beta = 1.0;
        launch_gemm(e, kbase + c * kF32Chunk, fast, akc, bkc, stream, nullptr, prof);
      }
      return;
    }
    splits = std::max(splits, nchunk);
    if (splits > 65535) throw Error(-2, "gemm: too many FP32 k chunks");
  }
  p.kchunk = ((kt + splits - 1) / splits) * bk;
  splits = (klen + p.kchunk - 1) / p.kchunk;
  p.splits = splits;
  p.kbase = kbase;
  p.out = nullptr;
  if (splits > 1) p.out = ws->get(size_t(splits) * d.batch * size_t(d.M) * d.N, stream);
  p.work = tiles * splits * int64_t(d.batch);
  p.per = p.work >= 64 ? (p.work + 7) / 8 : 0;
  const int64_t nblocks = p.per > 0 ? 8 * p.per : p.work;
  if (nblocks > 0x7fffffff) throw Error(-2, "gemm grid too large");
  p.fast = fast ? 1 : 0;
  p.nfast = (p.tiles_n <= 16 && p.tiles_m >= 4 * p.tiles_n) ? 1 : 0;
  p.group = (!p.nfast && p.tiles_m > 2 * kTileGroup && p.tiles_n > 2) ? kTileGroup : 0;
  const dim3 grid{static_cast<unsigned>(nblocks)};
  const bool timing = prof && prof->on;
  ProfileRec rec;
  if (timing) prof->begin(rec, stream);
  std::string kname;
  if (BM == 128 && BN == 128) {
    launch_cfg<128, 128>(p, akc, bkc, grid, stream);
    if (timing) kname = cfg_name<128, 128>(akc, bkc, d.ks != nullptr, d.prec);
  } else if (BM == 128) {
    launch_cfg<128, 64>(p, akc, bkc, grid, stream);
    if (timing) kname = cfg_name<128, 64>(akc, bkc, d.ks != nullptr, d.prec);
  } else if (BN == 128) {
    launch_cfg<64, 128>(p, akc, bkc, grid, stream);
    if (timing) kname = cfg_name<64, 128>(akc, bkc, d.ks != nullptr, d.prec);
  } else {
    launch_cfg<64, 64>(p, akc, bkc, grid, stream);
    if (timing) kname = cfg_name<64, 64>(akc, bkc, d.ks != nullptr, d.prec);
  }
  UW12G_HIP(hipGetLastError());
  if (timing) prof->end(rec, stream, kname, 2.0 * d.M * double(d.N) * klen * d.batch, 0.0);
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


// ---- BF16 split forms over plane tensors

// Largest plane workspace (bytes) one product may use (g_plane_bytes_max);
// larger products run in pieces over the batch, the output rows or the output
// columns.
constexpr int64_t kPlaneBytesMax = int64_t(16) << 30;

int64_t round_up(int64_t x, int64_t m) { return (x + m - 1) / m * m; }

// Plane kinds (GemmDesc::prec of the split forms): 2 two bfloat16 planes, 3
// three bfloat16 planes, 4 two half planes of row-scaled values.
constexpr int planes_of(int prec) { return prec == 3 ? 3 : 2; }

template <int NP, bool F16>
void launch_split(const PlaneDesc& s, hipStream_t stream, Profiler* prof) {
  const bool timing = prof && prof->on;
  ProfileRec rec;
  if (timing) prof->begin(rec, stream);
  // grid: row blocks, runs of k (at most 65535 of them), batches
  if (s.Xk.s0 == 1 && s.Xr.s0 != 1) {
    const int tiles = s.KQ / 4;  // 32-k tiles
    const int kc = std::max(8, (tiles + 65534) / 65535);
    const dim3 g(unsigned(s.Rp / 64), unsigned((tiles + kc - 1) / kc), unsigned(s.nb));
    hipLaunchKernelGGL((split_planes_kc_kernel<NP, F16>), g, dim3(256), 0, stream, s, kc);
  } else {
    const int qc = 2 * std::max(8, (s.KQ + 2 * 65535 - 1) / (2 * 65535));  // even
    const dim3 g(unsigned(s.Rp / 128), unsigned((s.KQ + qc - 1) / qc), unsigned(s.nb));
    hipLaunchKernelGGL((split_planes_kernel<NP, F16>), g, dim3(256), 0, stream, s, qc);
  }
  UW12G_HIP(hipGetLastError());
  if (timing)
    prof->end(rec, stream, "split_planes_kernel<" + std::to_string(NP) + "," + (F16 ? "1" : "0") + ">", 0.0,
              (8.0 + 2.0 * NP) * double(s.R) * s.K * s.nb);
}

// Row scales of one operand (FP16 form) from bounds on its row maxima.
void launch_hint_scale(const PlaneDesc& s, const double* hint, int64_t hb, double* scale, double* inv,
                       hipStream_t stream) {
  const int64_t rows = int64_t(s.nb) * s.Rp;
  hipLaunchKernelGGL(hint_scale_kernel, dim3(unsigned((rows + 255) / 256)), dim3(256), 0, stream, hint, hb, s.R, s.Rp,
                     s.nb, scale, inv);
  UW12G_HIP(hipGetLastError());
}

// Row scales of one operand (FP16 form): scale and inverse per row (b, r), r < Rp.
void launch_row_scale(const PlaneDesc& s, double* scale, double* inv, hipStream_t stream, Profiler* prof) {
  const int64_t rows = int64_t(s.nb) * s.Rp;
  const bool kc = s.Xk.s0 == 1 && s.Xr.s0 != 1;  // k contiguous: a wave per row and chunk
  // chunks of k: about 2^19 threads (KC: 2^13 waves) in all, chunks of at least 64 k
  const int64_t want = kc ? (int64_t(1) << 13) : (int64_t(1) << 19);
  const int kparts = int(std::max<int64_t>(1, std::min<int64_t>((want + rows - 1) / rows, (s.K + 63) / 64)));
  const int kchunk = (s.K + kparts - 1) / kparts;
  const int64_t threads = rows * kparts * (kc ? 64 : 1);
  const int64_t blocks = (threads + 255) / 256;
  if (blocks > 0x7fffffff) throw Error(-2, "row scale grid too large");
  const bool timing = prof && prof->on;
  ProfileRec rec;
  if (timing) prof->begin(rec, stream);
  fill(scale, rows, 0.0, stream);
  unsigned long long* mx = reinterpret_cast<unsigned long long*>(scale);
  if (kc)
    hipLaunchKernelGGL(row_max_kernel<true>, dim3(unsigned(blocks)), dim3(256), 0, stream, s, kparts, kchunk, mx);
  else
    hipLaunchKernelGGL(row_max_kernel<false>, dim3(unsigned(blocks)), dim3(256), 0, stream, s, kparts, kchunk, mx);
  UW12G_HIP(hipGetLastError());
  hipLaunchKernelGGL(scale_from_max_kernel, dim3(unsigned((rows + 255) / 256)), dim3(256), 0, stream, rows, scale, inv);
  UW12G_HIP(hipGetLastError());
  if (timing) prof->end(rec, stream, "row_max_kernel", 0.0, 8.0 * double(s.R) * s.K * s.nb);
}

template <int BM, int BN, int NP, bool F16>
const void* planes_kernel_ptr() {
  return reinterpret_cast<const void*>(&gemm_planes_kernel<BM, BN, NP, F16>);
}

template <int BM, int BN>
void launch_planes_cfg(const Params& p, const PlaneArgs& pa, int kind, dim3 grid, hipStream_t s) {
  if (kind == 2) hipLaunchKernelGGL((gemm_planes_kernel<BM, BN, 2, false>), grid, dim3(NT), 0, s, p, pa);
  else if (kind == 3) hipLaunchKernelGGL((gemm_planes_kernel<BM, BN, 3, false>), grid, dim3(NT), 0, s, p, pa);
  else hipLaunchKernelGGL((gemm_planes_kernel<BM, BN, 2, true>), grid, dim3(NT), 0, s, p, pa);
}

template <int BM, int BN>
int64_t planes_slots(int kind) {
  static std::mutex mu;
  static std::map<int, int64_t> cache;
  std::lock_guard<std::mutex> lock(mu);
  auto it = cache.find(kind);
  if (it != cache.end()) return it->second;
  const void* f = kind == 2   ? planes_kernel_ptr<BM, BN, 2, false>()
                  : kind == 3 ? planes_kernel_ptr<BM, BN, 3, false>()
                              : planes_kernel_ptr<BM, BN, 2, true>();
  int per_cu = 0, dev = 0, cus = 0;
  UW12G_HIP(hipGetDevice(&dev));
  UW12G_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  UW12G_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, f, NT, 0));
  const int64_t slots = int64_t(std::max(per_cu, 1)) * std::max(cus, 1);
  cache[kind] = slots;
  return slots;
}

// Splits one operand into its plane tensor s.out; the FP16 form first sets the
// rows' scales (scale, inv: nb Rp each) from the bounds `hint` when given, else
// from a pass over the operand.
void split_operand(PlaneDesc& s, int kind, const double* hint, int64_t hb, double* scale, double* inv,
                   hipStream_t stream, Profiler* prof) {
  if (kind == 4) {
    if (hint) launch_hint_scale(s, hint, hb, scale, inv, stream);
    else launch_row_scale(s, scale, inv, stream, prof);
    s.scale = scale;
    launch_split<2, true>(s, stream, prof);
  } else if (planes_of(kind) == 2) {
    launch_split<2, false>(s, stream, prof);
  } else {
    launch_split<3, false>(s, stream, prof);
  }
}

// Plane-tensor layout of operand A of a split-form product: its batches, k
// units, padded rows, plane bytes and (FP16) scale bytes.
struct APlanes {
  int nb = 1;
  int64_t KQ = 0, Rp = 0, bytes = 0, scale_bytes = 0;
};
APlanes a_planes(const GemmDesc& d) {
  APlanes a;
  a.nb = (d.Ab == 0 && (!d.ks || d.Kb == 0)) ? 1 : d.batch;
  a.KQ = 4 * ((int64_t(d.K) + BKB - 1) / BKB);
  a.Rp = round_up(d.M, kWideTile);
  a.bytes = int64_t(planes_of(d.prec)) * a.nb * a.KQ * a.Rp * 16;
  a.scale_bytes = d.prec == 4 ? 16 * int64_t(a.nb) * a.Rp : 0;
  return a;
}

PlaneDesc a_desc(const GemmDesc& d, const APlanes& a, u32x4* out) {
  PlaneDesc s;
  s.X = d.A;
  s.Xr = d.Am;
  s.Xk = d.Ak;
  s.Xb = d.Ab;
  s.ks = d.ks;
  s.Km = d.Km;
  s.Kb = d.Kb;
  s.R = d.M;
  s.K = d.K;
  s.nb = a.nb;
  s.KQ = int(a.KQ);
  s.Rp = int(a.Rp);
  s.out = out;
  s.x32 = d.a32;
  return s;
}

// One product C = alpha A B + beta C of a split form over plane tensors, or
// false when the workspace cannot hold its planes (then the caller runs the
// in-kernel split of gemm_bf16x_kernel).  Large products run in pieces.  With
// d.planesA, A's planes (and FP16 scales) come from a split_a_planes buffer.
bool gemm_bf16_planes(const GemmDesc& d, hipStream_t stream, Workspace* ws, Profiler* prof) {
  if (!ws || ws->carve) return false;
  const int kind = d.prec, np = planes_of(d.prec);
  const bool f16 = kind == 4;
  const APlanes ap = a_planes(d);
  const int nbA = ap.nb;
  const int nbB = d.Bb == 0 ? 1 : d.batch;
  const int64_t KQ = ap.KQ;
  const int64_t RpA = ap.Rp, RpB = round_up(d.N, kWideTile);
  const int64_t bytesA = d.planesA ? 0 : ap.bytes, bytesB = int64_t(np) * nbB * KQ * RpB * 16;
  if (KQ > 0x7fffffff / 8) return false;

  // tile shape as for the other forms
  const bool small = int64_t((d.M + 63) / 64) * ((d.N + 63) / 64) * d.batch <= 64;
  int BM = 128, BN = 128;
  double best = -1.0;
  for (int bm : {128, 64})
    for (int bn : {128, 64}) {
      if (g_force_tile[0] && g_force_tile[0] != kWideTile && (bm != g_force_tile[0] || bn != g_force_tile[1]))
        continue;
      const double padded = double((d.M + bm - 1) / bm * bm) * double((d.N + bn - 1) / bn * bn);
      const double speed = (bm == 128 && bn == 128) ? 1.0 : (bm == 64 && bn == 64) ? (small ? 0.9 : 0.7)
                                                                                    : (small ? 0.95 : 0.9);
      const double score = speed * double(d.M) * double(d.N) / padded;
      if (score > best + 1e-12) {
        best = score;
        BM = bm;
        BN = bn;
      }
    }
  // two planes: 256 x 256 tiles (gemm_planes_wide_kernel) where they pay
  // (beta != 0: the 128 x 128 tiles, whose two workgroups per CU overlap one's read-modify-write
  // epilogue with the other's MFMAs; 20480 x 5120 x 960 with beta = 1: 0.95 against 1.00 ms)
  const bool wide = np == 2 && (g_force_tile[0] == kWideTile ||
                                (!g_force_tile[0] && BM == 128 && BN == 128 && d.beta == 0.0 &&
                                 wide_pays(d.M, d.N, d.K, int(KQ / 4), d.batch, planes_slots<128, 128>(kind),
                                           wide_planes_slots(kind))));
  if (wide) BM = BN = kWideTile;
  Params p;
  p.d = d;
  p.tiles_m = (d.M + BM - 1) / BM;
  p.tiles_n = (d.N + BN - 1) / BN;
  const int64_t tiles = int64_t(p.tiles_m) * p.tiles_n;
  if (tiles > 0x7fffffff) return false;
  {
    // vector epilogue: unit column stride within periods of a multiple of 4, every other C offset
    // (rows, periods, batches) a multiple of 4 elements, and C 16-byte aligned
    auto m4 = [](int64_t x) { return x % 4 == 0; };
    const bool one = d.Cn.n0 == 0xffffffffu;
    p.cvec = d.Cn.s0 == 1 && (one || (d.Cn.n0 % 4 == 0 && m4(d.Cn.s1))) && m4(d.Cm.s0) &&
             (d.Cm.n0 == 0xffffffffu || m4(d.Cm.s1)) && m4(d.Cb) &&
             reinterpret_cast<uintptr_t>(d.C) % 16 == 0 ? 1 : 0;
  }
  const int kt = int(KQ / 4);
  const int64_t slots = wide        ? wide_planes_slots(kind)
                        : BM == 128 ? (BN == 128 ? planes_slots<128, 128>(kind) : planes_slots<128, 64>(kind))
                                    : (BN == 128 ? planes_slots<64, 128>(kind) : planes_slots<64, 64>(kind));
  int splits = kt >= 16 ? choose_splits(tiles * d.batch, kt, d.K, d.batch, slots) : 1;
  // FP32 sums over at most kF32Chunk k values: that many chunks at least
  const int nchunk = (d.K + kF32Chunk - 1) / kF32Chunk;
  bool sequential = false;
  if (nchunk > 1) {
    if (int64_t(std::max(splits, nchunk)) * d.batch * d.M * d.N > kSplitWsMax) sequential = true;
    else splits = std::max(splits, nchunk);
  }
  if (sequential) splits = 1;
  const int64_t part = splits > 1 ? int64_t(splits) * d.batch * d.M * d.N * 8 : 0;
  // FP16: scale and inverse scale per row of A and of B
  const int64_t bytesS = f16 ? round_up(16 * (int64_t(nbA) * RpA), 256) + round_up(16 * (int64_t(nbB) * RpB), 256) : 0;
  const int64_t need = round_up(bytesA, 256) + round_up(bytesB, 256) + bytesS + part;

  // workspace budget: what it holds already, or half of the free device memory
  size_t free_b = 0, total_b = 0;
  UW12G_HIP(hipMemGetInfo(&free_b, &total_b));
  const int64_t cap = g_plane_bytes_max > 0 ? g_plane_bytes_max : kPlaneBytesMax;
  const int64_t budget =
      std::min<int64_t>(cap, std::max<int64_t>(int64_t(ws->bytes), int64_t(ws->bytes) + int64_t(free_b) / 2));
  if (need > budget && d.planesA) {  // pieces re-split A from its values
    GemmDesc e = d;
    e.planesA = nullptr;
    return gemm_bf16_planes(e, stream, ws, prof);
  }
  if (need > budget) {
    // pieces: batch halves, else row / column halves (whole periods of the maps)
    if (d.batch > 1) {
      const int h = d.batch / 2;
      for (int b0 : {0, h}) {
        GemmDesc e = d;
        e.batch = b0 == 0 ? h : d.batch - h;
        e.A = advance(d.A, d.Ab * b0, d.a32);
        e.B = advance(d.B, d.Bb * b0, d.b32);
        e.C = advance(d.C, d.Cb * b0, d.c32);
        if (d.ks) e.ks = d.ks + d.Kb * b0;
        if (d.hintA) e.hintA = d.hintA + d.hAb * b0;
        if (d.hintB) e.hintB = d.hintB + d.hBb * b0;
        if (!gemm_bf16_planes(e, stream, ws, prof)) gemm_f64(e, stream, nullptr, prof);
      }
      return true;
    }
    const bool rows = bytesA >= bytesB;
    const Map2& om = rows ? d.Am : d.Bn;
    const Map2& cm = rows ? d.Cm : d.Cn;
    const int64_t ext = rows ? d.M : d.N;
    const int64_t pa_ = period(om, ext), pc_ = period(cm, ext);
    const int64_t per = pa_ / std::gcd(pa_, pc_) * pc_;
    int64_t half = round_up((ext + 1) / 2, std::max<int64_t>(per, 128));
    if (half >= ext || ext <= 128) return false;
    for (int64_t x0 = 0; x0 < ext; x0 += half) {
      GemmDesc e = d;
      const int len = int(std::min<int64_t>(half, ext - x0));
      if (rows) {
        e.M = len;
        e.A = advance(d.A, d.Am.off(uint32_t(x0)), d.a32);
        e.C = advance(d.C, d.Cm.off(uint32_t(x0)), d.c32);
        if (d.hintA) e.hintA = d.hintA + x0;
      } else {
        e.N = len;
        e.B = advance(d.B, d.Bn.off(uint32_t(x0)), d.b32);
        e.C = advance(d.C, d.Cn.off(uint32_t(x0)), d.c32);
        if (d.hintB) e.hintB = d.hintB + x0;
      }
      if (!gemm_bf16_planes(e, stream, ws, prof)) gemm_f64(e, stream, nullptr, prof);
    }
    return true;
  }

  char* base = reinterpret_cast<char*>(ws->get(size_t((need + 7) / 8), stream));
  PlaneArgs pa;
  pa.A = reinterpret_cast<const u32x4*>(base);
  pa.B = reinterpret_cast<const u32x4*>(base + round_up(bytesA, 256));
  char* sbase = base + round_up(bytesA, 256) + round_up(bytesB, 256);
  double* partials = part ? reinterpret_cast<double*>(sbase + bytesS) : nullptr;
  double* scA = f16 ? reinterpret_cast<double*>(sbase) : nullptr;
  double* scB = f16 ? reinterpret_cast<double*>(sbase + round_up(16 * (int64_t(nbA) * RpA), 256)) : nullptr;
  if (d.planesA) {  // A's planes, then its scales and inverse scales
    const char* c = reinterpret_cast<const char*>(d.planesA);
    pa.A = reinterpret_cast<const u32x4*>(c);
    if (f16) scA = const_cast<double*>(reinterpret_cast<const double*>(c + round_up(ap.bytes, 256)));
  }
  if (f16) {
    pa.invA = scA + int64_t(nbA) * RpA;
    pa.invB = scB + int64_t(nbB) * RpB;
  }
  pa.nbA = nbA;
  pa.nbB = nbB;
  pa.RpA = int(RpA);
  pa.RpB = int(RpB);
  pa.KQ = int(KQ);
  PlaneDesc sa = a_desc(d, ap, const_cast<u32x4*>(pa.A));
  PlaneDesc sb;
  sb.X = d.B;
  sb.Xr = d.Bn;
  sb.Xk = d.Bk;
  sb.Xb = d.Bb;
  sb.R = d.N;
  sb.K = d.K;
  sb.nb = nbB;
  sb.KQ = int(KQ);
  sb.Rp = int(RpB);
  sb.out = const_cast<u32x4*>(pa.B);
  sb.x32 = d.b32;
  if (!d.planesA) split_operand(sa, kind, d.hintA, d.hAb, scA, const_cast<double*>(pa.invA), stream, prof);
  split_operand(sb, kind, d.hintB, d.hBb, scB, const_cast<double*>(pa.invB), stream, prof);

  const int nlaunch = sequential ? nchunk : 1;
  for (int c = 0; c < nlaunch; ++c) {
    Params q = p;
    q.kbase = sequential ? c * kF32Chunk : 0;
    const int klen = (sequential ? std::min(d.K, (c + 1) * kF32Chunk) : d.K) - q.kbase;
    q.d.K = q.kbase + klen;
    if (sequential && c > 0) q.d.beta = 1.0;
    const int ktc = (klen + BKB - 1) / BKB;
    q.kchunk = ((ktc + splits - 1) / splits) * BKB;
    q.splits = (klen + q.kchunk - 1) / q.kchunk;
    q.out = q.splits > 1 ? partials : nullptr;
    q.work = tiles * q.splits * int64_t(d.batch);
    q.per = q.work >= 64 ? (q.work + 7) / 8 : 0;
    const int64_t nblocks = q.per > 0 ? 8 * q.per : q.work;
    if (nblocks > 0x7fffffff) throw Error(-2, "gemm grid too large");
    q.fast = 1;
    q.nfast = (q.tiles_n <= 16 && q.tiles_m >= 4 * q.tiles_n) ? 1 : 0;
    q.group = (!q.nfast && q.tiles_m > 2 * kTileGroup && q.tiles_n > 2) ? kTileGroup : 0;
    const dim3 grid{static_cast<unsigned>(nblocks)};
    const bool timing = prof && prof->on;
    ProfileRec rec;
    if (timing) prof->begin(rec, stream);
    if (wide) launch_planes_wide(q, pa, kind, grid, stream);
    else if (BM == 128 && BN == 128) launch_planes_cfg<128, 128>(q, pa, kind, grid, stream);
    else if (BM == 128) launch_planes_cfg<128, 64>(q, pa, kind, grid, stream);
    else if (BN == 128) launch_planes_cfg<64, 128>(q, pa, kind, grid, stream);
    else launch_planes_cfg<64, 64>(q, pa, kind, grid, stream);
    UW12G_HIP(hipGetLastError());
    if (timing)
      prof->end(rec, stream,
                (wide ? std::string("gemm_planes_wide_kernel<") + (f16 ? "1" : "0") + ">"
                      : "gemm_planes_kernel<" + std::to_string(BM) + "," + std::to_string(BN) + "," +
                            std::to_string(np) + "," + (f16 ? "1" : "0") + ">"),
                2.0 * d.M * double(d.N) * klen * d.batch,
                // compulsory HBM bytes: both operands' planes over this k range once, the output
                // (read too for beta), the split-K partials written
                2.0 * np * double(klen) * (double(d.M) * nbA + double(d.N) * nbB) +
                    double(d.M) * d.N * d.batch * (d.c32 ? 4.0 : 8.0) * (q.d.beta != 0.0 && q.splits == 1 ? 2.0 : 1.0) +
                    (q.splits > 1 ? 8.0 * q.splits * double(d.M) * d.N * d.batch : 0.0));
    if (q.splits > 1) {
      const int64_t MN = int64_t(d.M) * d.N;
      dim3 g(unsigned(std::min<int64_t>((MN + 255) / 256, 4096)), unsigned(d.batch));
      ProfileRec r2;
      if (timing) prof->begin(r2, stream);
      hipLaunchKernelGGL(splitk_reduce_kernel, g, dim3(256), 0, stream, q.d, partials, q.splits);
      UW12G_HIP(hipGetLastError());
      if (timing)
        prof->end(r2, stream, "splitk_reduce_kernel", double(q.splits) * d.batch * MN,
                  8.0 * (q.splits + 1) * d.batch * MN);
    }
  }
  return true;
}

}  // namespace

size_t plane_cache_bytes(const GemmDesc& d) {
  if (d.prec < 2 || !g_bf16_planes || d.M <= 0 || d.K <= 0) return 0;
  const APlanes a = a_planes(d);
  if (a.KQ > 0x7fffffff / 8) return 0;
  return size_t(round_up(a.bytes, 256) + a.scale_bytes);
}

void split_a_planes(const GemmDesc& d, void* out, hipStream_t stream, Profiler* prof) {
  if (plane_cache_bytes(d) == 0) throw Error(-2, "split_a_planes: not a split-form product");
  const APlanes a = a_planes(d);
  char* c = reinterpret_cast<char*>(out);
  double* scale = a.scale_bytes ? reinterpret_cast<double*>(c + round_up(a.bytes, 256)) : nullptr;
  PlaneDesc s = a_desc(d, a, reinterpret_cast<u32x4*>(c));
  split_operand(s, d.prec, d.hintA, d.hAb, scale, scale ? scale + int64_t(a.nb) * a.Rp : nullptr, stream, prof);
}

namespace {
// FP64 / FP32 products whose k maps share an inner period n0 that is not a multiple of the slab
// depth (the (m, A) contractions over the ABS rank r, k scales repeating every o occupied
// orbitals) would take the predicated general path:
// the first n0 - n0 % bk of every period run as one product with slab-affine maps, the rest of
// each period as a second product accumulated onto it.  False when that split does not apply.
bool split_period_remainder(const GemmDesc& d, hipStream_t stream, Workspace* ws, Profiler* prof) {
  const int bk = slab_depth(d.prec);
  if (d.prec >= 2) return false;
  // the common period of the two-level k maps (a one-level map of stride s is the two-level map
  // (n0; s, n0 s) for any n0)
  const uint32_t one = 0xffffffffu;
  uint32_t n0 = one;
  for (const Map2* k : {&d.Ak, &d.Bk, d.ks ? &d.Km : nullptr}) {
    if (!k || k->n0 == one) continue;
    if (n0 != one && k->n0 != n0) return false;
    n0 = k->n0;
  }
  if (n0 == one || n0 % bk == 0 || n0 < uint32_t(bk) || d.K % int64_t(n0) != 0) return false;
  auto s1 = [&](const Map2& k) { return k.n0 == one ? int64_t(n0) * k.s0 : k.s1; };
  const int q = int(d.K / int64_t(n0));
  const uint32_t n1 = n0 - n0 % bk, r = n0 % bk;
  GemmDesc m = d;
  m.K = int(n1) * q;
  m.Ak = two(n1, d.Ak.s0, s1(d.Ak));
  m.Bk = two(n1, d.Bk.s0, s1(d.Bk));
  if (d.ks) m.Km = two(n1, d.Km.s0, s1(d.Km));
  if (!fast_maps_ok(m)) return false;
  GemmDesc t = d;
  t.K = int(r) * q;
  t.beta = 1.0;
  t.A = d.A + int64_t(n1) * d.Ak.s0;
  t.B = d.B + int64_t(n1) * d.Bk.s0;
  t.Ak = two(r, d.Ak.s0, s1(d.Ak));
  t.Bk = two(r, d.Bk.s0, s1(d.Bk));
  if (d.ks) {
    t.ks = d.ks + int64_t(n1) * d.Km.s0;
    t.Km = two(r, d.Km.s0, s1(d.Km));
  }
  gemm_f64(m, stream, ws, prof);
  gemm_f64(t, stream, ws, prof);
  return true;
}
}  // namespace

void gemm_f64(const GemmDesc& d, hipStream_t stream, Workspace* ws, Profiler* prof) {
  if (d.M <= 0 || d.N <= 0 || d.batch <= 0) return;
  check_desc(d);
  if (d.batch > 65535) throw Error(-2, "gemm batch exceeds 65535");
  const bool f32s = d.a32 || d.b32 || d.c32;
  if (f32s && (d.prec < 2 || d.K <= 0))
    throw Error(-2, "gemm: FP32 operand storage is read only by the split forms over plane tensors");
  if (d.K <= 0) {  // C <- beta C
    GemmDesc z = d;
    z.K = 0;
    gemm_f64_naive(z, stream);
    return;
  }
  // BF16 split forms: operands split once into plane tensors when the workspace holds them
  if (d.prec >= 2 && g_bf16_planes && gemm_bf16_planes(d, stream, ws, prof)) return;
  if (f32s) throw Error(-2, "gemm: FP32 operand storage needs the plane-tensor workspace of the split forms");
  if (d.prec == 4) {
    // the FP16 form needs its row scales and planes in the workspace; without one the
    // product runs as the in-kernel split with three bfloat16 planes (2^-22 per product)
    GemmDesc e = d;
    e.prec = 3;
    gemm_f64(e, stream, ws, prof);
    return;
  }
  if (!fast_maps_ok(d) && split_period_remainder(d, stream, ws, prof)) return;
  // coalesce along k only when that index has the unit stride and the other does not
  const bool akc = unit(d.Ak) && !unit(d.Am);
  const bool bkc = unit(d.Bk) && !unit(d.Bn);
  // Operands whose row / column offsets pass the buffer path's 32-bit lane offsets:
  // split the output into blocks of rows / columns, each launched from a 64-bit
  // base pointer (A, B and C shifted to the block's first row / column), so the
  // lane offsets of every block stay small and the block keeps the fast path.
  const int bk = slab_depth(d.prec);
  if (d.K >= bk && fast_maps_ok(d) && !fast_ok(d, akc, bkc)) {
    const int64_t ka = (bk - 1) * d.Ak.s0, kb = (bk - 1) * d.Bk.s0;
    const bool kfit = !(d.ks && akc) || (bk - 1) * d.Km.s0 < kLaneLim;
    const int64_t cm = chunk_len(d.Am, d.Cm, d.M, kLaneLim - 1 - ka);
    const int64_t cn = chunk_len(d.Bn, d.Cn, d.N, kLaneLim - 1 - kb);
    if (kfit && cm > 0 && cn > 0 && (cm < d.M || cn < d.N)) {
      for (int64_t m0 = 0; m0 < d.M; m0 += cm)
        for (int64_t n0 = 0; n0 < d.N; n0 += cn) {
          GemmDesc c = d;
          c.M = int(std::min<int64_t>(cm, d.M - m0));
          c.N = int(std::min<int64_t>(cn, d.N - n0));
          c.A = d.A + d.Am.off(uint32_t(m0));
          c.B = d.B + d.Bn.off(uint32_t(n0));
          c.C = d.C + d.Cm.off(uint32_t(m0)) + d.Cn.off(uint32_t(n0));
          if (d.hintA) c.hintA = d.hintA + m0;
          if (d.hintB) c.hintB = d.hintB + n0;
          gemm_f64(c, stream, ws, prof);
        }
      return;
    }
  }
  // with k along the lanes the unpredicated path needs whole slabs: the k
  // remainder then runs on the general path, accumulated onto them
  const int krem = (akc || bkc) ? d.K % bk : 0;
  if (d.K >= bk && fast_ok(d, akc, bkc)) {
    GemmDesc m = d;
    m.K = d.K - krem;
    launch_gemm(m, 0, true, akc, bkc, stream, ws, prof);
    if (krem) {
      GemmDesc r = d;
      r.beta = 1.0;
      launch_gemm(r, d.K - krem, false, akc, bkc, stream, ws, prof);
    }
  } else {
    launch_gemm(d, 0, false, akc, bkc, stream, ws, prof);
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
