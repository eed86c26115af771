// Two-plane split-form products (BF16x3, FP16x3) on 256 x 256 output tiles.
//
// Eight waves of 128 x 64 each (two per SIMD), so that a workgroup loads each
// operand unit of its tile once from the plane tensors: a third of the bytes
// per MFMA that the 128 x 128 tiles move from L2.  The units move from the
// plane tensors straight into LDS with global_load_lds_dwordx4 (no staging
// registers): a slab is 4 k units of 256 rows per operand and plane, in the
// order of the plane tensors, and LDS holds two slabs (128 KiB).  One barrier
// per slab; after it each wave reads the first product's fragments of slab s,
// issues its transfers of slab s + 1 into the other buffer (they have one slab
// of MFMAs, 3 x 32 per wave, to land), then reads its other fragments paired
// with the first product's MFMAs.  128 accumulators and 96 fragment
// registers per wave fit the 256 VGPRs of two waves per SIMD.  Per accumulator
// the plane products and slabs run in the order of gemm_planes_kernel, and the
// epilogue is the same: the results equal those of the 128 x 128 tiles bit for
// bit.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "gemm_planes.h"

namespace uw12g {
namespace {

template <bool F16>
__global__ __launch_bounds__(2 * NT, 2) void gemm_planes_wide_kernel(Params p, PlaneArgs pa) {
  constexpr int BM = kWideTile, BN = kWideTile, NP = 2;
  constexpr int WM = BM / 2, WN = BN / 4;
  constexpr int MB = WM / 16, NB = WN / 16;
  // [buffer][operand][plane][k unit][row]: the units of a k slab in the order
  // of the plane tensors, 64 KiB per buffer
  __shared__ u32x4 L[2][2][NP][4][BM];
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
  // slab s into buffer buf: wave w moves rows 64 (w % 4) .. 64 (w % 4) + 63 of
  // operand w / 4 (A, B) in every plane and k unit, one 1 KiB global_load_lds each
  const int op = wave >> 2, rc = 64 * (wave & 3);
  const uint32_t ro = rc + lane;
  const u32x4* __restrict__ Xb = op ? Bb : Ab;
  const int64_t planeX = op ? planeB : planeA;
  const int RpX = op ? pa.RpB : pa.RpA;
  auto issue = [&](int s, int buf) {
    const int qs = q0 + 4 * s;
#pragma unroll
    for (int pl = 0; pl < NP; ++pl)
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const u32x4* g = Xb + (pl * planeX + int64_t(qs + q) * RpX);
        __builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void*)(g + ro),
                                         (__attribute__((address_space(3))) void*)&L[buf][op][pl][q][rc], 16, 0, 0);
      }
  };

  f4 acc[MB][NB];
#pragma unroll
  for (int i = 0; i < MB; ++i)
#pragma unroll
    for (int j = 0; j < NB; ++j) acc[i][j] = f4{0, 0, 0, 0};
  const int wm0 = (wave & 1) * WM, wn0 = (wave >> 1) * WN;
  const int lr = lane & 15, lk = lane >> 4;
  auto mma = [](const u32x4& bv, const u32x4& av, f4 c) {
    if constexpr (F16)
      return __builtin_amdgcn_mfma_f32_16x16x32_f16(__builtin_bit_cast(h8, bv), __builtin_bit_cast(h8, av), c, 0, 0, 0);
    else
      return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf8, bv), __builtin_bit_cast(bf8, av), c, 0,
                                                     0, 0);
  };
  // slab s from buffer s % 2; the transfers of slab s + 1 are issued once the
  // reads of the first product's fragments are in flight (the last slab
  // repeats its own transfers into the idle buffer, unread, so that the slab
  // has no branch; they are waited for before the epilogue)
  auto compute = [&](int s) {
    const int buf = s & 1;
    u32x4 af[MB][NP], bf[NB][NP];
#pragma unroll
    for (int j = 0; j < NB; ++j) bf[j][1] = L[buf][1][1][lk][wn0 + 16 * j + lr];
#pragma unroll
    for (int i = 0; i < MB; ++i) af[i][0] = L[buf][0][0][lk][wm0 + 16 * i + lr];
    issue(min(s + 1, nslab - 1), buf ^ 1);
#pragma unroll
    for (int j = 0; j < NB; ++j) bf[j][0] = L[buf][1][0][lk][wn0 + 16 * j + lr];
#pragma unroll
    for (int i = 0; i < MB; ++i) af[i][1] = L[buf][0][1][lk][wm0 + 16 * i + lr];
#pragma unroll
    for (int i = 0; i < MB; ++i)
#pragma unroll
      for (int j = 0; j < NB; ++j) acc[i][j] = mma(bf[j][1], af[i][0], acc[i][j]);
#pragma unroll
    for (int i = 0; i < MB; ++i)
#pragma unroll
      for (int j = 0; j < NB; ++j) acc[i][j] = mma(bf[j][0], af[i][1], acc[i][j]);
#pragma unroll
    for (int i = 0; i < MB; ++i)
#pragma unroll
      for (int j = 0; j < NB; ++j) acc[i][j] = mma(bf[j][0], af[i][0], acc[i][j]);
    // the first product's fragment reads, the transfers of the next slab, then
    // each of the other fragment reads paired with one of the first product's
    // MFMAs: every read is in flight before the first MFMA that needs it,
    // instead of reads and waits spread through the slab
    __builtin_amdgcn_sched_group_barrier(0x100, MB + NB, 0);
    __builtin_amdgcn_sched_group_barrier(0x020, 4 * NP, 0);
#pragma unroll
    for (int i = 0; i < MB + NB; ++i) {
      __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
      __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);
    }
    __builtin_amdgcn_sched_group_barrier(0x008, 3 * MB * NB - MB - NB, 0);
  };

  if (nslab > 0) issue(0, 0);
  for (int s = 0; s < nslab; ++s) {
    // slab s complete in LDS (each wave waits for its own transfers, the
    // barrier for everyone's); buffer (s + 1) % 2 no longer read (slab s - 1)
    __builtin_amdgcn_s_waitcnt(0x0f70);  // vmcnt(0)
    __syncthreads();
    compute(s);
  }
  __builtin_amdgcn_s_waitcnt(0x0f70);  // vmcnt(0): no transfer into LDS outlives the workgroup
  planes_store<MB, NB, F16>(p, pa, acc, m0 + wm0, n0 + wn0, lr, lk, split, b);
}

}  // namespace

int64_t wide_planes_slots(int kind) {
  static int64_t slots[2] = {0, 0};
  int64_t& n = slots[kind == 2 ? 0 : 1];
  if (n == 0) {
    const void* f = kind == 2 ? reinterpret_cast<const void*>(&gemm_planes_wide_kernel<false>)
                              : reinterpret_cast<const void*>(&gemm_planes_wide_kernel<true>);
    int per_cu = 0, dev = 0, cus = 0;
    UW12G_HIP(hipGetDevice(&dev));
    UW12G_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    UW12G_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, f, 2 * NT, 0));
    n = int64_t(std::max(per_cu, 1)) * std::max(cus, 1);
  }
  return n;
}

void launch_planes_wide(const Params& p, const PlaneArgs& pa, int kind, dim3 grid, hipStream_t s) {
  // every workgroup reads whole 256-row panels of both plane tensors
  if (pa.RpA % kWideTile != 0 || pa.RpB % kWideTile != 0 || kind == 3)
    throw Error(-2, "wide plane GEMM: plane rows not padded to 256, or three planes");
  if (kind == 2) hipLaunchKernelGGL((gemm_planes_wide_kernel<false>), grid, dim3(2 * NT), 0, s, p, pa);
  else hipLaunchKernelGGL((gemm_planes_wide_kernel<true>), grid, dim3(2 * NT), 0, s, p, pa);
}

}  // namespace uw12g
