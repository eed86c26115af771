// Definitions shared by the GEMM kernels of gemm_f64.hip and the wide-tile
// plane GEMM of gemm_planes_wide.hip.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "uw12g_internal.h"

namespace uw12g {

typedef float f4 __attribute__((ext_vector_type(4)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
typedef _Float16 h8 __attribute__((ext_vector_type(8)));

constexpr int NT = 256;         // threads per GEMM workgroup
constexpr int BKB = 32;         // k depth of one slab of the split forms (one MFMA k step)
constexpr int kWideTile = 256;  // rows and columns of a gemm_planes_wide_kernel tile

struct Params {
  GemmDesc d;
  int tiles_m, tiles_n, splits, kchunk;
  double* out;  // split-K partials (dense, M x N per split and batch) or nullptr
  int64_t work;  // output tiles x k chunks x batches: one workgroup each
  int64_t per;   // > 0: work items per XCD (contiguous ranges; the grid is 8 per)
  int fast;     // 1: buffer-load main path (fast_ok); every slab of every chunk is full
  int nfast;    // 1: tiles ordered with the column index fastest
  int group;    // > 1 (with nfast 0): tiles ordered in groups of this many tile rows, row index fastest
  int kbase;    // first k of chunk 0
  int cvec = 0;  // plane GEMMs: C columns 4 n .. 4 n + 3 contiguous and 4-element aligned (vector epilogue)
};

struct PlaneArgs {
  const u32x4* A = nullptr;
  const u32x4* B = nullptr;
  int nbA = 1, nbB = 1, RpA = 0, RpB = 0, KQ = 0;
  const double* invA = nullptr;  // FP16 form: 1 / row scale of A (b RpA + m) and of B (b RpB + n)
  const double* invB = nullptr;
};

// Epilogue of the plane GEMMs: one wave's MB x NB fragments (rows m0w + 16 i + lr,
// columns n0w + 16 j + 4 lk + r) times alpha (FP16: and the inverse row scales of
// A and B) into the split-K partials, or into C with beta.
template <int MB, int NB, bool F16>
__device__ __forceinline__ void planes_store(const Params& p, const PlaneArgs& pa, const f4 (&acc)[MB][NB], int m0w,
                                             int n0w, int lr, int lk, int split, int b) {
  const GemmDesc& d = p.d;
  auto ncol = [&](int j, int r) { return n0w + 16 * j + 4 * lk + r; };
  // FP16: alpha / (s_row s_col), a power of two times alpha
  const double* invA = F16 ? pa.invA + int64_t(pa.nbA > 1 ? b : 0) * pa.RpA : nullptr;
  const double* invB = F16 ? pa.invB + int64_t(pa.nbB > 1 ? b : 0) * pa.RpB : nullptr;
  double colf[NB][4];
#pragma unroll
  for (int j = 0; j < NB; ++j)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int n = ncol(j, r);
      colf[j][r] = F16 ? (n < d.N ? invB[n] : 0.0) * d.alpha : d.alpha;
    }
  if (p.out) {
    double* W = p.out + (int64_t(split) * d.batch + b) * int64_t(d.M) * d.N;
#pragma unroll
    for (int i = 0; i < MB; ++i) {
      const int m = m0w + 16 * i + lr;
      if (m >= d.M) continue;
      const double rf = F16 ? invA[m] : 1.0;
#pragma unroll
      for (int j = 0; j < NB; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int n = ncol(j, r);
          if (n < d.N) W[m + int64_t(d.M) * n] = colf[j][r] * rf * double(acc[i][j][r]);
        }
    }
    return;
  }
  const int64_t cb = d.Cb * b;
  // a lane's four columns of fragment j contiguous in C: one 16-byte access (FP32) or two (FP64)
  auto vec = [&](int j) { return p.cvec && ncol(j, 3) < d.N; };
#pragma unroll
  for (int i = 0; i < MB; ++i) {
    const int m = m0w + 16 * i + lr;
    if (m >= d.M) continue;
    const int64_t mo = d.Cm.off(m);
    const double rf = F16 ? invA[m] : 1.0;
    // beta != 0: the old values of the row group are all loaded before any of
    // its stores, so that a wave waits on memory once per row group instead of
    // once per fragment
    double old[NB][4];
    if (d.beta != 0.0) {
#pragma unroll
      for (int j = 0; j < NB; ++j) {
        if (vec(j)) {
          const int64_t off = cb + mo + d.Cn.off(ncol(j, 0));
          if (d.c32) {
            const float4 x = *reinterpret_cast<const float4*>(reinterpret_cast<const float*>(d.C) + off);
            old[j][0] = x.x, old[j][1] = x.y, old[j][2] = x.z, old[j][3] = x.w;
          } else {
            const double2* c = reinterpret_cast<const double2*>(d.C + off);
            const double2 x0 = c[0], x1 = c[1];
            old[j][0] = x0.x, old[j][1] = x0.y, old[j][2] = x1.x, old[j][3] = x1.y;
          }
        } else {
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int n = ncol(j, r);
            const int64_t off = cb + mo + d.Cn.off(n < d.N ? n : 0);
            old[j][r] = n >= d.N ? 0.0 : d.c32 ? double(reinterpret_cast<const float*>(d.C)[off]) : d.C[off];
          }
        }
      }
    }
#pragma unroll
    for (int j = 0; j < NB; ++j) {
      double v[4];
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        v[r] = colf[j][r] * rf * double(acc[i][j][r]);
        if (d.beta != 0.0) v[r] = v[r] + d.beta * old[j][r];
      }
      if (vec(j)) {
        const int64_t off = cb + mo + d.Cn.off(ncol(j, 0));
        if (d.c32)
          *reinterpret_cast<float4*>(reinterpret_cast<float*>(d.C) + off) =
              float4{float(v[0]), float(v[1]), float(v[2]), float(v[3])};
        else {
          double2* c = reinterpret_cast<double2*>(d.C + off);
          c[0] = double2{v[0], v[1]};
          c[1] = double2{v[2], v[3]};
        }
        continue;
      }
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int n = ncol(j, r);
        if (n >= d.N) continue;
        const int64_t off = cb + mo + d.Cn.off(n);
        if (d.c32) reinterpret_cast<float*>(d.C)[off] = float(v[r]);
        else d.C[off] = v[r];
      }
    }
  }
}

// Tile, split-K part and batch of this workgroup (blockIdx.x in XCD-major order
// when p.per > 0); false for the padding workgroups past p.work.
__device__ __forceinline__ bool plane_tile(const Params& p, int& tm, int& tn, int& split, int& b) {
  int64_t w = blockIdx.x;
  if (p.per > 0) {
    w = int64_t(blockIdx.x & 7) * p.per + (blockIdx.x >> 3);
    if (w >= p.work) return false;
  }
  const int tiles = p.tiles_m * p.tiles_n;
  const int tile = int(w % tiles);
  const int64_t q = w / tiles;
  split = int(q % p.splits);
  b = int(q / p.splits);
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
  return true;
}

// gemm_planes_wide_kernel for the two-plane kinds (2: BF16x3, 4: FP16x3);
// RpA and RpB multiples of kWideTile.
void launch_planes_wide(const Params& p, const PlaneArgs& pa, int kind, dim3 grid, hipStream_t s);
// Workgroups of gemm_planes_wide_kernel resident on the device at once.
int64_t wide_planes_slots(int kind);

}  // namespace uw12g
