// Bandwidth-bound and reduction kernels of the UW12 evaluation (FP64).
// All reductions run in a fixed order (per-block partials, then one block),
// so results are bit-reproducible run to run.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "kernels.h"
#include "uw12g_internal.h"

namespace uw12g {
namespace {

constexpr int RB = 256;  // reduction block size

__device__ inline double block_sum(double v) {
  __shared__ double sh[RB / 64];
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) sh[w] = v;
  __syncthreads();
  double r = 0.0;
  if (threadIdx.x == 0)
    for (int i = 0; i < RB / 64; ++i) r += sh[i];
  __syncthreads();
  return r;
}

// Gamma AO+RI assembly for a block of DF columns [a0, a0+na):
// G[mu' + M alpha + M N j] = X3 (packed, mu' < N) or X3ri[r + Nr alpha] (mu' = N + r).
__global__ void assemble_gamma_kernel(const double* __restrict__ X3p, const double* __restrict__ X3ri,
                                      double* __restrict__ G, int N, int Nr, int na) {
  const int M = N + Nr;
  const int64_t per = int64_t(M) * N;
  const int64_t total = per * na;
  for (int64_t e = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; e < total;
       e += int64_t(gridDim.x) * blockDim.x) {
    const int j = int(e / per);
    const int64_t rem = e - int64_t(j) * per;
    const int alpha = int(rem / M), mu = int(rem - int64_t(alpha) * M);
    double v;
    if (mu < N) {
      const int hi = max(mu, alpha), lo = min(mu, alpha);
      v = X3p[int64_t(hi) * (hi + 1) / 2 + lo + int64_t(N) * (N + 1) / 2 * j];
    } else {
      v = X3ri[(mu - N) + int64_t(Nr) * alpha + int64_t(Nr) * N * j];
    }
    G[e] = v;
  }
}

// T[:, j] = Q[:, keep[j]] * scale[j]
__global__ void gather_scale_cols_kernel(const double* Q, int n, const int* keep, const double* scale,
                                         double* T, int nk) {
  const int64_t total = int64_t(n) * nk;
  for (int64_t e = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; e < total;
       e += int64_t(gridDim.x) * blockDim.x) {
    const int j = int(e / n), i = int(e - int64_t(j) * n);
    T[e] = Q[i + int64_t(n) * keep[j]] * scale[j];
  }
}

// y[c][A] = sum_r G[r + ldg A] x[c][r] for nvec <= 2 vectors.  Block (A, part)
// sums rows [part * rpb, (part + 1) * rpb) of column A into
// partial[part + parts (A + ncols c)]; gemv_t_finish adds the parts in order.
// Each lane streams two consecutive rows per load (16 bytes) with four loads
// in flight.  The rows are read once per pass; x (both vectors) stays in L2.
constexpr int kGemvUnroll = 4;
__global__ __launch_bounds__(RB) void gemv_t_kernel(const double* __restrict__ G, int64_t rows, int64_t ldg,
                                                    const double* __restrict__ x, int64_t ldx, int nvec,
                                                    double* __restrict__ partial, int parts, int64_t rpb,
                                                    int ncols) {
  const int A = blockIdx.x, part = blockIdx.y;
  const double* g = G + ldg * A;
  const int64_t r0 = int64_t(part) * rpb, r1 = min(rows, r0 + rpb);
  double acc0 = 0.0, acc1 = 0.0;
  const bool aligned = ((reinterpret_cast<uintptr_t>(g + r0) | reinterpret_cast<uintptr_t>(x + r0) |
                         (nvec > 1 ? reinterpret_cast<uintptr_t>(x + ldx + r0) : 0)) & 15) == 0;
  // row pairs (r0 + 2p, r0 + 2p + 1), p < npair, as 16-byte loads, kGemvUnroll per lane in flight
  const int64_t npair = aligned ? (r1 - r0) / 2 : 0;
  for (int64_t p = threadIdx.x; p < npair; p += int64_t(RB) * kGemvUnroll) {
    double2 gv[kGemvUnroll];
#pragma unroll
    for (int u = 0; u < kGemvUnroll; ++u) {
      const int64_t pp = p + int64_t(RB) * u;
      gv[u] = pp < npair ? *reinterpret_cast<const double2*>(g + r0 + 2 * pp) : make_double2(0.0, 0.0);
    }
#pragma unroll
    for (int u = 0; u < kGemvUnroll; ++u) {
      const int64_t pp = p + int64_t(RB) * u;
      if (pp < npair) {
        const double2 xv = *reinterpret_cast<const double2*>(x + r0 + 2 * pp);
        acc0 += gv[u].x * xv.x + gv[u].y * xv.y;
        if (nvec > 1) {
          const double2 yv = *reinterpret_cast<const double2*>(x + ldx + r0 + 2 * pp);
          acc1 += gv[u].x * yv.x + gv[u].y * yv.y;
        }
      }
    }
  }
  // the rows not covered by pairs (an odd last row, or every row of unaligned operands)
  for (int64_t q = r0 + 2 * npair + threadIdx.x; q < r1; q += RB) {
    const double gv = g[q];
    acc0 += gv * x[q];
    if (nvec > 1) acc1 += gv * x[ldx + q];
  }
  const double s0 = block_sum(acc0);
  const double s1 = nvec > 1 ? block_sum(acc1) : 0.0;
  if (threadIdx.x == 0) {
    partial[part + int64_t(parts) * A] = s0;
    if (nvec > 1) partial[part + int64_t(parts) * (A + ncols)] = s1;
  }
}

__global__ void gemv_t_finish_kernel(const double* __restrict__ partial, int parts, int nvec, int ncols,
                                     double* __restrict__ y, int64_t ldy) {
  for (int64_t e = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; e < int64_t(ncols) * nvec;
       e += int64_t(gridDim.x) * blockDim.x) {
    const int c = int(e / ncols), A = int(e - int64_t(c) * ncols);
    const double* p = partial + int64_t(parts) * (A + int64_t(ncols) * c);
    double s = 0.0;
    for (int q = 0; q < parts; ++q) s += p[q];
    y[ldy * c + A] = s;
  }
}

// Partial y over a chunk of columns: part[(chunk, c)][r] = sum_{A in chunk}
// G[r + ldg A] v[c][A].  Block (row tile, chunk); each lane owns two
// consecutive rows (16-byte loads) and runs the chunk's columns four at a time,
// so eight loads per lane are in flight.  gemv_n_finish adds the chunks in
// order into y.
constexpr int kGemvRowsPerBlock = 2 * RB;
__global__ __launch_bounds__(RB) void gemv_n_kernel(const double* __restrict__ G, int64_t rows, int64_t ldg,
                                                    const double* __restrict__ v, int64_t ldv, int nvec,
                                                    double* __restrict__ part, int ncols, int cpb) {
  const int64_t r = int64_t(blockIdx.x) * kGemvRowsPerBlock + 2 * threadIdx.x;
  const int chunk = blockIdx.y;
  const int A0 = chunk * cpb, A1 = min(ncols, A0 + cpb);
  double a0x = 0.0, a0y = 0.0, a1x = 0.0, a1y = 0.0;
  if (r + 1 < rows && ((reinterpret_cast<uintptr_t>(G + r) | (uintptr_t(ldg) * 8)) & 15) == 0) {
    int A = A0;
    for (; A + kGemvUnroll <= A1; A += kGemvUnroll) {
      double2 gv[kGemvUnroll];
#pragma unroll
      for (int u = 0; u < kGemvUnroll; ++u) gv[u] = *reinterpret_cast<const double2*>(G + r + ldg * (A + u));
#pragma unroll
      for (int u = 0; u < kGemvUnroll; ++u) {
        const double w0 = v[A + u];
        a0x += gv[u].x * w0;
        a0y += gv[u].y * w0;
        if (nvec > 1) {
          const double w1 = v[ldv + A + u];
          a1x += gv[u].x * w1;
          a1y += gv[u].y * w1;
        }
      }
    }
    for (; A < A1; ++A) {
      const double2 gv = *reinterpret_cast<const double2*>(G + r + ldg * A);
      a0x += gv.x * v[A];
      a0y += gv.y * v[A];
      if (nvec > 1) {
        a1x += gv.x * v[ldv + A];
        a1y += gv.y * v[ldv + A];
      }
    }
  } else {
    for (int A = A0; A < A1; ++A) {
      if (r < rows) {
        const double g0 = G[r + ldg * A];
        a0x += g0 * v[A];
        if (nvec > 1) a1x += g0 * v[ldv + A];
      }
      if (r + 1 < rows) {
        const double g1 = G[r + 1 + ldg * A];
        a0y += g1 * v[A];
        if (nvec > 1) a1y += g1 * v[ldv + A];
      }
    }
  }
  double* p = part + rows * (int64_t(chunk) * nvec);
  if (r < rows) {
    p[r] = a0x;
    if (nvec > 1) p[rows + r] = a1x;
  }
  if (r + 1 < rows) {
    p[r + 1] = a0y;
    if (nvec > 1) p[rows + r + 1] = a1y;
  }
}

__global__ void gemv_n_finish_kernel(const double* __restrict__ part, int64_t rows, int nvec, int chunks,
                                     double* __restrict__ y, int64_t ldy) {
  for (int64_t e = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; e < rows * nvec;
       e += int64_t(gridDim.x) * blockDim.x) {
    const int c = int(e / rows);
    const int64_t r = e - int64_t(c) * rows;
    double s = 0.0;
    for (int q = 0; q < chunks; ++q) s += part[r + rows * (int64_t(q) * nvec + c)];
    y[ldy * c + r] += s;
  }
}

// Two-electron direct term from X_D (two_electron.cpp:11-81), on one block:
//   E += sum_{s,t} 1/2 f[s][t] sum_A X_D[A, s] sg[A] X_D[A, t]
//   v[A, s] = sum_t 1/2 n_spin f[s][t] sg[A] X_D[A, t]   (the Fock's DF vector)
__global__ __launch_bounds__(RB) void xd_combine_kernel(const double* __restrict__ XD, int D, int ns, double f00,
                                                        double f01, double f10, double f11,
                                                        const double* __restrict__ sg, double* __restrict__ v,
                                                        double* __restrict__ E) {
  const double f[2][2] = {{f00, f01}, {f10, f11}};
  double d[2][2] = {{0.0, 0.0}, {0.0, 0.0}};
  for (int A = threadIdx.x; A < D; A += RB) {
    const double x0 = XD[A], x1 = ns > 1 ? XD[D + A] : 0.0, w = sg[A];
    d[0][0] += x0 * w * x0;
    if (ns > 1) {
      d[0][1] += x0 * w * x1;
      d[1][0] += x1 * w * x0;
      d[1][1] += x1 * w * x1;
    }
    if (v) {
      v[A] = 0.5 * ns * w * (f[0][0] * x0 + (ns > 1 ? f[0][1] * x1 : 0.0));
      if (ns > 1) v[D + A] = 0.5 * ns * w * (f[1][0] * x0 + f[1][1] * x1);
    }
  }
  double e = 0.0;
  for (int s = 0; s < ns; ++s)
    for (int t = 0; t < ns; ++t) {
      const double sum = block_sum(d[s][t]);
      e += 0.5 * f[s][t] * sum;
    }
  if (threadIdx.x == 0) *E += e;
}

// partial[b] = sum over this block's share of  x[i] * y[i] * w(i)
//   w(i) = (ws ? ws[(i / wdiv) % wmod] : 1) * (vs ? vs[(i % vmodr) ...]) -- see dot_desc
__global__ void dot_kernel(DotDesc d, double* partial) {
  double acc = 0.0;
  for (int64_t i = blockIdx.x * int64_t(RB) + threadIdx.x; i < d.n; i += int64_t(gridDim.x) * RB) {
    int64_t ix = i, iy = i;
    if (d.xs > 0 || d.ys > 0) {
      const int64_t col = i / d.ld, r = i - col * d.ld;
      if (d.xs > 0) ix = r + col * d.xs;
      if (d.ys > 0) iy = r + col * d.ys;
    }
    double v = d.x[ix] * (d.y ? d.y[iy] : d.x[ix]);
    if (d.row_w) v *= d.row_w[i % d.ld];
    if (d.col_w) v *= d.col_w[(i / d.ld) % d.col_n];
    acc += v;
  }
  const double s = block_sum(acc);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

__global__ void finish_sum_kernel(const double* partial, int n, double scale, double* out, int accumulate) {
  double acc = 0.0;
  for (int i = threadIdx.x; i < n; i += RB) acc += partial[i];
  const double s = block_sum(acc);
  if (threadIdx.x == 0) *out = accumulate ? *out + scale * s : scale * s;
}

// 4e indirect: sum_{i,j<a; k,l<o} W4[(i + a l) + ao (j + a k)] V4[(i + a k) + ao (j + a l)]
__global__ void e4i_kernel(const double* __restrict__ W4, const double* __restrict__ V4, int a, int o,
                           double* partial) {
  const int64_t ao = int64_t(a) * o;
  const int64_t total = ao * ao;
  double acc = 0.0;
  for (int64_t e = blockIdx.x * int64_t(RB) + threadIdx.x; e < total; e += int64_t(gridDim.x) * RB) {
    const int64_t row = e % ao, col = e / ao;  // V4 element (i + a k, j + a l)
    const int i = int(row % a), k = int(row / a);
    const int j = int(col % a), l = int(col / a);
    acc += W4[(i + int64_t(a) * l) + ao * (j + int64_t(a) * k)] * V4[e];
  }
  const double s = block_sum(acc);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

// 3e indirect energy: sum GW[mu + M i + M a (j + a kk)] GV[mu + M j + M a (i + a kk)]
// over mu < M, i, j < a and kk in [0, nk) (a k-block of the occupied index).
__global__ void e3i_kernel(const double* __restrict__ GW, const double* __restrict__ GV, int M, int a, int nk,
                           double* partial) {
  const int64_t Ma = int64_t(M) * a;
  const int64_t total = Ma * a * nk;
  double acc = 0.0;
  for (int64_t e = blockIdx.x * int64_t(RB) + threadIdx.x; e < total; e += int64_t(gridDim.x) * RB) {
    const int64_t r = e % Ma, c = e / Ma;  // GW row (mu, i), col (j, kk)
    const int mu = int(r % M), i = int(r / M);
    const int j = int(c % a), kk = int(c / a);
    acc += GW[e] * GV[mu + int64_t(M) * j + Ma * (i + int64_t(a) * kk)];
  }
  const double s = block_sum(acc);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

// Direct-term matrices on the DF x DF grid:
//   Tb = sW sV sum_u c_u t_u ;  Z = Tb - sW sV sum_u c_u X_u
__global__ void combine_dd_kernel(int nW, int nV, const double* t0, const double* t1, const double* x0,
                                  const double* x1, double c0, double c1, const double* sW, const double* sV,
                                  double* Tb, double* Z) {
  const int64_t total = int64_t(nW) * nV;
  for (int64_t e = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; e < total;
       e += int64_t(gridDim.x) * blockDim.x) {
    const int A = int(e % nW), B = int(e / nW);
    const double sg = sW[A] * sV[B];
    double tb = c0 * t0[e];
    double xb = c0 * x0[e];
    if (t1) {
      tb += c1 * t1[e];
      xb += c1 * x1[e];
    }
    tb *= sg;
    Tb[e] = tb;
    Z[e] = tb - sg * xb;
  }
}

// F_out[s] (+)= 0.5 (F + F^T)
__global__ void symmetrise_add_kernel(const double* F, double* out, int N, int accumulate) {
  const int64_t total = int64_t(N) * N;
  for (int64_t e = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; e < total;
       e += int64_t(gridDim.x) * blockDim.x) {
    const int i = int(e % N), j = int(e / N);
    const double v = 0.5 * (F[e] + F[j + int64_t(N) * i]);
    out[e] = accumulate ? out[e] + v : v;
  }
}

// out[r + rows c] = in[r + rows c] * sc[c]
__global__ void scale_cols_kernel(const double* __restrict__ in, int64_t rows, int64_t cols,
                                  const double* __restrict__ sc, double* __restrict__ out) {
  const int64_t total = rows * cols;
  for (int64_t e = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; e < total;
       e += int64_t(gridDim.x) * blockDim.x)
    out[e] = in[e] * sc[e / rows];
}

// One block per output column (i, b), the threads along the rows: reads and
// writes are both contiguous runs of `rows` doubles.
// in and out may be the same array (in place: each element is read, then written,
// by one thread).
__global__ __launch_bounds__(256) void scale_rows_kernel(const double* in, int64_t rows, int64_t ni, int64_t ncol,
                                                         int64_t si, int64_t sb, const double* __restrict__ d,
                                                         const double* __restrict__ cb, double* out) {
  for (int64_t col = blockIdx.x; col < ncol; col += gridDim.x) {
    const int64_t i = col % ni, b = col / ni;
    const double c = cb ? cb[b] : 1.0;
    const double* src = in + si * i + sb * b;
    double* dst = out + rows * col;
    for (int64_t m = threadIdx.x; m < rows; m += blockDim.x) dst[m] = c * d[m] * src[m];
  }
}

// Column q = kk + nk jj of one block of occupied pairs: j = j0 + jj, k = kk for
// kk < c (core), else kk + j0; out[r + rows q] = in[r + rows (j + a k)].  One
// block per column, the rows along the threads.
__global__ __launch_bounds__(256) void pair_gather_kernel(const double* __restrict__ in, int64_t rows, int a, int c,
                                                          int j0, int nk, int64_t ncol, double* __restrict__ out) {
  for (int64_t q = blockIdx.x; q < ncol; q += gridDim.x) {
    const int kk = int(q % nk), jj = int(q / nk);
    const int k = kk < c ? kk : kk + j0;
    const double* src = in + rows * (int64_t(j0 + jj) + int64_t(a) * k);
    double* dst = out + rows * q;
    for (int64_t r = threadIdx.x; r < rows; r += blockDim.x) dst[r] = src[r];
  }
}

// out[c + cols r] = in[r + rows c] sc[c]: 64 x 64 tiles through LDS, so both the
// reads (along r) and the writes (along c) are coalesced.
__global__ __launch_bounds__(256) void transpose_scale_kernel(const double* __restrict__ in, int64_t rows,
                                                              int64_t cols, const double* __restrict__ sc,
                                                              double* __restrict__ out) {
  __shared__ double tile[64][65];
  const int64_t r0 = int64_t(blockIdx.x) * 64, c0 = int64_t(blockIdx.y) * 64;
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;  // 64 x 4 threads
  for (int j = ty; j < 64; j += 4) {
    const int64_t r = r0 + tx, c = c0 + j;
    tile[j][tx] = (r < rows && c < cols) ? in[r + rows * c] * sc[c] : 0.0;
  }
  __syncthreads();
  for (int j = ty; j < 64; j += 4) {
    const int64_t c = c0 + tx, r = r0 + j;
    if (r < rows && c < cols) out[c + cols * r] = tile[tx][j];
  }
}

__global__ void fill_kernel(double* p, int64_t n, double v) {
  for (int64_t e = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; e < n; e += int64_t(gridDim.x) * blockDim.x)
    p[e] = v;
}

unsigned grid_for(int64_t n, int64_t cap = 8192) {
  return unsigned(std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, cap)));
}

}  // namespace

void assemble_gamma(const double* X3p, const double* X3ri, double* G, int N, int Nr, int na, hipStream_t s) {
  const int64_t total = int64_t(N + Nr) * N * na;
  hipLaunchKernelGGL(assemble_gamma_kernel, dim3(grid_for(total, 65536)), dim3(256), 0, s, X3p, X3ri, G, N, Nr, na);
  UW12G_HIP(hipGetLastError());
}

void gather_scale_cols(const double* Q, int n, const int* keep, const double* scale, double* T, int nk,
                       hipStream_t s) {
  hipLaunchKernelGGL(gather_scale_cols_kernel, dim3(grid_for(int64_t(n) * nk)), dim3(256), 0, s, Q, n, keep, scale,
                     T, nk);
  UW12G_HIP(hipGetLastError());
}

void gemv_t(const double* G, int64_t rows, int64_t ldg, const double* x, int64_t ldx, int nvec, double* y,
            int64_t ldy, int ncols, double* partial, int64_t partial_cap, hipStream_t s) {
  if (ncols <= 0) return;
  if (nvec > 2) throw Error(-2, "gemv_t supports at most 2 vectors");
  // enough blocks to fill the chip several times over: columns x row parts
  int parts = int(std::max<int64_t>(1, std::min<int64_t>((4096 + ncols - 1) / ncols,
                                                         (rows + 2 * RB * kGemvUnroll - 1) / (2 * RB * kGemvUnroll))));
  while (parts > 1 && int64_t(parts) * ncols * nvec > partial_cap) --parts;
  if (int64_t(parts) * ncols * nvec > partial_cap) throw Error(-2, "gemv_t: partial buffer too small");
  // row parts of whole 16-byte pairs
  const int64_t rpb = ((rows + parts - 1) / parts + 1) / 2 * 2;
  parts = int((rows + rpb - 1) / rpb);
  if (parts < 1) parts = 1;
  if (parts > 65535) throw Error(-2, "gemv_t: too many row parts");
  hipLaunchKernelGGL(gemv_t_kernel, dim3(unsigned(ncols), unsigned(parts)), dim3(RB), 0, s, G, rows, ldg, x, ldx,
                     nvec, partial, parts, rpb, ncols);
  hipLaunchKernelGGL(gemv_t_finish_kernel, dim3(grid_for(int64_t(ncols) * nvec)), dim3(256), 0, s, partial, parts,
                     nvec, ncols, y, ldy);
  UW12G_HIP(hipGetLastError());
}

int64_t gemv_n_scratch(int64_t rows, int nvec, int ncols) {
  const int64_t tiles = (rows + kGemvRowsPerBlock - 1) / kGemvRowsPerBlock;
  const int chunks = int(std::max<int64_t>(1, std::min<int64_t>((2048 + tiles - 1) / tiles, (ncols + 15) / 16)));
  return rows * nvec * chunks;
}

void gemv_n(const double* G, int64_t rows, int64_t ldg, const double* v, int64_t ldv, int nvec, double* y,
            int64_t ldy, int ncols, double* part, int64_t part_cap, hipStream_t s) {
  if (rows <= 0 || ncols <= 0) return;
  if (nvec > 2) throw Error(-2, "gemv_n supports at most 2 vectors");
  const int64_t tiles = (rows + kGemvRowsPerBlock - 1) / kGemvRowsPerBlock;
  if (tiles > 0x7fffffff) throw Error(-2, "gemv_n: too many rows");
  // column chunks: about 2048 blocks in all, at least 16 columns per chunk
  int chunks = int(std::max<int64_t>(1, std::min<int64_t>((2048 + tiles - 1) / tiles, (ncols + 15) / 16)));
  while (chunks > 1 && rows * nvec * chunks > part_cap) --chunks;
  if (rows * nvec * chunks > part_cap) throw Error(-2, "gemv_n: partial buffer too small");
  const int cpb = (ncols + chunks - 1) / chunks;
  chunks = (ncols + cpb - 1) / cpb;
  hipLaunchKernelGGL(gemv_n_kernel, dim3(unsigned(tiles), unsigned(chunks)), dim3(RB), 0, s, G, rows, ldg, v, ldv,
                     nvec, part, ncols, cpb);
  hipLaunchKernelGGL(gemv_n_finish_kernel, dim3(grid_for(rows * nvec, 1 << 16)), dim3(256), 0, s, part, rows, nvec,
                     chunks, y, ldy);
  UW12G_HIP(hipGetLastError());
}

void xd_combine(const double* XD, int D, int ns, const double f[2][2], const double* sg, double* v, double* E,
                hipStream_t s) {
  hipLaunchKernelGGL(xd_combine_kernel, dim3(1), dim3(RB), 0, s, XD, D, ns, f[0][0], f[0][1], f[1][0], f[1][1], sg,
                     v, E);
  UW12G_HIP(hipGetLastError());
}

void dot(const DotDesc& d, double scale, double* out, int accumulate, double* partial, hipStream_t s) {
  const int nb = int(std::min<int64_t>((d.n + RB - 1) / RB, kMaxPartials));
  if (d.n <= 0) {
    hipLaunchKernelGGL(fill_kernel, dim3(1), dim3(64), 0, s, partial, int64_t(1), 0.0);
    hipLaunchKernelGGL(finish_sum_kernel, dim3(1), dim3(RB), 0, s, partial, 1, scale, out, accumulate);
  } else {
    hipLaunchKernelGGL(dot_kernel, dim3(nb), dim3(RB), 0, s, d, partial);
    hipLaunchKernelGGL(finish_sum_kernel, dim3(1), dim3(RB), 0, s, partial, nb, scale, out, accumulate);
  }
  UW12G_HIP(hipGetLastError());
}

void e4i_energy(const double* W4, const double* V4, int a, int o, double scale, double* out, double* partial,
                hipStream_t s) {
  const int64_t total = int64_t(a) * o * a * o;
  const int nb = int(std::max<int64_t>(1, std::min<int64_t>((total + RB - 1) / RB, kMaxPartials)));
  hipLaunchKernelGGL(e4i_kernel, dim3(nb), dim3(RB), 0, s, W4, V4, a, o, partial);
  hipLaunchKernelGGL(finish_sum_kernel, dim3(1), dim3(RB), 0, s, partial, nb, scale, out, 1);
  UW12G_HIP(hipGetLastError());
}

void e3i_energy(const double* GW, const double* GV, int M, int a, int nk, double scale, double* out,
                double* partial, hipStream_t s) {
  const int64_t total = int64_t(M) * a * a * nk;
  const int nb = int(std::max<int64_t>(1, std::min<int64_t>((total + RB - 1) / RB, kMaxPartials)));
  hipLaunchKernelGGL(e3i_kernel, dim3(nb), dim3(RB), 0, s, GW, GV, M, a, nk, partial);
  hipLaunchKernelGGL(finish_sum_kernel, dim3(1), dim3(RB), 0, s, partial, nb, scale, out, 1);
  UW12G_HIP(hipGetLastError());
}

void combine_dd(int nW, int nV, const double* t0, const double* t1, const double* x0, const double* x1, double c0,
                double c1, const double* sW, const double* sV, double* Tb, double* Z, hipStream_t s) {
  hipLaunchKernelGGL(combine_dd_kernel, dim3(grid_for(int64_t(nW) * nV)), dim3(256), 0, s, nW, nV, t0, t1, x0, x1,
                     c0, c1, sW, sV, Tb, Z);
  UW12G_HIP(hipGetLastError());
}

void symmetrise_add(const double* F, double* out, int N, int accumulate, hipStream_t s) {
  hipLaunchKernelGGL(symmetrise_add_kernel, dim3(grid_for(int64_t(N) * N)), dim3(256), 0, s, F, out, N, accumulate);
  UW12G_HIP(hipGetLastError());
}

void transpose_scale(const double* in, int64_t rows, int64_t cols, const double* sc, double* out, hipStream_t s) {
  if (rows <= 0 || cols <= 0) return;
  const int64_t gr = (rows + 63) / 64, gc = (cols + 63) / 64;
  if (gr > 0x7fffffff || gc > 65535) throw Error(-2, "transpose_scale: matrix too large");
  hipLaunchKernelGGL(transpose_scale_kernel, dim3(unsigned(gr), unsigned(gc)), dim3(256), 0, s, in, rows, cols, sc,
                     out);
  UW12G_HIP(hipGetLastError());
}

void scale_cols(const double* in, int64_t rows, int64_t cols, const double* sc, double* out, hipStream_t s) {
  if (rows <= 0 || cols <= 0) return;
  hipLaunchKernelGGL(scale_cols_kernel, dim3(grid_for(rows * cols)), dim3(256), 0, s, in, rows, cols, sc, out);
  UW12G_HIP(hipGetLastError());
}

void scale_rows(const double* in, int64_t rows, int64_t ni, int64_t nb, int64_t si, int64_t sb, const double* d,
                const double* cb, double* out, hipStream_t s) {
  const int64_t ncol = ni * nb;
  if (rows <= 0 || ncol <= 0) return;
  const unsigned grid = unsigned(std::min<int64_t>(ncol, 65536));
  hipLaunchKernelGGL(scale_rows_kernel, dim3(grid), dim3(256), 0, s, in, rows, ni, ncol, si, sb, d, cb, out);
  UW12G_HIP(hipGetLastError());
}

void pair_gather(const double* in, int64_t rows, int a, int c, int j0, int nj, int nk, double* out, hipStream_t s) {
  const int64_t ncol = int64_t(nj) * nk;
  if (rows <= 0 || ncol <= 0) return;
  const unsigned grid = unsigned(std::min<int64_t>(ncol, 65536));
  hipLaunchKernelGGL(pair_gather_kernel, dim3(grid), dim3(256), 0, s, in, rows, a, c, j0, nk, ncol, out);
  UW12G_HIP(hipGetLastError());
}

// Euclidean norm of each row of a column-major rows x nt matrix with column stride `stride`.
__global__ void row_norms_kernel(const double* x, int64_t rows, int64_t stride, int nt, double* out) {
  const int64_t r = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (r >= rows) return;
  double acc = 0.0;
  for (int t = 0; t < nt; ++t) {
    const double v = x[r + stride * t];
    acc += v * v;
  }
  out[r] = sqrt(acc);
}

// Maxima of a column-major n0 x n1 matrix: over its columns (out[i], i < n0) or over its rows (out[j], j < n1).
__global__ void max_rows_kernel(const double* x, int64_t n0, int64_t n1, double* out) {
  const int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (i >= n0) return;
  double m = 0.0;
  for (int64_t j = 0; j < n1; ++j) m = fmax(m, x[i + n0 * j]);
  out[i] = m;
}
__global__ void max_cols_kernel(const double* x, int64_t n0, int64_t n1, double* out) {
  const int64_t j = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (j >= n1) return;
  double m = 0.0;
  for (int64_t i = 0; i < n0; ++i) m = fmax(m, x[i + n0 * j]);
  out[j] = m;
}

// Largest magnitude of each row of a column-major n0 x n1 matrix (out[i], i < n0).
__global__ void abs_max_rows_kernel(const double* x, int64_t n0, int64_t n1, double* out) {
  const int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (i >= n0) return;
  double m = 0.0;
  for (int64_t j = 0; j < n1; ++j) m = fmax(m, fabs(x[i + n0 * j]));
  out[i] = m;
}

__global__ void outer_kernel(const double* u, int64_t nu, const double* v, int64_t nv, double c, double* out) {
  const int64_t t = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (t >= nu * nv) return;
  out[t] = c * u[t % nu] * v[t / nu];
}

void row_norms(const double* x, int64_t rows, int64_t stride, int nt, double* out, hipStream_t s) {
  if (rows <= 0) return;
  hipLaunchKernelGGL(row_norms_kernel, dim3(unsigned((rows + 255) / 256)), dim3(256), 0, s, x, rows, stride, nt, out);
  UW12G_HIP(hipGetLastError());
}

void max_over(const double* x, int64_t n0, int64_t n1, bool over_cols, double* out, hipStream_t s) {
  const int64_t n = over_cols ? n0 : n1;
  if (n <= 0) return;
  if (over_cols)
    hipLaunchKernelGGL(max_rows_kernel, dim3(unsigned((n + 255) / 256)), dim3(256), 0, s, x, n0, n1, out);
  else
    hipLaunchKernelGGL(max_cols_kernel, dim3(unsigned((n + 255) / 256)), dim3(256), 0, s, x, n0, n1, out);
  UW12G_HIP(hipGetLastError());
}

void abs_max_rows(const double* x, int64_t n0, int64_t n1, double* out, hipStream_t s) {
  if (n0 <= 0) return;
  hipLaunchKernelGGL(abs_max_rows_kernel, dim3(unsigned((n0 + 255) / 256)), dim3(256), 0, s, x, n0, n1, out);
  UW12G_HIP(hipGetLastError());
}

void outer(const double* u, int64_t nu, const double* v, int64_t nv, double c, double* out, hipStream_t s) {
  const int64_t n = nu * nv;
  if (n <= 0) return;
  hipLaunchKernelGGL(outer_kernel, dim3(unsigned((n + 255) / 256)), dim3(256), 0, s, u, nu, v, nv, c, out);
  UW12G_HIP(hipGetLastError());
}

void fill(double* p, int64_t n, double v, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(fill_kernel, dim3(grid_for(n)), dim3(256), 0, s, p, n, v);
  UW12G_HIP(hipGetLastError());
}

}  // namespace uw12g
