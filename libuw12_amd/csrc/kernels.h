// Host launchers for the bandwidth-bound / reduction kernels (kernels.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace uw12g {

constexpr int kMaxPartials = 4096;  // scratch doubles needed by every reduction

// sum_i x[i] * (y ? y[i] : x[i]) * (row_w ? row_w[i % ld] : 1) * (col_w ? col_w[(i / ld) % col_n] : 1)
// sum_i x(i) y(i) row_w[i % ld] col_w[(i / ld) % col_n] over i < n; element i of
// x sits at x[i % ld + (i / ld) xs] when xs > 0 (else x[i]), likewise y with ys.
struct DotDesc {
  const double* x = nullptr;
  const double* y = nullptr;
  int64_t n = 0;
  const double* row_w = nullptr;
  const double* col_w = nullptr;
  int64_t ld = 1;
  int64_t col_n = 1;
  int64_t xs = 0, ys = 0;
};

void assemble_gamma(const double* X3p, const double* X3ri, double* G, int N, int Nr, int na, hipStream_t s);
void gather_scale_cols(const double* Q, int n, const int* keep, const double* scale, double* T, int nk,
                       hipStream_t s);
// y[c][A] = sum_r G[r + ldg A] x[c][r] (nvec <= 2 vectors); `partial` holds at
// least min(partial_cap, ...) doubles of per-block sums, added in a fixed order.
void gemv_t(const double* G, int64_t rows, int64_t ldg, const double* x, int64_t ldx, int nvec, double* y,
            int64_t ldy, int ncols, double* partial, int64_t partial_cap, hipStream_t s);
// y[c][r] += sum_A G[r + ldg A] v[c][A] (nvec <= 2); `part` holds the per-column-chunk
// partial rows (gemv_n_scratch doubles), added in a fixed order.
void gemv_n(const double* G, int64_t rows, int64_t ldg, const double* v, int64_t ldv, int nvec, double* y,
            int64_t ldy, int ncols, double* part, int64_t part_cap, hipStream_t s);
int64_t gemv_n_scratch(int64_t rows, int nvec, int ncols);
// E += sum_{s,t} 1/2 f[s][t] sum_A XD[A,s] sg[A] XD[A,t];  v[A,s] = sum_t 1/2 ns f[s][t] sg[A] XD[A,t]
// (v may be null).  One block: D is the DF dimension.
void xd_combine(const double* XD, int D, int ns, const double f[2][2], const double* sg, double* v, double* E,
                hipStream_t s);
// *out = (accumulate ? *out : 0) + scale * dot
void dot(const DotDesc& d, double scale, double* out, int accumulate, double* partial, hipStream_t s);
void e4i_energy(const double* W4, const double* V4, int a, int o, double scale, double* out, double* partial,
                hipStream_t s);
void e3i_energy(const double* GW, const double* GV, int M, int a, int nk, double scale, double* out,
                double* partial, hipStream_t s);
void combine_dd(int nW, int nV, const double* t0, const double* t1, const double* x0, const double* x1, double c0,
                double c1, const double* sW, const double* sV, double* Tb, double* Z, hipStream_t s);
void symmetrise_add(const double* F, double* out, int N, int accumulate, hipStream_t s);
void scale_cols(const double* in, int64_t rows, int64_t cols, const double* sc, double* out, hipStream_t s);
// out = (in diag(sc))^T: in is rows x cols (column-major), out is cols x rows.
void transpose_scale(const double* in, int64_t rows, int64_t cols, const double* sc, double* out, hipStream_t s);
// out[m + rows (i + ni b)] = d[m] cb[b] in[m + si i + sb b] for m < rows, i < ni, b < nb
// (cb == nullptr: 1).  out may be in itself when si = rows and sb = rows ni.
void scale_rows(const double* in, int64_t rows, int64_t ni, int64_t nb, int64_t si, int64_t sb, const double* d,
                const double* cb, double* out, hipStream_t s);
// Columns of one block of occupied pairs (j, k), j in [j0, j0 + nj) active, k in
// [0, c) (core) and [c + j0, c + a) (active from j0 on): out[r + rows (kk + nk jj)]
// = in[r + rows (j + a k)] with j = j0 + jj, k = kk (kk < c) or kk + j0, nk = c + a - j0.
void pair_gather(const double* in, int64_t rows, int a, int c, int j0, int nj, int nk, double* out, hipStream_t s);
void fill(double* p, int64_t n, double v, hipStream_t s);
// out[r] = || x[r + stride t], t < nt ||_2 for r < rows.
void row_norms(const double* x, int64_t rows, int64_t stride, int nt, double* out, hipStream_t s);
// x column-major n0 x n1: out[i] = max_j x[i + n0 j] (over_cols) or out[j] = max_i x[i + n0 j].
void max_over(const double* x, int64_t n0, int64_t n1, bool over_cols, double* out, hipStream_t s);
// out[i + nu j] = c u[i] v[j].
// out[i] = max_j |x[i + n0 j]| (a column-major n0 x n1 matrix, i < n0).
void abs_max_rows(const double* x, int64_t n0, int64_t n1, double* out, hipStream_t s);
void outer(const double* u, int64_t nu, const double* v, int64_t nv, double c, double* out, hipStream_t s);

}  // namespace uw12g
