// Synthetic s-Gaussian three-index integrals generated on the device.
//
// The same integrals as the host generator (synth.cpp): normalised s-type
// primitives g_a(r) = (2a/pi)^{3/4} exp(-a |r - A|^2), operators 1/r12
// (kind 0), the Gaussian geminal -sum_i c_i exp(-g_i r12^2) (kind 1) and the
// geminal over r12 (kind 2).  For the charge distribution of a function pair
// (exponent p = a + b, centre P, prefactor K) and a DF function (exponent q,
// centre Q), with al = pq / (p + q) and R = |P - Q| (Helgaker, Jorgensen &
// Olsen, Molecular Electronic-Structure Theory, ch. 9):
//   1/r12:           2 pi^{5/2} / (pq sqrt(p + q)) F0(al R^2)
//   e^{-g r^2}:      (pi^2 / ((p + q)(al + g)))^{3/2} exp(-al g R^2 / (al + g))
//   e^{-g r^2} / r:  (pi / (p + q))^{3/2} 2 pi / (al + g) exp(-al g R^2 / (al + g))
//                    F0(al^2 R^2 / (al + g))
// with F0(t) = 1/2 sqrt(pi / t) erf(sqrt t).
//
// uw12s_generate has the uw12g_three_index_gen_fn signature (include/uw12g.h):
// the library calls it for each block of DF functions it needs, and the
// kernel writes the block straight into the library's device buffer, on the
// library's stream.  This is input generation for tests and benchmarks (the
// role a quantum-chemistry integral package plays for a real caller).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <algorithm>
#include <vector>

namespace {

constexpr double kPi = 3.14159265358979323846;

struct Basis {  // device arrays
  const double* ex = nullptr;   // exponents
  const double* nrm = nullptr;  // (2a/pi)^{3/4}
  const double* cen = nullptr;  // 3 per function
  int n = 0;
};

struct Gen {
  int kind = 0;
  int ngem = 0;
  double c[8] = {}, g[8] = {};
  Basis ao, df, ri;
  int device = 0;
  std::vector<void*> allocs;
};

__device__ inline double boys0(double t) {
  if (t < 1e-12) return 1.0 - t / 3.0;
  const double st = sqrt(t);
  return 0.5 * sqrt(kPi / t) * erf(st);
}

struct GemPar {
  int n;
  double c[8], g[8];
};

// out[mu' + m (nu + n_ao j)] = (mu' nu | x | df0 + j) for mu' < m (AO rows, then
// RI rows when m > n_ao), nu < n_ao, j < n.  One thread per (mu', nu), looping
// over the block's DF functions: consecutive threads write consecutive rows.
__global__ __launch_bounds__(256) void synth_block_kernel(int kind, GemPar gem, const double* __restrict__ aex,
                                                          const double* __restrict__ anrm,
                                                          const double* __restrict__ acen, int n_ao,
                                                          const double* __restrict__ rex,
                                                          const double* __restrict__ rnrm,
                                                          const double* __restrict__ rcen, int m,
                                                          const double* __restrict__ qex,
                                                          const double* __restrict__ qnrm,
                                                          const double* __restrict__ qcen, int df0, int n,
                                                          double* __restrict__ out) {
  const int64_t pairs = int64_t(m) * n_ao;
  const int64_t e = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (e >= pairs) return;
  const int mu = int(e % m), nu = int(e / m);
  double a, na, A[3];
  if (mu < n_ao) {
    a = aex[mu];
    na = anrm[mu];
    for (int x = 0; x < 3; ++x) A[x] = acen[3 * mu + x];
  } else {
    const int r = mu - n_ao;
    a = rex[r];
    na = rnrm[r];
    for (int x = 0; x < 3; ++x) A[x] = rcen[3 * r + x];
  }
  const double b = aex[nu], nb = anrm[nu];
  const double B[3] = {acen[3 * nu], acen[3 * nu + 1], acen[3 * nu + 2]};
  const double p = a + b;
  double AB2 = 0.0, P[3];
  {
    // no fused multiply-add in the pair centre: (a A + b B) / p is then bitwise the
    // same for (mu, nu) and (nu, mu), so the AO block is exactly symmetric
#pragma clang fp contract(off)
    for (int x = 0; x < 3; ++x) {
      const double d = A[x] - B[x];
      AB2 += d * d;
      P[x] = (a * A[x] + b * B[x]) / p;
    }
  }
  const double K = na * nb * exp(-a * b / p * AB2);
  double* o = out + mu + int64_t(m) * nu;
  const int64_t stride = pairs;
  for (int j = 0; j < n; ++j) {
    const int Bf = df0 + j;
    const double q = qex[Bf];
    double R2 = 0.0;
    for (int x = 0; x < 3; ++x) {
      const double d = P[x] - qcen[3 * Bf + x];
      R2 += d * d;
    }
    const double pq = p * q, s = p + q, al = pq / s;
    double v;
    if (kind == 0) {
      v = 2.0 * kPi * kPi * sqrt(kPi) / (pq * sqrt(s)) * boys0(al * R2);
    } else {
      v = 0.0;
      for (int i = 0; i < gem.n; ++i) {
        const double gi = gem.g[i], ag = al + gi;
        const double ex = exp(-al * gi * R2 / ag);
        if (kind == 1) {
          const double t = kPi * kPi / (s * ag);
          v -= gem.c[i] * t * sqrt(t) * ex;
        } else {
          const double t = kPi / s;
          v -= gem.c[i] * t * sqrt(t) * (2.0 * kPi / ag) * ex * boys0(al * al * R2 / ag);
        }
      }
    }
    o[stride * j] = K * qnrm[Bf] * v;
  }
}

double* upload(Gen* g, const std::vector<double>& h) {
  void* p = nullptr;
  if (hipMalloc(&p, sizeof(double) * std::max<size_t>(h.size(), 1)) != hipSuccess) return nullptr;
  g->allocs.push_back(p);
  if (!h.empty() && hipMemcpy(p, h.data(), sizeof(double) * h.size(), hipMemcpyHostToDevice) != hipSuccess)
    return nullptr;
  return static_cast<double*>(p);
}

bool make_basis(Gen* g, const double* ex, const double* cen, int n, Basis* b) {
  std::vector<double> e(ex, ex + n), nr(n), c(cen, cen + 3 * size_t(n));
  for (int i = 0; i < n; ++i) nr[i] = std::pow(2.0 * ex[i] / kPi, 0.75);
  b->n = n;
  b->ex = upload(g, e);
  b->nrm = upload(g, nr);
  b->cen = upload(g, c);
  return b->ex && b->nrm && b->cen;
}

}  // namespace

extern "C" {

// A generator for one operator on device `device`: AO, DF and RI bases as
// exponent arrays and 3 n centre arrays (x, y, z per function); n_ri may be 0.
// kind 0: 1/r12, 1: -sum_i c_i exp(-g_i r12^2), 2: the geminal over r12.
int uw12s_create(int device, int kind, const double* ao_ex, const double* ao_cen, int n_ao, const double* df_ex,
                 const double* df_cen, int n_df, const double* ri_ex, const double* ri_cen, int n_ri, int ngem,
                 const double* gem_c, const double* gem_g, void** out) {
  if (!out || n_ao <= 0 || n_df <= 0 || n_ri < 0 || kind < 0 || kind > 2 || ngem < 0 || ngem > 8) return -1;
  if (hipSetDevice(device) != hipSuccess) return -3;
  Gen* g = new Gen();
  g->device = device;
  g->kind = kind;
  g->ngem = ngem;
  for (int i = 0; i < ngem; ++i) {
    g->c[i] = gem_c[i];
    g->g[i] = gem_g[i];
  }
  bool ok = make_basis(g, ao_ex, ao_cen, n_ao, &g->ao) && make_basis(g, df_ex, df_cen, n_df, &g->df);
  if (ok && n_ri > 0) ok = make_basis(g, ri_ex, ri_cen, n_ri, &g->ri);
  if (!ok) {
    for (void* p : g->allocs) (void)hipFree(p);
    delete g;
    return -5;
  }
  *out = g;
  return 0;
}

void uw12s_destroy(void* h) {
  Gen* g = static_cast<Gen*>(h);
  if (!g) return;
  (void)hipSetDevice(g->device);
  (void)hipDeviceSynchronize();
  for (void* p : g->allocs) (void)hipFree(p);
  delete g;
}

// uw12g_three_index_gen_fn: DF functions [df0, df0 + n), AO rows (with_ri 0)
// or AO then RI rows (with_ri 1), enqueued on `stream`.
int uw12s_generate(void* user, size_t df0, size_t n, int with_ri, double* out, void* stream) {
  const Gen* g = static_cast<const Gen*>(user);
  if (!g || !out || df0 + n > size_t(g->df.n)) return -1;
  if (with_ri && g->ri.n == 0) return -1;
  if (n == 0) return 0;
  const int m = g->ao.n + (with_ri ? g->ri.n : 0);
  const int64_t pairs = int64_t(m) * g->ao.n;
  GemPar gp;
  gp.n = g->ngem;
  for (int i = 0; i < 8; ++i) {
    gp.c[i] = g->c[i];
    gp.g[i] = g->g[i];
  }
  const int64_t blocks = (pairs + 255) / 256;
  if (blocks > 0x7fffffff) return -2;
  hipLaunchKernelGGL(synth_block_kernel, dim3(unsigned(blocks)), dim3(256), 0, static_cast<hipStream_t>(stream),
                     g->kind, gp, g->ao.ex, g->ao.nrm, g->ao.cen, g->ao.n, g->ri.ex, g->ri.nrm, g->ri.cen, m, g->df.ex,
                     g->df.nrm, g->df.cen, int(df0), int(n), out);
  return hipGetLastError() == hipSuccess ? 0 : -3;
}

// Test hook: the same block generated into a device buffer of this library's
// own and copied to host memory `host_out` (rows x n_ao x n doubles).
int uw12s_generate_host(void* user, size_t df0, size_t n, int with_ri, double* host_out) {
  const Gen* g = static_cast<const Gen*>(user);
  if (!g || !host_out) return -1;
  if (hipSetDevice(g->device) != hipSuccess) return -3;
  const size_t rows = size_t(g->ao.n) + (with_ri ? size_t(g->ri.n) : 0);
  const size_t count = rows * size_t(g->ao.n) * n;
  double* d = nullptr;
  if (hipMalloc(&d, sizeof(double) * std::max<size_t>(count, 1)) != hipSuccess) return -5;
  int rc = uw12s_generate(user, df0, n, with_ri, d, nullptr);
  if (rc == 0 && hipDeviceSynchronize() != hipSuccess) rc = -3;
  if (rc == 0 && count && hipMemcpy(host_out, d, sizeof(double) * count, hipMemcpyDeviceToHost) != hipSuccess) rc = -3;
  (void)hipFree(d);
  return rc;
}

}  // extern "C"
