// CPU oracle for the UW12 energy + Fock evaluation.
//
// TEST INFRASTRUCTURE ONLY.  The shipped library never links, loads or calls
// this file; only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline
// leg load it, and only as the checker or as the timed CPU baseline.
//
// It restates the reference algorithm (zwill22/libuw12) in plain C++17 +
// OpenMP + FP64 BLAS, following the reference's own contraction order:
//   driver                 src/uw12.cpp:22-119
//   orbitals / core        src/utils/utils.hpp:235-310
//   DF metric (eig, pinv)  src/integrals/base_integrals.cpp:155-166,
//                          src/utils/linalg.hpp:469-481,1112-1145
//   MO transforms          src/integrals/transformations.hpp:33-192,
//                          src/integrals/integrals.cpp:213-378
//   two-electron           src/two_electron/two_electron.cpp:11-152
//   four-electron          src/four_electron/four_electron.cpp:11-312,
//                          src/four_electron/four_electron_utils.cpp:9-79
//   ABS+ projectors        src/three_electron/ri_utils.hpp:45-74,
//                          src/utils/linalg.hpp:1157-1197
//   three-electron direct  src/three_electron/direct_utils.cpp:31-347,
//                          src/three_electron/three_electron.cpp:32-115
//   three-electron indir.  src/three_electron/indirect_utils.cpp:19-345,
//                          src/three_electron/three_electron.cpp:125-189
// Deviations from the reference's loop structure (same arithmetic per output
// element, different batching):
//   * J3 = X3_0 P2 and J3_ri = X3_ri_0 P2 are formed once with a GEMM instead
//     of X3_0 * col(P2, A) per A (the reference's own `has_J3()` path,
//     direct_utils.cpp:206-240, indirect_utils.cpp:261-295).
//   * In fock_i the per-pair GEMVs J3 * x_jk are batched over j for fixed k
//     into one GEMM (indirect_utils.cpp:159-191).
// What pins each function (tests/test_oracle.py): every energy term against
// brute-force four-index sums and a long-double evaluation; the all-electron
// Fock against finite differences of the energy; the frozen-core Fock
// (n_active < n_occ; four_electron.cpp:63-76,227-248, direct_utils.cpp:185-203,
// indirect_utils.cpp:311-342) against finite differences along orbital paths
// that move only the active orbitals, of which the reference's F_all + F_act
// sums are the exact derivative (the same Fock fails along paths that also
// move the core).
// BLAS/LAPACK (dgemm_, dsyevd_) are resolved at run time from MKL
// (libmkl_rt) when present; otherwise a built-in blocked GEMM and a cyclic
// Jacobi eigensolver are used.
#include <dlfcn.h>
#include <omp.h>

#include <algorithm>
#include <atomic>
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

namespace {

// --------------------------------------------------------------- BLAS shim
using dgemm_t = void (*)(const char*, const char*, const int*, const int*, const int*,
                         const double*, const double*, const int*, const double*,
                         const int*, const double*, double*, const int*);
using dsyevd_t = void (*)(const char*, const char*, const int*, double*, const int*,
                          double*, double*, const int*, int*, const int*, int*);
dgemm_t g_dgemm = nullptr;
dsyevd_t g_dsyevd = nullptr;
bool g_blas_init = false;
int g_force_builtin = 0;
// Stride over the outermost same-spin pair index (1 = every pair).  Values > 1
// evaluate a uniform sample of the pair loops; the CPU-baseline timing uses it
// to bound its run time and extrapolates.
int g_stride = 1;
int g_progress = 0;  // print completed pair-loop iterations to stderr (long runs)

// Completion count of one parallel pair loop, printed every ~5% when enabled.
struct Progress {
  const char* what;
  long total, step;
  std::atomic<long> done{0};
  Progress(const char* w, long t) : what(w), total(t), step(std::max(1L, t / 20)) {}
  void tick() {
    if (!g_progress) return;
    const long d = ++done;
    if (d % step == 0 || d == total) {
      std::fprintf(stderr, "[oracle] %s %ld/%ld\n", what, d, total);
      std::fflush(stderr);
    }
  }
};

void blas_init() {
  if (g_blas_init) return;
  g_blas_init = true;
  if (g_force_builtin) return;
  const char* names[] = {"libmkl_rt.so", "/opt/conda/lib/libmkl_rt.so", "libmkl_rt.so.1",
                         "/opt/conda/lib/libmkl_rt.so.1"};
  for (const char* n : names) {
    void* h = dlopen(n, RTLD_NOW | RTLD_LOCAL);
    if (!h) continue;
    auto set_layer = reinterpret_cast<int (*)(int)>(dlsym(h, "MKL_Set_Threading_Layer"));
    if (set_layer) set_layer(3);  // MKL_THREADING_GNU (3): share libgomp with this file
    g_dgemm = reinterpret_cast<dgemm_t>(dlsym(h, "dgemm_"));
    g_dsyevd = reinterpret_cast<dsyevd_t>(dlsym(h, "dsyevd_"));
    if (g_dgemm && g_dsyevd) return;
    g_dgemm = nullptr;
    g_dsyevd = nullptr;
  }
}

// C = alpha op(A) op(B) + beta C, column-major.
void gemm(char ta, char tb, int m, int n, int k, double alpha, const double* A, int lda,
          const double* B, int ldb, double beta, double* C, int ldc) {
  if (m <= 0 || n <= 0) return;
  if (k <= 0) {
    for (int j = 0; j < n; ++j)
      for (int i = 0; i < m; ++i) C[i + (size_t)ldc * j] = beta == 0 ? 0 : beta * C[i + (size_t)ldc * j];
    return;
  }
  blas_init();
  if (g_dgemm) {
    g_dgemm(&ta, &tb, &m, &n, &k, &alpha, A, &lda, B, &ldb, &beta, C, &ldc);
    return;
  }
  const bool tA = ta == 'T', tB = tb == 'T';
  auto a = [&](int i, int p) { return tA ? A[p + (size_t)lda * i] : A[i + (size_t)lda * p]; };
  auto b = [&](int p, int j) { return tB ? B[j + (size_t)ldb * p] : B[p + (size_t)ldb * j]; };
  const bool par = !omp_in_parallel() && (double)m * n * k > 1e6;
#pragma omp parallel for schedule(static) if (par)
  for (int j = 0; j < n; ++j) {
    std::vector<double> acc(m, 0.0);
    for (int p = 0; p < k; ++p) {
      const double bp = b(p, j);
      if (bp == 0) continue;
      for (int i = 0; i < m; ++i) acc[i] += a(i, p) * bp;
    }
    for (int i = 0; i < m; ++i) {
      double& c = C[i + (size_t)ldc * j];
      c = alpha * acc[i] + (beta == 0 ? 0 : beta * c);
    }
  }
}

// Symmetric eigendecomposition, ascending eigenvalues (LAPACK dsyevd, which
// Armadillo's eig_sym uses by default; linalg.hpp:1112-1145).
void eig_sym(int n, const double* A, double* w, double* V) {
  std::memcpy(V, A, sizeof(double) * (size_t)n * n);
  blas_init();
  if (g_dsyevd && n > 0) {
    int lwork = -1, liwork = -1, info = 0;
    double wq;
    int iwq;
    g_dsyevd("V", "L", &n, V, &n, w, &wq, &lwork, &iwq, &liwork, &info);
    lwork = (int)wq;
    liwork = iwq;
    std::vector<double> work(lwork);
    std::vector<int> iwork(liwork);
    g_dsyevd("V", "L", &n, V, &n, w, work.data(), &lwork, iwork.data(), &liwork, &info);
    if (info != 0) throw std::runtime_error("eigen-decomposition failed");
    return;
  }
  // cyclic Jacobi (Golub & Van Loan, Alg. 8.5.3) on a copy
  std::vector<double> a(A, A + (size_t)n * n);
  std::vector<double> v((size_t)n * n, 0.0);
  for (int i = 0; i < n; ++i) v[i + (size_t)n * i] = 1.0;
  for (int sweep = 0; sweep < 100; ++sweep) {
    double off = 0, tot = 0;
    for (int j = 0; j < n; ++j)
      for (int i = 0; i < n; ++i) {
        const double x = a[i + (size_t)n * j] * a[i + (size_t)n * j];
        tot += x;
        if (i != j) off += x;
      }
    if (off <= 1e-30 * tot || off == 0) break;
    for (int p = 0; p < n - 1; ++p)
      for (int q = p + 1; q < n; ++q) {
        const double apq = a[p + (size_t)n * q];
        if (apq == 0) continue;
        const double app = a[p + (size_t)n * p], aqq = a[q + (size_t)n * q];
        const double tau = (aqq - app) / (2 * apq);
        const double t = (tau >= 0 ? 1.0 : -1.0) / (std::fabs(tau) + std::sqrt(1 + tau * tau));
        const double c = 1 / std::sqrt(1 + t * t), s = t * c;
        for (int k = 0; k < n; ++k) {  // A <- A J
          const double akp = a[k + (size_t)n * p], akq = a[k + (size_t)n * q];
          a[k + (size_t)n * p] = c * akp - s * akq;
          a[k + (size_t)n * q] = s * akp + c * akq;
        }
        for (int k = 0; k < n; ++k) {  // A <- J^T A
          const double apk = a[p + (size_t)n * k], aqk = a[q + (size_t)n * k];
          a[p + (size_t)n * k] = c * apk - s * aqk;
          a[q + (size_t)n * k] = s * apk + c * aqk;
        }
        for (int k = 0; k < n; ++k) {
          const double vkp = v[k + (size_t)n * p], vkq = v[k + (size_t)n * q];
          v[k + (size_t)n * p] = c * vkp - s * vkq;
          v[k + (size_t)n * q] = s * vkp + c * vkq;
        }
      }
  }
  std::vector<int> idx(n);
  for (int i = 0; i < n; ++i) idx[i] = i;
  std::sort(idx.begin(), idx.end(), [&](int x, int y) { return a[x + (size_t)n * x] < a[y + (size_t)n * y]; });
  for (int j = 0; j < n; ++j) {
    w[j] = a[idx[j] + (size_t)n * idx[j]];
    std::memcpy(V + (size_t)n * j, v.data() + (size_t)n * idx[j], sizeof(double) * n);
  }
}

// ------------------------------------------------------------ small matrix
struct Mat {
  int r = 0, c = 0;
  std::vector<double> d;
  Mat() = default;
  Mat(int r_, int c_) : r(r_), c(c_), d((size_t)r_ * c_, 0.0) {}
  double& operator()(int i, int j) { return d[i + (size_t)r * j]; }
  double operator()(int i, int j) const { return d[i + (size_t)r * j]; }
  double* col(int j) { return d.data() + (size_t)r * j; }
  const double* col(int j) const { return d.data() + (size_t)r * j; }
};

Mat mul(const Mat& A, char ta, const Mat& B, char tb, double alpha = 1.0) {
  const int m = ta == 'N' ? A.r : A.c, k = ta == 'N' ? A.c : A.r;
  const int n = tb == 'N' ? B.c : B.r;
  Mat C(m, n);
  gemm(ta, tb, m, n, k, alpha, A.d.data(), std::max(1, A.r), B.d.data(), std::max(1, B.r), 0.0,
       C.d.data(), std::max(1, m));
  return C;
}

double dot(const double* a, const double* b, size_t n) {
  double s = 0;
  for (size_t i = 0; i < n; ++i) s += a[i] * b[i];
  return s;
}

void axpy(Mat& y, double a, const Mat& x) {
  for (size_t i = 0; i < y.d.size(); ++i) y.d[i] += a * x.d[i];
}

// square() of a packed lower column, utils.hpp:23-48
void square_into(const double* v, int n, double* out, double factor = 1.0) {
  size_t idx = 0;
  for (int i = 0; i < n; ++i)
    for (int j = 0; j <= i; ++j) {
      const double x = v[idx++];
      if (i != j) {
        out[i + (size_t)n * j] = factor * x;
        out[j + (size_t)n * i] = factor * x;
      } else {
        out[i + (size_t)n * i] = x;
      }
    }
}

Mat symmetrise(const Mat& F) {
  Mat out(F.r, F.c);
  for (int j = 0; j < F.c; ++j)
    for (int i = 0; i < F.r; ++i) out(i, j) = 0.5 * (F(i, j) + F(j, i));
  return out;
}

double spin_factor(int n_spin, int s, int sp, double k0, double k1) {  // four_electron_utils.cpp:9-30
  if (n_spin == 1) return 2 * (k0 + k1);
  return s == sp ? k1 : k0;
}

// ------------------------------------------------------------ integrals
struct Base {  // base_integrals.cpp
  int N = 0, Nd = 0, Nr = 0;
  Mat X3, X3ri;  // packed (P x Nd), (N*Nr x Nd)
  Mat P2;        // Q pinv(diag vals)
  std::vector<double> vals;
  Mat J3, J3ri;  // X3 P2, X3ri P2
};

void base_setup(Base& b, const double* X3, const double* X2, const double* X3ri, int N, int Nd,
                int Nr) {
  b.N = N;
  b.Nd = Nd;
  b.Nr = X3ri ? Nr : 0;
  const int P = N * (N + 1) / 2;
  b.X3 = Mat(P, Nd);
  std::memcpy(b.X3.d.data(), X3, sizeof(double) * (size_t)P * Nd);
  if (X3ri) {
    b.X3ri = Mat(N * Nr, Nd);
    std::memcpy(b.X3ri.d.data(), X3ri, sizeof(double) * (size_t)N * Nr * Nd);
  }
  Mat Q(Nd, Nd);
  b.vals.assign(Nd, 0.0);
  eig_sym(Nd, X2, b.vals.data(), Q.d.data());
  b.P2 = Mat(Nd, Nd);
  for (int A = 0; A < Nd; ++A) {
    const double l = b.vals[A];
    const double inv = std::fabs(l) >= 1e-10 ? 1.0 / l : 0.0;  // pinv tol, linalg.hpp:469-481
    for (int i = 0; i < Nd; ++i) b.P2(i, A) = Q(i, A) * inv;
  }
  b.J3 = mul(b.X3, 'N', b.P2, 'N');
  if (X3ri) b.J3ri = mul(b.X3ri, 'N', b.P2, 'N');
}

struct Ints {  // integrals.cpp: Integrals(base, Co, Ca)
  const Base* b = nullptr;
  std::vector<Mat> Co, Ca;
  std::vector<Mat> one, one_ri, two;
};

// transformations.hpp:33-54, then * P2 (integrals.cpp:226-232)
Mat one_trans(const Base& b, const Mat& C) {
  const int N = b.N, o = C.c, Nd = b.Nd;
  Mat out(N * o, Nd);
#pragma omp parallel
  {
    std::vector<double> sq((size_t)N * N);
#pragma omp for schedule(dynamic)
    for (int A = 0; A < Nd; ++A) {
      square_into(b.X3.col(A), N, sq.data());
      gemm('N', 'N', N, o, N, 1.0, sq.data(), N, C.d.data(), N, 0.0, out.col(A), N);
    }
  }
  return mul(out, 'N', b.P2, 'N');
}

// transformations.hpp:130-158 then * P2 (integrals.cpp:263-269)
Mat one_trans_ri(const Base& b, const Mat& C) {
  const int N = b.N, Nr = b.Nr, a = C.c, Nd = b.Nd;
  Mat out(Nr * a, Nd);
#pragma omp parallel for schedule(dynamic)
  for (int A = 0; A < Nd; ++A)
    gemm('N', 'N', Nr, a, N, 1.0, b.X3ri.col(A), Nr, C.d.data(), N, 0.0, out.col(A), Nr);
  return mul(out, 'N', b.P2, 'N');
}

// transformations.hpp:79-106
Mat two_trans(const Mat& one, const Mat& C, int N) {
  const int o = one.r / N, a = C.c, Nd = one.c;
  Mat out(a * o, Nd);
#pragma omp parallel for schedule(dynamic)
  for (int A = 0; A < Nd; ++A)
    gemm('T', 'N', a, o, N, 1.0, C.d.data(), N, one.col(A), N, 0.0, out.col(A), a);
  return out;
}

void ints_build(Ints& I, const Base& b, const std::vector<Mat>& Co, const std::vector<Mat>& Ca,
                bool need_ri) {
  I.b = &b;
  I.Co = Co;
  I.Ca = Ca;
  const int ns = (int)Co.size();
  I.one.resize(ns);
  I.one_ri.resize(ns);
  I.two.resize(ns);
  for (int s = 0; s < ns; ++s) {
    I.one[s] = one_trans(b, Co[s]);
    if (need_ri) I.one_ri[s] = one_trans_ri(b, Ca[s]);
    I.two[s] = two_trans(I.one[s], Ca[s], b.N);
  }
}

struct ABS {
  Mat rr, ra, ar, aa;  // s_inv_ri_ri, s_inv_ri_ao, s_inv_ao_ri, s_inv_ao_ao
};

ABS abs_setup(const double* S, int N, int Nr, double rel = 1e-8, double absthr = 1e-6) {
  const int M = N + Nr;
  Mat V(M, M);
  std::vector<double> w(M);
  eig_sym(M, S, w.data(), V.d.data());
  double mx = 0;
  for (double x : w) mx = std::max(mx, std::fabs(x));
  const double thr = std::max(mx * rel, absthr);
  std::vector<int> keep;
  for (int i = 0; i < M; ++i)
    if (std::fabs(w[i]) > thr) keep.push_back(i);
  const int K = (int)keep.size();
  Mat Vr(Nr, K), Va(N, K), Vri(Nr, K), Vai(N, K);
  for (int k = 0; k < K; ++k) {
    const double inv = 1.0 / w[keep[k]];
    for (int i = 0; i < N; ++i) {
      Va(i, k) = V(i, keep[k]);
      Vai(i, k) = V(i, keep[k]) * inv;
    }
    for (int i = 0; i < Nr; ++i) {
      Vr(i, k) = V(N + i, keep[k]);
      Vri(i, k) = V(N + i, keep[k]) * inv;
    }
  }
  ABS p;
  p.rr = mul(Vri, 'N', Vr, 'T');
  p.ra = mul(Vri, 'N', Va, 'T');
  p.ar = mul(Vai, 'N', Vr, 'T');
  p.aa = mul(Vai, 'N', Va, 'T');
  return p;
}

struct Result {
  std::vector<Mat> F;
  double e2 = 0, e3 = 0, e4 = 0;
  double e2d = 0, e2i = 0, e3d = 0, e3i = 0, e4d = 0, e4i = 0;
};

std::vector<Mat> zeros(int ns, int N) { return std::vector<Mat>(ns, Mat(N, N)); }

// ------------------------------------------------------------ two-electron
void two_electron(const Base& WVb, const std::vector<Mat>& Ca, bool indirect, bool calc_fock,
                  double k0, double k1, Result& R) {
  const int ns = (int)Ca.size(), N = WVb.N, Nd = WVb.Nd;
  Ints I;
  ints_build(I, WVb, Ca, Ca, false);
  // get_X_D, integrals.cpp:347-378
  std::vector<std::vector<double>> XD(ns, std::vector<double>(Nd)), XDt(ns, std::vector<double>(Nd));
  const int P = N * (N + 1) / 2;
  for (int s = 0; s < ns; ++s) {
    const double g = ns == 1 ? 2.0 : 1.0;
    Mat D = mul(Ca[s], 'N', Ca[s], 'T', g);
    std::vector<double> low(P);
    size_t idx = 0;
    for (int i = 0; i < N; ++i) {
      for (int j = 0; j < i; ++j) low[idx++] = 2.0 * D(i, j);
      low[idx++] = D(i, i);
    }
    std::vector<double> tmp(Nd);
    gemm('T', 'N', Nd, 1, P, 1.0, WVb.X3.d.data(), P, low.data(), P, 0.0, tmp.data(), Nd);
    gemm('T', 'N', Nd, 1, Nd, 0.5 * ns, WVb.P2.d.data(), Nd, tmp.data(), Nd, 0.0, XD[s].data(), Nd);
    for (int A = 0; A < Nd; ++A) XDt[s][A] = WVb.vals[A] * XD[s][A];
  }
  std::vector<Mat> F = zeros(ns, N);
  double E = 0;
  for (int s = 0; s < ns; ++s)
    for (int t = 0; t < ns; ++t) {
      const double sf = s == t ? k1 : k0;
      const double f = ns == 1 ? 2 * (k0 + k1) : sf;
      if (f == 0) continue;
      E += 0.5 * f * dot(XD[s].data(), XDt[t].data(), Nd);
      if (calc_fock) {  // calculate_direct_fock, two_electron.cpp:26-35
        std::vector<double> v(Nd), col(P);
        gemm('N', 'N', Nd, 1, Nd, 1.0, WVb.P2.d.data(), Nd, XDt[t].data(), Nd, 0.0, v.data(), Nd);
        gemm('N', 'N', P, 1, Nd, 1.0, WVb.X3.d.data(), P, v.data(), Nd, 0.0, col.data(), P);
        std::vector<double> sq((size_t)N * N);
        square_into(col.data(), N, sq.data());
        for (size_t i = 0; i < sq.size(); ++i) F[s].d[i] += 0.5 * ns * f * sq[i];
      }
    }
  R.e2d += E;
  double Ei = 0;
  if (indirect && k1 != 0) {  // indirect_fock, two_electron.cpp:87-132
    const double esf = ns == 1 ? 2 : 1;
    for (int s = 0; s < ns; ++s) {
      const int o = Ca[s].c;
      if (o == 0) continue;
      const Mat& T2 = I.two[s];
      double acc = 0;
      for (int A = 0; A < Nd; ++A) acc += WVb.vals[A] * dot(T2.col(A), T2.col(A), T2.r);
      Ei -= k1 * 0.5 * esf * acc;
      if (calc_fock) {
        Mat t2(I.one[s].r, Nd);
        for (int A = 0; A < Nd; ++A)
          for (int i = 0; i < t2.r; ++i) t2(i, A) = I.one[s](i, A) * WVb.vals[A];
        Mat f(N, N);
        gemm('N', 'T', N, N, o * Nd, -1.0, I.one[s].d.data(), N, t2.d.data(), N, 0.0, f.d.data(), N);
        axpy(F[s], k1, f);
      }
    }
  }
  R.e2i += Ei;
  for (int s = 0; s < ns; ++s) axpy(R.F[s], 1.0, F[s]);
}

// ------------------------------------------------------------ four-electron
std::vector<Mat> calc_tab(const Ints& W, const Ints& V) {  // four_electron_utils.cpp:33-49
  std::vector<Mat> t;
  for (size_t s = 0; s < W.two.size(); ++s) t.push_back(mul(W.two[s], 'T', V.two[s], 'N'));
  return t;
}

std::vector<Mat> calc_ttilde(const Ints& W, const Ints& V, const std::vector<Mat>& t) {  // :51-79
  std::vector<Mat> out;
  for (const Mat& x : t) {
    Mat y(x.r, x.c);
    for (int B = 0; B < x.c; ++B)
      for (int A = 0; A < x.r; ++A) y(A, B) = W.b->vals[A] * x(A, B) * V.b->vals[B];
    out.push_back(y);
  }
  return out;
}

Mat direct_fock_contribution(const double* W1, const double* V1, int rows, const Mat& tt,
                             double esf, int N, int ns) {  // four_electron.cpp:11-41
  const int o = rows / N, Nd = tt.r;
  // W1 and V1 are column blocks with leading dimension ld passed via rows? use packed copies
  Mat Wt(rows, Nd);
  gemm('N', 'N', rows, Nd, Nd, 1.0, W1, rows, tt.d.data(), Nd, 0.0, Wt.d.data(), rows);
  const double fsf = (ns == 1 ? 0.5 : 1.0) * esf;
  Mat F(N, N);
  gemm('N', 'T', N, N, o * Nd, 2 * fsf, V1, N, Wt.d.data(), N, 0.0, F.d.data(), N);
  return F;
}

Mat tail_rows(const Mat& X, int n) {
  Mat out(n, X.c);
  for (int j = 0; j < X.c; ++j) std::memcpy(out.col(j), X.col(j) + (X.r - n), sizeof(double) * n);
  return out;
}

void four_electron(const Ints& W, const Ints& V, bool indirect, bool calc_fock, double k0, double k1,
                   Result& R, const std::vector<Mat>& tab, const std::vector<Mat>& tt) {
  const int ns = (int)W.Co.size(), N = W.b->N;
  std::vector<Mat> F = zeros(ns, N);
  double E = 0;
  for (int s = 0; s < ns; ++s) {
    const int a = W.Ca[s].c, o = W.Co[s].c;
    for (int t = 0; t < ns; ++t) {
      const double f = spin_factor(ns, s, t, k0, k1);
      if (f == 0) continue;
      E += 0.5 * f * dot(tt[t].d.data(), tab[s].d.data(), tt[t].d.size());
      if (!calc_fock) continue;
      // direct_fock_matrix, four_electron.cpp:43-77
      Mat fk = direct_fock_contribution(W.one[s].d.data(), V.one[s].d.data(), W.one[s].r, tt[t], f, N, ns);
      if (!(o == a || a == 0)) {
        Mat Wa = tail_rows(W.one[s], N * a), Va = tail_rows(V.one[s], N * a);
        Mat ff = direct_fock_contribution(Wa.d.data(), Va.d.data(), Wa.r, tt[t], f, N, ns);
        for (size_t i = 0; i < fk.d.size(); ++i) fk.d[i] = 0.5 * (fk.d[i] + ff.d[i]);
      }
      axpy(F[s], 1.0, fk);
    }
  }
  for (int s = 0; s < ns; ++s) axpy(R.F[s], 1.0, symmetrise(F[s]));
  R.e4d += E;
  if (!(indirect && k1 != 0)) return;
  // indirect_four_el_fock, four_electron.cpp:253-290
  double Ei = 0;
  std::vector<Mat> Fi = zeros(ns, N);
  for (int s = 0; s < ns; ++s) {
    const int o = W.Co[s].c, a = W.Ca[s].c, Nd = W.b->Nd;
    if (a == 0 || o == 0 || N == 0) continue;
    auto four = [&](const Ints& X) {  // integrals.cpp:339-345
      Mat t2(X.two[s].r, Nd);
      for (int A = 0; A < Nd; ++A)
        for (int i = 0; i < t2.r; ++i) t2(i, A) = X.two[s](i, A) * X.b->vals[A];
      return mul(t2, 'N', X.two[s], 'T');
    };
    Mat W4 = four(W), V4 = four(V);
    // indirect_energy, four_electron.cpp:143-179
    double acc = 0;
#pragma omp parallel for reduction(+ : acc) collapse(2)
    for (int k = 0; k < o; k += g_stride)
      for (int l = 0; l < o; ++l) {
        double v = 0;
        for (int jj = 0; jj < a; ++jj)
          for (int ii = 0; ii < a; ++ii) v += W4(l * a + ii, k * a + jj) * V4(k * a + ii, l * a + jj);
        acc += -0.5 * (ns == 1 ? 2 : 1) * v;
      }
    Ei += k1 * acc;
    if (!calc_fock) continue;
    auto three = [&](const Ints& X) {  // integrals.cpp:333-337
      Mat t2(X.one[s].r, Nd);
      for (int A = 0; A < Nd; ++A)
        for (int i = 0; i < t2.r; ++i) t2(i, A) = X.one[s](i, A) * X.b->vals[A];
      return mul(t2, 'N', X.two[s], 'T');
    };
    Mat W3 = three(W), V3 = three(V);  // (N o) x (a o)
    const int core = o - a;
    // indirect_fock_matrix, four_electron.cpp:181-249
    Mat fc(N, N);
    {
      const int ld3 = W3.r;
      for (int i = 0; i < a; i += g_stride)
        for (int k = 0; k < o; ++k) {
          // W4_ik = sub_mat(W3, (core+i)*N, a*k, N, a); V4_ik = reshape(tail(col(V3, i + a k), N a), N, a)
          const double* Wp = W3.d.data() + (size_t)(core + i) * N + (size_t)ld3 * (a * k);
          const double* Vp = V3.col(i + a * k) + (size_t)(ld3 - N * a);
          gemm('N', 'T', N, N, a, -2.0, Wp, ld3, Vp, N, 1.0, fc.d.data(), N);
        }
    }
    if (core != 0) {
      Mat fr(N, N);
      const int ld3 = W3.r;
      for (int k = 0; k < o; k += g_stride)
        for (int l = 0; l < o; ++l) {
          const double* Wp = W3.d.data() + (size_t)l * N + (size_t)ld3 * (k * a);
          const double* Vp = V3.d.data() + (size_t)k * N + (size_t)ld3 * (l * a);
          gemm('N', 'T', N, N, a, -2.0, Wp, ld3, Vp, ld3, 1.0, fr.d.data(), N);
        }
      for (size_t i = 0; i < fc.d.size(); ++i) fc.d[i] = 0.5 * (fc.d[i] + fr.d[i]);
    }
    axpy(Fi[s], k1, fc);
  }
  for (int s = 0; s < ns; ++s) axpy(R.F[s], 1.0, symmetrise(Fi[s]));
  R.e4i += Ei;
}

// ------------------------------------------------------------ three-electron
// calculate_xab_component, direct_utils.cpp:31-52
void xab_component(const double* W, int ldw, const double* V, int ldv, int n_orb, int a, const Mat& proj,
                   int Nd, Mat& out) {
#pragma omp parallel
  {
    std::vector<double> vp((size_t)proj.r * a);
#pragma omp for schedule(dynamic)
    for (int A = 0; A < Nd; ++A) {
      gemm('N', 'N', proj.r, a, n_orb, 1.0, proj.d.data(), proj.r, V + (size_t)ldv * A, n_orb, 0.0,
           vp.data(), proj.r);
      gemm('T', 'N', Nd, 1, proj.r * a, 1.0, W, ldw, vp.data(), proj.r * a, 1.0, out.col(A), Nd);
    }
  }
}

std::vector<Mat> calc_xab(const Ints& W, const Ints& V, const ABS& P) {  // direct_utils.cpp:62-133
  const int ns = (int)W.Co.size(), N = W.b->N, Nr = W.b->Nr, Nd = W.b->Nd;
  std::vector<Mat> x(ns, Mat(Nd, V.b->Nd));
  for (int s = 0; s < ns; ++s) {
    const int a = W.Ca[s].c;
    const Mat &W1 = W.one[s], &V1 = V.one[s], &Wr = W.one_ri[s], &Vr = V.one_ri[s];
    xab_component(Wr.d.data(), Wr.r, Vr.d.data(), Vr.r, Nr, a, P.rr, Nd, x[s]);
    if (a == 0) continue;
    const size_t off = (size_t)W1.r - (size_t)N * a;  // tail rows
    xab_component(W1.d.data() + off, W1.r, Vr.d.data(), Vr.r, Nr, a, P.ar, Nd, x[s]);
    xab_component(W1.d.data() + off, W1.r, V1.d.data() + off, V1.r, N, a, P.aa, Nd, x[s]);
    xab_component(Wr.d.data(), Wr.r, V1.d.data() + off, V1.r, N, a, P.ra, Nd, x[s]);
  }
  return x;
}

// calculate_xab_dttilde, direct_utils.cpp:164-204
Mat xab_dttilde(const Mat& W1, const Mat& V1, const Mat& xab, const std::vector<double>& wv,
                const std::vector<double>& vv, int a, int N) {
  const int Nd = (int)wv.size(), o = W1.r / N;
  Mat Wl = W1, Vl = V1;
  for (int A = 0; A < Nd; ++A)
    for (int i = 0; i < W1.r; ++i) {
      Wl(i, A) *= wv[A];
      Vl(i, A) *= vv[A];
    }
  Mat w = mul(Wl, 'N', xab, 'N');
  Mat F(N, N);
  const int nth = omp_get_max_threads();
  std::vector<Mat> part(nth, Mat(N, N));
#pragma omp parallel
  {
    Mat& f = part[omp_get_thread_num()];
#pragma omp for schedule(static)
    for (int A = 0; A < (int)Vl.c; ++A) {
      gemm('N', 'T', N, N, o, -1.0, w.col(A), N, Vl.col(A), N, 1.0, f.d.data(), N);
      if (a == 0) continue;
      gemm('N', 'T', N, N, a, -1.0, w.col(A) + (size_t)N * (o - a), N, Vl.col(A) + (size_t)N * (o - a), N,
           1.0, f.d.data(), N);
    }
  }
  for (auto& p : part) axpy(F, 1.0, p);
  return F;
}

// calculate_ttilde_dxab_incore: s term (direct_utils.cpp:256-286) + p term (:302-334)
Mat ttilde_dxab(const Base& Wb, const Base& Vb, const Mat& tt, const ABS& P) {
  const int N = Wb.N, Nr = Wb.Nr, NdW = Wb.Nd;
  Mat s = mul(Vb.J3, 'N', tt, 'T');    // (PP x NdW)
  Mat p = mul(Vb.J3ri, 'N', tt, 'T');  // (N Nr x NdW)
  const int nth = omp_get_max_threads();
  std::vector<Mat> part(nth, Mat(N, N));
#pragma omp parallel
  {
    Mat& f = part[omp_get_thread_num()];
    std::vector<double> sqW((size_t)N * N), sA((size_t)N * N), t1((size_t)N * N), t2((size_t)Nr * N);
#pragma omp for schedule(dynamic)
    for (int A = 0; A < NdW; ++A) {
      square_into(Wb.J3.col(A), N, sqW.data());
      const double* WriA = Wb.J3ri.col(A);  // reshape (Nr x N)
      square_into(s.col(A), N, sA.data());
      const double* pA = p.col(A);  // reshape (Nr x N)
      // s term: -sq(W3_A) S_aa sA - W3ri_A^T S_ra sA
      gemm('N', 'N', N, N, N, 1.0, P.aa.d.data(), N, sA.data(), N, 0.0, t1.data(), N);
      gemm('N', 'N', N, N, N, -1.0, sqW.data(), N, t1.data(), N, 1.0, f.d.data(), N);
      gemm('N', 'N', Nr, N, N, 1.0, P.ra.d.data(), Nr, sA.data(), N, 0.0, t2.data(), Nr);
      gemm('T', 'N', N, N, Nr, -1.0, WriA, Nr, t2.data(), Nr, 1.0, f.d.data(), N);
      // p term: -sq(W3_A) S_ar pA - W3ri_A^T S_rr pA
      gemm('N', 'N', N, N, Nr, 1.0, P.ar.d.data(), N, pA, Nr, 0.0, t1.data(), N);
      gemm('N', 'N', N, N, N, -1.0, sqW.data(), N, t1.data(), N, 1.0, f.d.data(), N);
      gemm('N', 'N', Nr, N, Nr, 1.0, P.rr.d.data(), Nr, pA, Nr, 0.0, t2.data(), Nr);
      gemm('T', 'N', N, N, Nr, -1.0, WriA, Nr, t2.data(), Nr, 1.0, f.d.data(), N);
    }
  }
  Mat F(N, N);
  for (auto& x : part) axpy(F, 1.0, x);
  return F;
}

void three_direct(const Ints& W, const Ints& V, const ABS& P, bool calc_fock, double k0, double k1,
                  Result& R, const std::vector<Mat>& tt) {
  const int ns = (int)W.Co.size(), N = W.b->N;
  std::vector<Mat> x = calc_xab(W, V, P);
  double E = 0;  // calculate_direct_energy, direct_utils.cpp:136-157
  for (int s = 0; s < ns; ++s)
    for (int t = 0; t < ns; ++t)
      E -= spin_factor(ns, s, t, k0, k1) * dot(x[s].d.data(), tt[t].d.data(), x[s].d.size());
  R.e3d += E;
  if (!calc_fock) return;
  std::vector<Mat> F = zeros(ns, N);  // calculate_direct_fock_matrix, three_electron.cpp:32-83
  for (int s = 0; s < ns; ++s) {
    const int a = W.Ca[s].c;
    Mat fsp = ttilde_dxab(*W.b, *V.b, tt[s], P);
    for (int t = 0; t < ns; ++t) {
      const double dfsf = (ns == 1 ? 0.5 : 1.0) * spin_factor(ns, s, t, k0, k1);
      Mat fx = xab_dttilde(W.one[s], V.one[s], x[t], W.b->vals, V.b->vals, a, N);
      axpy(F[s], dfsf, fx);
      axpy(F[t], dfsf, fsp);
    }
  }
  for (int s = 0; s < ns; ++s) axpy(R.F[s], 1.0, symmetrise(F[s]));
}

// indirect_3el_energy, indirect_utils.cpp:19-133
double indirect_3el_energy(const Ints& W, const Ints& V, const ABS& P, int s) {
  const int N = W.b->N, Nr = W.b->Nr, NdW = W.b->Nd, NdV = V.b->Nd;
  const int o = W.Co[s].c, a = W.Ca[s].c, core = o - a, ns = (int)W.Co.size();
  const Mat &W1 = W.one[s], &V1 = V.one[s], &Wr = W.one_ri[s], &Vr = V.one_ri[s];
  const Mat &W2 = W.two[s], &V2 = V.two[s];
  double acc = 0;
  Progress prog("indirect 3e energy pairs", long((a + g_stride - 1) / g_stride) * a);
#pragma omp parallel reduction(+ : acc)
  {
    Mat x(NdW, NdV), t1(Nr, NdV), t2(N, NdV), Wt(o, NdW), Vt(o, NdV), ttij(NdW, NdV);
#pragma omp for collapse(2) schedule(dynamic)
    for (int i = 0; i < a; i += g_stride)
      for (int j = 0; j < a; ++j) {
        // calculate_xij
        gemm('N', 'N', Nr, NdV, Nr, 1.0, P.rr.d.data(), Nr, Vr.d.data() + (size_t)j * Nr, Vr.r, 0.0, t1.d.data(), Nr);
        gemm('T', 'N', NdW, NdV, Nr, 1.0, Wr.d.data() + (size_t)i * Nr, Wr.r, t1.d.data(), Nr, 0.0, x.d.data(), NdW);
        gemm('N', 'N', Nr, NdV, N, 1.0, P.ra.d.data(), Nr, V1.d.data() + (size_t)(j + core) * N, V1.r, 0.0, t1.d.data(), Nr);
        gemm('T', 'N', NdW, NdV, Nr, 1.0, Wr.d.data() + (size_t)i * Nr, Wr.r, t1.d.data(), Nr, 1.0, x.d.data(), NdW);
        gemm('N', 'N', N, NdV, Nr, 1.0, P.ar.d.data(), N, Vr.d.data() + (size_t)j * Nr, Vr.r, 0.0, t2.d.data(), N);
        gemm('T', 'N', NdW, NdV, N, 1.0, W1.d.data() + (size_t)(i + core) * N, W1.r, t2.d.data(), N, 1.0, x.d.data(), NdW);
        gemm('N', 'N', N, NdV, N, 1.0, P.aa.d.data(), N, V1.d.data() + (size_t)(j + core) * N, V1.r, 0.0, t2.d.data(), N);
        gemm('T', 'N', NdW, NdV, N, 1.0, W1.d.data() + (size_t)(i + core) * N, W1.r, t2.d.data(), N, 1.0, x.d.data(), NdW);
        // calculate_ttilde_ij: rows j + a k of W2, i + a k of V2, times vals
        for (int A = 0; A < NdW; ++A)
          for (int k = 0; k < o; ++k) Wt(k, A) = W2(j + a * k, A) * W.b->vals[A];
        for (int B = 0; B < NdV; ++B)
          for (int k = 0; k < o; ++k) Vt(k, B) = V2(i + a * k, B) * V.b->vals[B];
        gemm('T', 'N', NdW, NdV, o, 1.0, Wt.d.data(), o, Vt.d.data(), o, 0.0, ttij.d.data(), NdW);
        acc += (ns == 1 ? 2 : 1) * dot(x.d.data(), ttij.d.data(), x.d.size());
        prog.tick();
      }
  }
  return acc;
}

// fock_i summed over (j, k), indirect_utils.cpp:138-194 with X3 GEMVs batched over j.
// The k loop runs in order, its per-k products on all threads (threaded BLAS);
// the j loop inside runs in parallel.  So every k costs the same however many
// k are evaluated, which keeps a strided sample proportional to the full sum.
Mat fock_i_sum(const Ints& X, const Ints& Y, const ABS& P, int s) {
  const Base &Xb = *X.b, &Yb = *Y.b;
  const int N = Xb.N, Nr = Xb.Nr, NdX = Xb.Nd, NdY = Yb.Nd, PP = N * (N + 1) / 2;
  const int o = X.Co[s].c, a = X.Ca[s].c, core = o - a;
  const Mat &X2 = X.two[s], &Y1 = Y.one[s], &Yr = Y.one_ri[s];
  const int nth = omp_get_max_threads();
  std::vector<Mat> part(nth, Mat(N, N));
  Mat xt(NdX, a), xam(Nr * N, a), xap(PP, a), Yk(NdY, N);
  Progress prog("indirect 3e Fock (i, j terms) outer k", long((o + g_stride - 1) / g_stride));
  for (int k = 0; k < o; k += g_stride) {
    for (int j = 0; j < a; ++j)
      for (int A = 0; A < NdX; ++A) xt(A, j) = X2(k * a + j, A) * Xb.vals[A];
    gemm('N', 'N', Nr * N, a, NdX, 1.0, Xb.J3ri.d.data(), Nr * N, xt.d.data(), NdX, 0.0, xam.d.data(), Nr * N);
    gemm('N', 'N', PP, a, NdX, 1.0, Xb.J3.d.data(), PP, xt.d.data(), NdX, 0.0, xap.d.data(), PP);
    for (int b = 0; b < N; ++b)
      for (int B = 0; B < NdY; ++B) Yk(B, b) = Yb.vals[B] * Y1(k * N + b, B);
#pragma omp parallel
    {
      Mat& f = part[omp_get_thread_num()];
      Mat yvj(Nr, N), ypj(N, N), t1(Nr, N), t2(N, N), sq(N, N);
#pragma omp for schedule(dynamic)
      for (int j = 0; j < a; ++j) {
        const double* Xam = xam.col(j);  // (Nr x N)
        gemm('N', 'N', Nr, N, NdY, 1.0, Yr.d.data() + (size_t)Nr * j, Yr.r, Yk.d.data(), NdY, 0.0, yvj.d.data(), Nr);
        gemm('N', 'N', Nr, N, Nr, 1.0, P.rr.d.data(), Nr, yvj.d.data(), Nr, 0.0, t1.d.data(), Nr);
        gemm('T', 'N', N, N, Nr, 1.0, Xam, Nr, t1.d.data(), Nr, 1.0, f.d.data(), N);
        gemm('N', 'N', N, N, NdY, 1.0, Y1.d.data() + (size_t)(j + core) * N, Y1.r, Yk.d.data(), NdY, 0.0, ypj.d.data(), N);
        gemm('N', 'N', Nr, N, N, 1.0, P.ra.d.data(), Nr, ypj.d.data(), N, 0.0, t1.d.data(), Nr);
        gemm('T', 'N', N, N, Nr, 1.0, Xam, Nr, t1.d.data(), Nr, 1.0, f.d.data(), N);
        square_into(xap.col(j), N, sq.d.data());
        gemm('N', 'N', N, N, Nr, 1.0, P.ar.d.data(), N, yvj.d.data(), Nr, 0.0, t2.d.data(), N);
        gemm('N', 'N', N, N, N, 1.0, sq.d.data(), N, t2.d.data(), N, 1.0, f.d.data(), N);
        gemm('N', 'N', N, N, N, 1.0, P.aa.d.data(), N, ypj.d.data(), N, 0.0, t2.d.data(), N);
        gemm('N', 'N', N, N, N, 1.0, sq.d.data(), N, t2.d.data(), N, 1.0, f.d.data(), N);
      }
    }
    prog.tick();
  }
  Mat F(N, N);
  for (auto& x : part) axpy(F, 1.0, x);
  return F;
}

// fock_k summed over (i, j), indirect_utils.cpp:199-259
Mat fock_k_sum(const Ints& W, const Ints& V, const ABS& P, int s) {
  const int N = W.b->N, Nr = W.b->Nr, NdW = W.b->Nd, NdV = V.b->Nd;
  const int o = W.Co[s].c, a = W.Ca[s].c, core = o - a;
  const Mat &W1 = W.one[s], &V1 = V.one[s], &Wr = W.one_ri[s], &Vr = V.one_ri[s];
  const int nth = omp_get_max_threads();
  std::vector<Mat> part(nth, Mat(N, N));
  Progress prog("indirect 3e Fock (k term) pairs", long((a + g_stride - 1) / g_stride) * a);
#pragma omp parallel
  {
    Mat& f = part[omp_get_thread_num()];
    Mat Wj(NdW, N), Vi(NdV, N), Wjb_im(Nr, N), Vvj_ai(Nr, N), Vpj_ai(N, N), Wjb_ip(N, N), t1(Nr, N), t2(N, N);
#pragma omp for collapse(2) schedule(dynamic)
    for (int i = 0; i < a; i += g_stride)
      for (int j = 0; j < a; ++j) {
        for (int b = 0; b < N; ++b) {
          for (int A = 0; A < NdW; ++A) Wj(A, b) = W.b->vals[A] * W1((j + core) * N + b, A);
          for (int B = 0; B < NdV; ++B) Vi(B, b) = V.b->vals[B] * V1((i + core) * N + b, B);
        }
        gemm('N', 'N', Nr, N, NdW, 1.0, Wr.d.data() + (size_t)Nr * i, Wr.r, Wj.d.data(), NdW, 0.0, Wjb_im.d.data(), Nr);
        gemm('N', 'N', Nr, N, NdV, 1.0, Vr.d.data() + (size_t)Nr * j, Vr.r, Vi.d.data(), NdV, 0.0, Vvj_ai.d.data(), Nr);
        gemm('N', 'N', Nr, N, Nr, 1.0, P.rr.d.data(), Nr, Vvj_ai.d.data(), Nr, 0.0, t1.d.data(), Nr);
        gemm('T', 'N', N, N, Nr, 1.0, Wjb_im.d.data(), Nr, t1.d.data(), Nr, 1.0, f.d.data(), N);
        gemm('N', 'N', N, N, NdV, 1.0, V1.d.data() + (size_t)N * (j + core), V1.r, Vi.d.data(), NdV, 0.0, Vpj_ai.d.data(), N);
        gemm('N', 'N', Nr, N, N, 1.0, P.ra.d.data(), Nr, Vpj_ai.d.data(), N, 0.0, t1.d.data(), Nr);
        gemm('T', 'N', N, N, Nr, 1.0, Wjb_im.d.data(), Nr, t1.d.data(), Nr, 1.0, f.d.data(), N);
        gemm('N', 'N', N, N, NdW, 1.0, W1.d.data() + (size_t)N * (core + i), W1.r, Wj.d.data(), NdW, 0.0, Wjb_ip.d.data(), N);
        gemm('N', 'N', N, N, Nr, 1.0, P.ar.d.data(), N, Vvj_ai.d.data(), Nr, 0.0, t2.d.data(), N);
        gemm('T', 'N', N, N, N, 1.0, Wjb_ip.d.data(), N, t2.d.data(), N, 1.0, f.d.data(), N);
        gemm('N', 'N', N, N, N, 1.0, P.aa.d.data(), N, Vpj_ai.d.data(), N, 0.0, t2.d.data(), N);
        gemm('T', 'N', N, N, N, 1.0, Wjb_ip.d.data(), N, t2.d.data(), N, 1.0, f.d.data(), N);
        prog.tick();
      }
  }
  Mat F(N, N);
  for (auto& x : part) axpy(F, 1.0, x);
  return F;
}

void three_indirect(const Ints& W, const Ints& V, const ABS& P, bool calc_fock, double k1, Result& R) {
  const int ns = (int)W.Co.size();
  double E = 0;
  for (int s = 0; s < ns; ++s) E += indirect_3el_energy(W, V, P, s);
  R.e3i += k1 * E;
  if (!calc_fock) return;
  for (int s = 0; s < ns; ++s) {  // indirect_3el_fock_matrix, indirect_utils.cpp:297-345
    Mat f = fock_i_sum(W, V, P, s);
    axpy(f, 1.0, fock_i_sum(V, W, P, s));
    axpy(f, 1.0, fock_k_sum(W, V, P, s));
    axpy(R.F[s], k1, symmetrise(f));
  }
}

thread_local std::string g_err;

}  // namespace

struct oracle_system;

extern "C" {

const char* oracle_last_error() { return g_err.c_str(); }

// 0 = BLAS from MKL, 1 = built-in fallback.
int oracle_blas_kind() {
  blas_init();
  return g_dgemm ? 0 : 1;
}

void oracle_force_builtin_blas(int on) {
  g_force_builtin = on;
  g_blas_init = false;
  g_dgemm = nullptr;
  g_dsyevd = nullptr;
}

struct oracle_system {
  Base W, V, WV;
  ABS P;
  int N = 0;
};

// Density-independent setup (the reference's setup_base_integrals x3 and
// setup_abs_projectors, libuw12.cpp:22-56).
//   X3*: packed (n_ao(n_ao+1)/2 x n_df) column-major, X2*: n_df x n_df,
//   X3ri*: (n_ao n_ri x n_df) rows r + n_ri mu; S: (n_ao+n_ri)^2.
int oracle_setup(const double* W3, const double* W2, const double* W3ri, const double* V3, const double* V2,
                 const double* V3ri, const double* WV3, const double* WV2, const double* S, int n_ao, int n_df,
                 int n_ri, oracle_system** out) {
  try {
    auto* sys = new oracle_system();
    sys->N = n_ao;
    base_setup(sys->W, W3, W2, W3ri, n_ao, n_df, n_ri);
    base_setup(sys->V, V3, V2, V3ri, n_ao, n_df, n_ri);
    base_setup(sys->WV, WV3, WV2, nullptr, n_ao, n_df, 0);
    sys->P = abs_setup(S, n_ao, n_ri);
    *out = sys;
    return 0;
  } catch (const std::exception& e) {
    g_err = e.what();
    return -1;
  }
}

void oracle_destroy(oracle_system* sys) { delete sys; }

// form_fock (src/uw12.cpp:22-119) on a set-up system.
//   C: n_ao x (n_orb n_spin) (alpha block then beta); occ: alpha then beta.
//   terms_out[6]: e2 direct, e2 indirect, e3 direct, e3 indirect, e4 direct, e4 indirect.
//   fock_out: n_ao x (n_ao n_spin) or NULL (energy only).
int oracle_evaluate(const oracle_system* sys, const double* C, int n_orb, int n_spin, const double* occ,
                    const int* n_occ, const int* n_active, double k0, double k1, int indirect_term,
                    double* fock_out, double* terms_out) {
  try {
    const int N = sys->N;
    const int ns = n_spin;
    if (ns != 1 && ns != 2) throw std::runtime_error("Number of spin channels must be 1 or 2");
    std::vector<Mat> Co(ns), Ca(ns);
    size_t occ_off = 0;
    for (int s = 0; s < ns; ++s) {  // occupation_weighted_orbitals + freeze_core
      const int o = n_occ[s], a = n_active[s];
      if (a > o) throw std::runtime_error("n_active exceeds n_occ");
      const double g = ns == 1 ? 2.0 : 1.0;
      Co[s] = Mat(N, o);
      for (int k = 0; k < o; ++k) {
        const double w = occ[occ_off + k];
        if (!(w > 0)) throw std::runtime_error("occupations must be positive");
        const double sc = std::sqrt(w / g);
        for (int m = 0; m < N; ++m) Co[s](m, k) = C[m + (size_t)N * ((size_t)s * n_orb + k)] * sc;
      }
      occ_off += o;
      Ca[s] = Mat(N, a);
      for (int k = 0; k < a; ++k) std::memcpy(Ca[s].col(k), Co[s].col(o - a + k), sizeof(double) * N);
    }
    const bool calc_fock = fock_out != nullptr;
    const bool indirect = indirect_term != 0;
    Result R;
    R.F = zeros(ns, N);
    two_electron(sys->WV, Ca, indirect, calc_fock, k0, k1, R);
    Ints WI, VI;
    ints_build(WI, sys->W, Co, Ca, true);
    ints_build(VI, sys->V, Co, Ca, true);
    std::vector<Mat> tab = calc_tab(WI, VI);
    std::vector<Mat> tt = calc_ttilde(WI, VI, tab);
    three_direct(WI, VI, sys->P, calc_fock, k0, k1, R, tt);
    if (indirect && k1 != 0) three_indirect(WI, VI, sys->P, calc_fock, k1, R);
    four_electron(WI, VI, indirect, calc_fock, k0, k1, R, tab, tt);
    if (terms_out) {
      terms_out[0] = R.e2d;
      terms_out[1] = R.e2i;
      terms_out[2] = R.e3d;
      terms_out[3] = R.e3i;
      terms_out[4] = R.e4d;
      terms_out[5] = R.e4i;
    }
    if (fock_out)
      for (int s = 0; s < ns; ++s) std::memcpy(fock_out + (size_t)N * N * s, R.F[s].d.data(), sizeof(double) * N * N);
    return 0;
  } catch (const std::exception& e) {
    g_err = e.what();
    return -1;
  }
}

// Setup + evaluation in one call.
int oracle_uw12(const double* W3, const double* W2, const double* W3ri, const double* V3, const double* V2,
                const double* V3ri, const double* WV3, const double* WV2, const double* S, int n_ao, int n_df,
                int n_ri, const double* C, int n_orb, int n_spin, const double* occ, const int* n_occ,
                const int* n_active, double k0, double k1, int indirect_term, double* fock_out, double* terms_out) {
  oracle_system* sys = nullptr;
  int rc = oracle_setup(W3, W2, W3ri, V3, V2, V3ri, WV3, WV2, S, n_ao, n_df, n_ri, &sys);
  if (rc != 0) return rc;
  rc = oracle_evaluate(sys, C, n_orb, n_spin, occ, n_occ, n_active, k0, k1, indirect_term, fock_out, terms_out);
  oracle_destroy(sys);
  return rc;
}

void oracle_set_pair_stride(int stride) { g_stride = stride < 1 ? 1 : stride; }
void oracle_set_progress(int on) { g_progress = on; }

// The ABS+ projector blocks as one (n_ao+n_ri)^2 matrix, AO first.
int oracle_abs_projector(const double* S, int n_ao, int n_ri, double* out) {
  try {
    ABS P = abs_setup(S, n_ao, n_ri);
    const int M = n_ao + n_ri;
    for (int j = 0; j < M; ++j)
      for (int i = 0; i < M; ++i) {
        double v;
        if (i < n_ao && j < n_ao) v = P.aa(i, j);
        else if (i < n_ao) v = P.ar(i, j - n_ao);
        else if (j < n_ao) v = P.ra(i - n_ao, j);
        else v = P.rr(i - n_ao, j - n_ao);
        out[i + (size_t)M * j] = v;
      }
    return 0;
  } catch (const std::exception& e) {
    g_err = e.what();
    return -1;
  }
}

// Eigen-decomposition used by the oracle (ascending), exported for tests.
int oracle_eig_sym(int n, const double* A, double* w, double* V) {
  try {
    eig_sym(n, A, w, V);
    return 0;
  } catch (const std::exception& e) {
    g_err = e.what();
    return -1;
  }
}
}
