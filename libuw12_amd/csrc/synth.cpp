// Synthetic s-Gaussian integral generator (host, FP64, OpenMP).
//
// Produces the inputs of the UW12 evaluation for systems of normalised
// s-type primitives g_a(r) = (2a/pi)^{3/4} exp(-a |r - A|^2):
//   overlap S, and for an operator x(r12) the two-index (A|x|B), the packed
//   three-index (mu nu|x|A) and the RI three-index (r mu|x|A) integrals, in the
//   layouts the reference's setup_base_integrals takes (libuw12.cpp:12-49).
// Operators (kind): 0 = 1/r12, 1 = -sum_i c_i exp(-g_i r12^2) (Gaussian
// geminal), 2 = -sum_i c_i exp(-g_i r12^2) / r12.  All follow from one
// relative-coordinate formula: for charge distributions of exponent p (centre
// P, prefactor Kp) and q (Q, Kq), with al = pq/(p+q) and R = |P - Q|,
//   int int rho_p f(r12) rho_q = Kp Kq (pi^2/(pq))^{3/2} <f>,
//   <1/r>        = 2 sqrt(al/pi) F0(al R^2)
//   <e^{-g r^2}> = (al/(al+g))^{3/2} exp(-al g R^2/(al+g))
//   <e^{-g r^2}/r> = (al/pi)^{3/2} exp(-al g R^2/(al+g)) 2 pi/(al+g) F0(al^2 R^2/(al+g))
// with the Boys function F0(t) = 1/2 sqrt(pi/t) erf(sqrt t) (Helgaker,
// Jorgensen & Olsen, Molecular Electronic-Structure Theory, ch. 9).
#include <cmath>
#include <cstdint>
#include <cstring>

namespace {

constexpr double kPi = 3.14159265358979323846;

struct Geminal {
  int n;
  double c[8], g[8];
};
Geminal g_gem = {2, {0.4, 0.6}, {0.5, 2.0}};

inline double boys0(double t) {
  if (t < 1e-12) return 1.0 - t / 3.0;
  const double st = std::sqrt(t);
  return 0.5 * std::sqrt(kPi) / st * std::erf(st);
}

inline double norm(double a) { return std::pow(2.0 * a / kPi, 0.75); }

// Bare two-centre operator integrals of unit-prefactor Gaussian charge
// distributions exp(-p|r-P|^2), exp(-q|r-Q|^2), R2 = |P-Q|^2:
//   kind 0: 1/r12; kind 1: exp(-g r12^2); kind 2: exp(-g r12^2)/r12.
inline double bare(int kind, double p, double q, double R2, double g) {
  const double al = p * q / (p + q);
  const double pref = std::pow(kPi * kPi / (p * q), 1.5);
  if (kind == 0) return pref * 2.0 * std::sqrt(al / kPi) * boys0(al * R2);
  const double ex = std::exp(-al * g / (al + g) * R2);
  if (kind == 1) return pref * std::pow(al / (al + g), 1.5) * ex;
  return pref * std::pow(al / kPi, 1.5) * ex * (2.0 * kPi / (al + g)) * boys0(al * al * R2 / (al + g));
}

inline double kernel(int kind, double p, double q, double R2) {
  if (kind == 0) return bare(0, p, q, R2, 0.0);
  double out = 0.0;
  for (int i = 0; i < g_gem.n; ++i) out -= g_gem.c[i] * bare(kind, p, q, R2, g_gem.g[i]);
  return out;
}

struct Pair {
  double p, K, P[3];
};

inline Pair make_pair(double a, const double* A, double b, const double* B) {
  Pair r;
  r.p = a + b;
  double R2 = 0.0;
  for (int x = 0; x < 3; ++x) {
    const double d = A[x] - B[x];
    R2 += d * d;
    r.P[x] = (a * A[x] + b * B[x]) / r.p;
  }
  r.K = norm(a) * norm(b) * std::exp(-a * b / r.p * R2);
  return r;
}

inline double dist2(const double* A, const double* B) {
  double s = 0.0;
  for (int x = 0; x < 3; ++x) s += (A[x] - B[x]) * (A[x] - B[x]);
  return s;
}

}  // namespace

extern "C" {

int synth_set_geminal(int n, const double* c, const double* g) {
  if (n < 1 || n > 8) return -1;
  g_gem.n = n;
  for (int i = 0; i < n; ++i) {
    g_gem.c[i] = c[i];
    g_gem.g[i] = g[i];
  }
  return 0;
}

// S[i + n j] = <g_i | g_j>
void synth_overlap(const double* ex, const double* cen, int n, double* S) {
#pragma omp parallel for schedule(dynamic)
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < n; ++i) {
      const Pair pr = make_pair(ex[i], cen + 3 * i, ex[j], cen + 3 * j);
      S[i + (int64_t)n * j] = pr.K * std::pow(kPi / pr.p, 1.5);
    }
}

// X2[A + n B] = (A | x | B)
void synth_two_index(int kind, const double* ex, const double* cen, int n, double* X2) {
#pragma omp parallel for schedule(dynamic)
  for (int B = 0; B < n; ++B)
    for (int A = 0; A < n; ++A)
      X2[A + (int64_t)n * B] =
          norm(ex[A]) * norm(ex[B]) * kernel(kind, ex[A], ex[B], dist2(cen + 3 * A, cen + 3 * B));
}

// X3[mu(mu+1)/2 + nu + P A] = (mu nu | x | A), nu <= mu
void synth_three_index(int kind, const double* eao, const double* cao, int N, const double* edf,
                       const double* cdf, int Nd, double* X3) {
  const int64_t P = (int64_t)N * (N + 1) / 2;
#pragma omp parallel for schedule(dynamic)
  for (int mu = 0; mu < N; ++mu)
    for (int nu = 0; nu <= mu; ++nu) {
      const Pair pr = make_pair(eao[mu], cao + 3 * mu, eao[nu], cao + 3 * nu);
      const int64_t row = (int64_t)mu * (mu + 1) / 2 + nu;
      for (int A = 0; A < Nd; ++A)
        X3[row + P * A] = pr.K * norm(edf[A]) * kernel(kind, pr.p, edf[A], dist2(pr.P, cdf + 3 * A));
    }
}

// X3ri[r + Nr mu + N Nr A] = (r mu | x | A)
void synth_three_index_ri(int kind, const double* eri, const double* cri, int Nr, const double* eao,
                          const double* cao, int N, const double* edf, const double* cdf, int Nd, double* X3ri) {
  const int64_t rows = (int64_t)N * Nr;
#pragma omp parallel for schedule(dynamic) collapse(2)
  for (int mu = 0; mu < N; ++mu)
    for (int r = 0; r < Nr; ++r) {
      const Pair pr = make_pair(eri[r], cri + 3 * r, eao[mu], cao + 3 * mu);
      const int64_t row = r + (int64_t)Nr * mu;
      for (int A = 0; A < Nd; ++A)
        X3ri[row + rows * A] = pr.K * norm(edf[A]) * kernel(kind, pr.p, edf[A], dist2(pr.P, cdf + 3 * A));
    }
}

double synth_bare(int kind, double p, double q, double R2, double g) { return bare(kind, p, q, R2, g); }

// Value of the operator's two-centre integral for two primitives (tests).
double synth_pair_integral(int kind, double p, const double* P, double q, const double* Q) {
  return kernel(kind, p, q, dist2(P, Q)) * norm(p) * norm(q);
}
}
