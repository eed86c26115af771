// Evaluation over non-stored integrals: the opposite-spin UW12 terms with no
// three-index tensor of the size of X3_ri resident.
//
// The reference keeps J3 / J3_ri when it can and otherwise recomputes them
// from its ThreeIndexFn callbacks (base_integrals.hpp:52-62 storage flags;
// the direct t~ dX path, direct_utils.cpp:349-436, which loops over pairs of
// DF shells).  Here every three-index block comes from the integrals' device
// generator, in three passes:
//
//  pass 0 (2e direct, two_electron.cpp:11-81): X_D = T_WV^T (G0_WV^T D_a) and
//    F += G0_WV (T_WV v) over DF blocks of the raw AO integrals G0_WV.
//  pass 1 (AO rows, integrals.cpp:184-331, transformations.hpp:33-106):
//    Phi_X[mu, k, A] = sum_B (sum_nu G0_X[mu, nu, B] Co[nu, k]) T_X[B, A] over DF
//    blocks B, then (ik|A) = Ca^T Phi, t_AB, the 4e direct energy and
//    T~ = sum_u f t~_u (four_electron_utils.cpp:33-79).
//  pass 2 (ABS rows, direct_utils.cpp:62-347), blocked over the kept
//    eigenvectors m of the AO+RI overlap S (S+ = U diag(1/s) U^T):
//      Gh_X[m, nu, A] = sum_B (sum_mu' U[mu', m] G0_X[mu', nu, B]) T_X[B, A]
//    is formed block by block, and every term that pairs two ABS indices is a
//    sum over m of products of one block:
//      X_AB    += sum_{m, i} (Gh_W Ca)[m, i, A] (1/s_m) (Gh_V Ca)[m, i, B]
//      F_s     -= sum_{m, A} Gh_W[m, mu, A] (1/s_m) Xi[m, nu, A],
//      Xi[m, nu, A] = sum_B Gh_V[m, nu, B] T~_s[A, B]            (t~ dX Fock).
//  Then the 3e direct energy X . t~ and the X dt~ / 4e direct Fock
//  F += Phi_V (w_k Phi_W (T~ - X~))^T, as in the stored path.
//
// Each pass-2 block regenerates G0_W and G0_V in full (elementwise work,
// small next to the two products above).  The generation products
// (U^T G0 and (.) T) always run in FP64: T carries 1/sqrt|lambda| factors
// whose cancellation a reduced-precision operand would not survive.  The
// mixed form runs Xi on FP32 matrix cores and keeps the rest as in the
// stored path.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <string>
#include <vector>

#include "uw12g_impl.h"

namespace uw12g {

namespace {

// Run a generation product in FP64 whatever the evaluation's precision.
void run_gemm64(uw12g_context* ctx, const GemmDesc& d, const char* stage) {
  const int prec = ctx->prec;
  ctx->prec = 0;
  try {
    run_gemm(ctx, d, stage, true);
  } catch (...) {
    ctx->prec = prec;
    throw;
  }
  ctx->prec = prec;
}

// One block of three-index integrals from the integrals' device generator
// (timed as the "three_index_generator" kernel when profiling is on).
void generate(uw12g_context* ctx, const uw12g_base_integrals* X, int B0, int n, bool with_ri, double* out) {
  const double rows = double(X->N) + (with_ri ? double(X->Nr) : 0.0);
  timed(ctx, with_ri ? "st_generate_abs" : "st_generate_ao", "three_index_generator", 0.0, 8.0 * rows * X->N * n,
        [&] {
          const int rc = X->gen(X->gen_user, size_t(B0), size_t(n), with_ri ? 1 : 0, out, ctx->stream);
          require(rc == 0, "three-index generator failed for DF functions " + std::to_string(B0) + ".." +
                               std::to_string(B0 + n), -6);
        });
  UW12G_HIP(hipGetLastError());
}

}  // namespace

void evaluate_streamed(uw12g_context* ctx, const uw12g_base_integrals* W, const uw12g_base_integrals* V,
                       const uw12g_base_integrals* WV, const uw12g_abs_projectors* P, const std::vector<SpinData>& sd,
                       double k0, double k1, double* Fraw, double* Eacc, double* partial) {
  require(!W->stored && !V->stored, "W and V must both be stored or both be non-stored", -2);
  hipStream_t st = ctx->stream;
  const int ns = int(sd.size());
  const int N = W->N, M = P->M, Mr = P->r, DW = W->D, DV = V->D, DWV = WV->D;
  const int NdW = W->Nd, NdV = V->Nd, NdWV = WV->Nd;
  const size_t NN = size_t(N) * N;
  const bool calc_fock = Fraw != nullptr;
  require(W->Nr == P->Nr && V->Nr == P->Nr && W->Nr > 0, "ri integrals requested, but no function provided to calculate");

  // occupied orbitals of every spin side by side: Coall[:, off[s] + k]
  std::vector<int> off(ns);
  int O = 0;
  for (int s = 0; s < ns; ++s) {
    off[s] = O;
    O += sd[s].o;
  }
  double* Coall = ctx->buf("st_Co", size_t(N) * std::max(O, 1));
  for (int s = 0; s < ns; ++s)
    if (sd[s].o > 0)
      UW12G_HIP(hipMemcpyAsync(Coall + size_t(N) * off[s], sd[s].Co, sizeof(double) * N * sd[s].o,
                               hipMemcpyDeviceToDevice, st));
  auto Ca = [&](int s) { return Coall + size_t(N) * (off[s] + sd[s].c); };

  size_t free0 = 0, total_b = 0;
  UW12G_HIP(hipMemGetInfo(&free0, &total_b));
  // ---- block sizes follow the free device memory (the big buffers of an earlier
  // streamed evaluation were released at its end); a block cap set through the
  // test hook also bounds them (tiny caps force many blocks)
  const bool capped = ctx->block_cap != ctx->block_cap_default;
  const int64_t cap = capped ? int64_t(ctx->block_cap) : INT64_MAX / 4;
  int amax = 0;
  for (int s = 0; s < ns; ++s) amax = std::max(amax, sd[s].a);
  const int NdMax = std::max(std::max(NdW, NdV), NdWV);
  const int bs_ao = int(std::max<int64_t>(
      1, std::min<int64_t>({int64_t(NdMax), 1024, cap / std::max<int64_t>(int64_t(NN), 1)})));

  // ================= pass 0: two-electron direct term =================
  {
    double* DCa = ctx->buf("DCa", NN * ns);
    for (int s = 0; s < ns; ++s) {
      GemmDesc g;  // D_a = Ca Ca^T
      g.M = N;
      g.N = N;
      g.K = sd[s].a;
      g.A = Ca(s);
      g.Am = lin(1);
      g.Ak = lin(N);
      g.B = Ca(s);
      g.Bk = lin(N);
      g.Bn = lin(1);
      g.C = DCa + NN * s;
      g.Cm = lin(1);
      g.Cn = lin(N);
      if (g.K > 0) run_gemm(ctx, g, "2e_density", true);
      else fill(g.C, int64_t(NN), 0.0, st);
    }
    double* XD = ctx->buf("XD", size_t(std::max(DWV, 1)) * ns);
    const int64_t gpart_n = std::max<int64_t>(2 * (4096 + int64_t(std::max(DWV, NdWV))),
                                              calc_fock ? gemv_n_scratch(NN, ns, std::max(DWV, bs_ao)) : 0);
    double* gpart = ctx->buf("gemv_part", size_t(gpart_n));
    double* Gao = nullptr;
    if (WV->stored) {
      timed(ctx, "2e_XD", "gemv_t_kernel", 2.0 * NN * DWV * ns, 8.0 * NN * DWV, [&] {
        gemv_t(WV->gamma.p, int64_t(NN), int64_t(NN), DCa, int64_t(NN), ns, XD, DWV, DWV, gpart, gpart_n, st);
      });
    } else {
      // raw X_D0[B] = sum_{mu nu} G0_WV[mu, nu, B] D_a[mu, nu] over DF blocks, then X_D = T^T X_D0
      Gao = ctx->buf("st_Gao", NN * bs_ao);
      double* XD0 = ctx->buf("st_XD0", size_t(NdWV) * ns);
      for (int B0 = 0; B0 < NdWV; B0 += bs_ao) {
        const int bn = std::min(bs_ao, NdWV - B0);
        generate(ctx, WV, B0, bn, false, Gao);
        timed(ctx, "2e_XD", "gemv_t_kernel", 2.0 * NN * bn * ns, 8.0 * NN * bn, [&] {
          gemv_t(Gao, int64_t(NN), int64_t(NN), DCa, int64_t(NN), ns, XD0 + B0, NdWV, bn, gpart, gpart_n, st);
        });
      }
      GemmDesc g;
      g.M = DWV;
      g.N = ns;
      g.K = NdWV;
      g.A = WV->T.p;
      g.Am = lin(NdWV);
      g.Ak = lin(1);
      g.B = XD0;
      g.Bk = lin(1);
      g.Bn = lin(NdWV);
      g.C = XD;
      g.Cm = lin(1);
      g.Cn = lin(DWV);
      run_gemm(ctx, g, "2e_XD_metric", true);
    }
    double f2[2][2] = {{0.0, 0.0}, {0.0, 0.0}};
    for (int s = 0; s < ns; ++s)
      for (int t = 0; t < ns; ++t) f2[s][t] = spin_factor(ns, s, t, k0, k1);
    double* vd = calc_fock ? ctx->buf("XDv", size_t(std::max(DWV, 1)) * ns) : nullptr;
    if (DWV > 0) xd_combine(XD, DWV, ns, f2, WV->sign.p, vd, Eacc + UW12G_E2_DIRECT, st);
    if (calc_fock && DWV > 0) {
      if (WV->stored) {
        timed(ctx, "2e_fock", "gemv_n_kernel", 2.0 * NN * DWV * ns, 8.0 * NN * DWV, [&] {
          gemv_n(WV->gamma.p, int64_t(NN), int64_t(NN), vd, DWV, ns, Fraw, int64_t(NN), DWV, gpart, gpart_n, st);
        });
      } else {
        double* u = ctx->buf("st_XDu", size_t(NdWV) * ns);
        GemmDesc g;  // u = T v (the raw-basis DF vector)
        g.M = NdWV;
        g.N = ns;
        g.K = DWV;
        g.A = WV->T.p;
        g.Am = lin(1);
        g.Ak = lin(NdWV);
        g.B = vd;
        g.Bk = lin(1);
        g.Bn = lin(DWV);
        g.C = u;
        g.Cm = lin(1);
        g.Cn = lin(NdWV);
        run_gemm(ctx, g, "2e_fock_metric", true);
        for (int B0 = 0; B0 < NdWV; B0 += bs_ao) {
          const int bn = std::min(bs_ao, NdWV - B0);
          generate(ctx, WV, B0, bn, false, Gao);
          timed(ctx, "2e_fock", "gemv_n_kernel", 2.0 * NN * bn * ns, 8.0 * NN * bn, [&] {
            gemv_n(Gao, int64_t(NN), int64_t(NN), u + B0, NdWV, ns, Fraw, int64_t(NN), bn, gpart, gpart_n, st);
          });
        }
      }
    }
  }
  progress(ctx, "streamed: two-electron direct term");

  // ================= pass 1: AO rows, (ik|A), t, 4e direct energy, T~ =================
  double* phi[2] = {ctx->buf("st_phiW", size_t(N) * std::max(O, 1) * std::max(DW, 1)),
                    ctx->buf("st_phiV", size_t(N) * std::max(O, 1) * std::max(DV, 1))};
  {
    double* Gao = ctx->buf("st_Gao", NN * bs_ao);
    double* phi0 = ctx->buf("st_phi0", size_t(N) * std::max(O, 1) * bs_ao);
    for (int op = 0; op < 2; ++op) {
      const uw12g_base_integrals* X = op == 0 ? W : V;
      const int Nd = X->Nd, D = X->D;
      for (int B0 = 0; B0 < Nd && O > 0 && D > 0; B0 += bs_ao) {
        const int bn = std::min(bs_ao, Nd - B0);
        generate(ctx, X, B0, bn, false, Gao);
        GemmDesc g;  // Phi0[mu, k, j] = sum_nu G0[mu, nu, B0 + j] Co[nu, k]
        g.M = N;
        g.N = O;
        g.K = N;
        g.batch = bn;
        g.A = Gao;
        g.Am = lin(1);
        g.Ak = lin(N);
        g.Ab = int64_t(NN);
        g.B = Coall;
        g.Bk = lin(1);
        g.Bn = lin(N);
        g.C = phi0;
        g.Cm = lin(1);
        g.Cn = lin(N);
        g.Cb = int64_t(N) * O;
        run_gemm64(ctx, g, "st_phi0");
        GemmDesc h;  // Phi[(mu, k), A] += sum_j Phi0[(mu, k), j] T[B0 + j, A]
        h.M = N * O;
        h.N = D;
        h.K = bn;
        h.beta = B0 == 0 ? 0.0 : 1.0;
        h.A = phi0;
        h.Am = lin(1);
        h.Ak = lin(int64_t(N) * O);
        h.B = X->T.p + B0;
        h.Bk = lin(1);
        h.Bn = lin(Nd);
        h.C = phi[op];
        h.Cm = lin(1);
        h.Cn = lin(int64_t(N) * O);
        run_gemm64(ctx, h, "st_phi_metric");
      }
    }
  }
  std::vector<double*> t(ns), x(ns), Tb(ns);
  for (int s = 0; s < ns; ++s) {
    const int o = sd[s].o, a = sd[s].a;
    const std::string sfx = std::to_string(s);
    t[s] = ctx->buf("st_t" + sfx, size_t(std::max(DW, 1)) * std::max(DV, 1));
    x[s] = ctx->buf("st_x" + sfx, size_t(std::max(DW, 1)) * std::max(DV, 1));
    Tb[s] = ctx->buf("st_Tb" + sfx, size_t(std::max(DW, 1)) * std::max(DV, 1));
    double* o2[2] = {ctx->buf("st_o2W", size_t(std::max(a * o, 1)) * std::max(DW, 1)),
                     ctx->buf("st_o2V", size_t(std::max(a * o, 1)) * std::max(DV, 1))};
    for (int op = 0; op < 2; ++op) {
      const int D = op == 0 ? DW : DV;
      GemmDesc h;  // (ik|A) = sum_mu Ca[mu, i] Phi[mu, off + k, A]
      h.M = a;
      h.N = o * D;
      h.K = N;
      h.A = Ca(s);
      h.Am = lin(N);
      h.Ak = lin(1);
      h.B = phi[op] + size_t(N) * off[s];
      h.Bk = lin(1);
      h.Bn = two(uint32_t(std::max(o, 1)), N, int64_t(N) * O);
      h.C = o2[op];
      h.Cm = lin(1);
      h.Cn = lin(a);
      run_gemm(ctx, h, "o2_transform", true);
    }
    GemmDesc g;  // t = O2_W^T O2_V  (four_electron_utils.cpp:33-49)
    g.M = DW;
    g.N = DV;
    g.K = a * o;
    g.A = o2[0];
    g.Am = lin(int64_t(a) * o);
    g.Ak = lin(1);
    g.B = o2[1];
    g.Bk = lin(1);
    g.Bn = lin(int64_t(a) * o);
    g.C = t[s];
    g.Cm = lin(1);
    g.Cn = lin(DW);
    if (g.K > 0) run_gemm(ctx, g, "4e_tab", true);
    else fill(t[s], int64_t(DW) * DV, 0.0, st);
  }
  ctx->release("st_o2W");
  ctx->release("st_o2V");
  for (int s = 0; s < ns; ++s)
    for (int u = 0; u < ns; ++u) {
      const double f = spin_factor(ns, s, u, k0, k1);
      if (f == 0) continue;
      DotDesc dd;  // four_electron.cpp:118-119
      dd.n = int64_t(DW) * DV;
      dd.row_w = W->sign.p;
      dd.col_w = V->sign.p;
      dd.ld = DW;
      dd.col_n = DV;
      dd.x = t[u];
      dd.y = t[s];
      dot(dd, 0.5 * f, Eacc + UW12G_E4_DIRECT, 1, partial, st);
    }
  // T~_s = sW sV sum_u c_u t_u (the X part follows after pass 2)
  const double fs = ns == 1 ? 0.5 : 1.0;
  auto coef = [&](int s, int u) { return u < ns ? fs * spin_factor(ns, s, u, k0, k1) : 0.0; };
  double* Zs = ctx->buf("Z", size_t(std::max(DW, 1)) * std::max(DV, 1));
  if (calc_fock)
    for (int s = 0; s < ns; ++s)
      combine_dd(DW, DV, t[0], ns == 2 ? t[1] : nullptr, t[0], ns == 2 ? t[1] : nullptr, coef(s, 0), coef(s, 1),
                 W->sign.p, V->sign.p, Tb[s], Zs, st);
  progress(ctx, "streamed: AO-row transforms, t and the 4e direct energy");

  // ================= pass 2: ABS rows, blocked over S's kept eigenvectors =================
  ctx->release("st_Gao");
  ctx->release("st_phi0");
  if (Mr > 0 && DW > 0 && DV > 0 && O > 0) {
    const int NdX = std::max(NdW, NdV);
    // generator calls of bs DF functions; the metric product (.) T runs over
    // super-blocks of bsup of them (fewer read-modify-write passes over Gh)
    const int bs = int(std::max<int64_t>(1, std::min<int64_t>({int64_t(NdX), 256, cap / std::max<int64_t>(int64_t(M) * N, 1)})));
    const int bsup = int(std::max<int64_t>(bs, std::min<int64_t>(NdX, int64_t(4) * bs) / bs * bs));
    // W's rotated tensor Gh_W = U^T Gamma_W (r x N x D_W, the operator with fewer kept DF
    // directions on the recipe) is kept on the integrals across evaluations when it fits in
    // half of the device memory: pass 2 then regenerates only V
    const int64_t ghw_n = int64_t(Mr) * N * DW;
    const bool cache_w = capped ? ghw_n <= cap : ghw_n * 8 <= int64_t(total_b / 2);
    size_t free_b = 0, total_b2 = 0;
    auto plan_rows = [&](int64_t per_row, int64_t fixed) {
      UW12G_HIP(hipMemGetInfo(&free_b, &total_b2));
      int64_t rows = (int64_t(double(free_b) * 0.85 / 8.0) - fixed) / std::max<int64_t>(per_row, 1);
      rows = std::min<int64_t>(rows, Mr);
      if (capped)
        rows = std::min<int64_t>(rows, std::max<int64_t>(16, cap / std::max<int64_t>(int64_t(N) * std::max(DV, 1), 1)));
      if (rows >= 16) rows = rows / 16 * 16;  // whole 16-deep slabs of the (m, A) sums
      require(rows >= 1, "not enough device memory for one block of the ABS rows", -5);
      return int(rows);
    };
    // Gh[m, nu, A] for ABS rows [m0, m0 + bmm) of X: out[m + ldm nu + ldmn A]
    auto form_gh = [&](const uw12g_base_integrals* X, int m0, int bmm, double* out, int64_t ldm, int64_t ldmn,
                       double* Gfull, double* H) {
      const int Nd = X->Nd, D = X->D;
      const int64_t bN = int64_t(bmm) * N;
      for (int B0 = 0; B0 < Nd; B0 += bsup) {
        const int nsup = std::min(bsup, Nd - B0);
        for (int b0 = B0; b0 < B0 + nsup; b0 += bs) {
          const int bn = std::min(bs, B0 + nsup - b0);
          generate(ctx, X, b0, bn, true, Gfull);
          GemmDesc g;  // H[m, nu, j] = sum_mu' U[mu', m0 + m] G0[mu', nu, b0 + j]
          g.M = bmm;
          g.N = N;
          g.K = M;
          g.batch = bn;
          g.A = P->U.p + size_t(M) * m0;
          g.Am = lin(M);
          g.Ak = lin(1);
          g.B = Gfull;
          g.Bk = lin(1);
          g.Bn = lin(M);
          g.Bb = int64_t(M) * N;
          g.C = H + bN * (b0 - B0);
          g.Cm = lin(1);
          g.Cn = lin(bmm);
          g.Cb = bN;
          run_gemm64(ctx, g, "st_abs_rotate");
        }
        GemmDesc h;  // Gh[(m, nu), A] += sum_j H[(m, nu), j] T[B0 + j, A]
        h.M = int(bN);
        h.N = D;
        h.K = nsup;
        h.beta = B0 == 0 ? 0.0 : 1.0;
        h.A = H;
        h.Am = lin(1);
        h.Ak = lin(bN);
        h.B = X->T.p + B0;
        h.Bk = lin(1);
        h.Bn = lin(Nd);
        h.C = out;
        h.Cm = ldm == bmm ? lin(1) : two(uint32_t(bmm), 1, ldm);
        h.Cn = lin(ldmn);
        run_gemm64(ctx, h, "st_gh_metric");
      }
    };
    if (cache_w && W->gt_id != P->id) {
      W->gt_id = 0;
      W->gt.reset();
      W->gt.get(size_t(ghw_n));
      const int bmb = plan_rows(int64_t(N) * bsup, int64_t(M) * N * bs);
      double* Gfull = ctx->buf("st_G0", size_t(M) * N * bs);
      double* H = ctx->buf("st_H", size_t(bmb) * N * bsup);
      for (int m0 = 0; m0 < Mr; m0 += bmb) form_gh(W, m0, std::min(bmb, Mr - m0), W->gt.p + m0, Mr, int64_t(Mr) * N,
                                                  Gfull, H);
      ctx->release("st_H");
      W->gt_id = P->id;
      progress(ctx, "streamed: rotated W integrals kept (%.1f GB)", ghw_n * 8.0 / 1e9);
    }
    // per ABS row of a block: Gh_V, Xi, H, Psi_W, diag(1/s) Psi_V (and Gh_W when not kept)
    const int64_t per_row = int64_t(N) * (DV + (calc_fock ? DW : 0) + bsup + (cache_w ? 0 : DW)) +
                            int64_t(amax) * (DW + DV);
    const int bmx = plan_rows(per_row, int64_t(M) * N * bs);
    double* Gfull = ctx->buf("st_G0", size_t(M) * N * bs);
    double* H = ctx->buf("st_H", size_t(bmx) * N * bsup);
    double* GhW = cache_w ? nullptr : ctx->buf("st_GhW", size_t(bmx) * N * DW);
    double* GhV = ctx->buf("st_GhV", size_t(bmx) * N * DV);
    double* psiW = ctx->buf("st_psiW", size_t(bmx) * std::max(amax, 1) * DW);
    double* psiV = ctx->buf("st_psiV", size_t(bmx) * std::max(amax, 1) * DV);
    double* Xi = calc_fock ? ctx->buf("st_Xi", size_t(bmx) * N * DW) : nullptr;
    for (int m0 = 0; m0 < Mr; m0 += bmx) {
      const int bmm = std::min(bmx, Mr - m0);
      const int64_t bN = int64_t(bmm) * N;
      // Gh_W of the block: a slice of the kept tensor, or formed here
      const double* ghw = cache_w ? W->gt.p + m0 : GhW;
      const int64_t wm = cache_w ? Mr : bmm, wmn = cache_w ? int64_t(Mr) * N : bN;
      if (!cache_w) form_gh(W, m0, bmm, GhW, bmm, bN, Gfull, H);
      form_gh(V, m0, bmm, GhV, bmm, bN, Gfull, H);
      for (int s = 0; s < ns; ++s) {
        const int a = sd[s].a;
        if (a == 0) {
          if (m0 == 0) fill(x[s], int64_t(DW) * DV, 0.0, st);
        } else {
          for (int op = 0; op < 2; ++op) {
            GemmDesc g;  // Psi[m, i, A] = sum_nu Gh[m, nu, A] Ca[nu, i]
            g.M = bmm;
            g.N = a;
            g.K = N;
            g.batch = op == 0 ? DW : DV;
            g.A = op == 0 ? ghw : GhV;
            g.Am = lin(1);
            g.Ak = lin(op == 0 ? wm : bmm);
            g.Ab = op == 0 ? wmn : bN;
            g.B = Ca(s);
            g.Bk = lin(1);
            g.Bn = lin(N);
            g.C = op == 0 ? psiW : psiV;
            g.Cm = lin(1);
            g.Cn = lin(bmm);
            g.Cb = int64_t(bmm) * a;
            run_gemm(ctx, g, "psi_transform", true);
          }
          // diag(1/s) Psi_V in place
          scale_rows(psiV, bmm, a, DV, bmm, int64_t(bmm) * a, P->dinv.p + m0, nullptr, psiV, st);
          GemmDesc xg;  // X[A, B] += sum_{m, i} Psi_W[m, i, A] (1/s_m) Psi_V[m, i, B]  (direct_utils.cpp:62-133)
          xg.M = DW;
          xg.N = DV;
          xg.K = bmm * a;
          xg.beta = m0 == 0 ? 0.0 : 1.0;
          xg.A = psiW;
          xg.Am = lin(int64_t(bmm) * a);
          xg.Ak = lin(1);
          xg.B = psiV;
          xg.Bk = lin(1);
          xg.Bn = lin(int64_t(bmm) * a);
          xg.C = x[s];
          xg.Cm = lin(1);
          xg.Cn = lin(DW);
          run_gemm(ctx, xg, "3e_xab", true);
        }
        if (!calc_fock) continue;
        GemmDesc a1;  // Xi[(m, nu), A] = sum_B Gh_V[(m, nu), B] T~_s[A, B]
        a1.M = int(bN);
        a1.N = DW;
        a1.K = DV;
        a1.A = GhV;
        a1.Am = lin(1);
        a1.Ak = lin(bN);
        a1.B = Tb[s];
        a1.Bk = lin(DW);
        a1.Bn = lin(1);
        a1.C = Xi;
        a1.Cm = lin(1);
        a1.Cn = lin(bN);
        run_gemm(ctx, a1, "3e_ttdx_xi");
        GemmDesc a3;  // F -= sum_{m, A} Gh_W[m, mu, A] (1/s_m) Xi[m, nu, A]  (direct_utils.cpp:242-347)
        a3.M = N;
        a3.N = N;
        a3.K = bmm * DW;
        a3.alpha = -1.0;
        a3.beta = 1.0;
        a3.A = ghw;
        a3.Am = lin(wm);
        a3.Ak = two(uint32_t(bmm), 1, wmn);
        a3.ks = P->dinv.p + m0;
        a3.Km = two(uint32_t(bmm), 1, 0);
        a3.B = Xi;
        a3.Bk = two(uint32_t(bmm), 1, bN);
        a3.Bn = lin(bmm);
        a3.C = Fraw + NN * s;
        a3.Cm = lin(1);
        a3.Cn = lin(N);
        run_gemm(ctx, a3, "3e_ttdx_fock", true);
      }
      progress(ctx, "streamed: ABS rows %d-%d of %d", m0, m0 + bmm, Mr);
    }
    for (const char* nm : {"st_G0", "st_H", "st_GhW", "st_GhV", "st_psiW", "st_psiV", "st_Xi"}) ctx->release(nm);
  } else {
    for (int s = 0; s < ns; ++s) fill(x[s], int64_t(DW) * DV, 0.0, st);
  }

  // 3e direct energy (direct_utils.cpp:136-157)
  for (int s = 0; s < ns; ++s)
    for (int u = 0; u < ns; ++u) {
      const double f = spin_factor(ns, s, u, k0, k1);
      if (f == 0) continue;
      DotDesc dd;
      dd.n = int64_t(DW) * DV;
      dd.row_w = W->sign.p;
      dd.col_w = V->sign.p;
      dd.ld = DW;
      dd.col_n = DV;
      dd.x = x[s];
      dd.y = t[u];
      dot(dd, -f, Eacc + UW12G_E3_DIRECT, 1, partial, st);
    }
  if (!calc_fock) {
    ctx->release("st_phiW");
    ctx->release("st_phiV");
    return;
  }

  // X dt~ and 4e direct Fock: F += Phi_V (w_k Phi_W Z)^T, Z = T~ - sW sV sum_u c_u X_u
  // (four_electron.cpp:11-77, direct_utils.cpp:164-204)
  for (int s = 0; s < ns; ++s) {
    const int o = sd[s].o;
    if (o == 0) continue;
    combine_dd(DW, DV, t[0], ns == 2 ? t[1] : nullptr, x[0], ns == 2 ? x[1] : nullptr, coef(s, 0), coef(s, 1),
               W->sign.p, V->sign.p, Tb[s], Zs, st);
    double* G = ctx->buf("G", size_t(N) * o * DV);
    GemmDesc g;  // G[(beta, k), B] = sum_A Phi_W[beta, off + k, A] Z[A, B]
    g.M = N * o;
    g.N = DV;
    g.K = DW;
    g.A = phi[0] + size_t(N) * off[s];
    g.Am = lin(1);
    g.Ak = lin(int64_t(N) * O);
    g.B = Zs;
    g.Bk = lin(1);
    g.Bn = lin(DW);
    g.C = G;
    g.Cm = lin(1);
    g.Cn = lin(int64_t(N) * o);
    run_gemm(ctx, g, "direct_fock_G");
    GemmDesc f;  // F += sum_{k, B} Phi_V[alpha, off + k, B] w_k G[beta, k, B]
    f.M = N;
    f.N = N;
    f.K = o * DV;
    f.beta = 1.0;
    f.A = phi[1] + size_t(N) * off[s];
    f.Am = lin(1);
    f.Ak = two(uint32_t(o), N, int64_t(N) * O);
    f.ks = sd[s].wk;
    f.Km = two(uint32_t(o), 1, 0);
    f.B = G;
    f.Bk = lin(N);
    f.Bn = lin(1);
    f.C = Fraw + NN * s;
    f.Cm = lin(1);
    f.Cn = lin(N);
    run_gemm(ctx, f, "direct_fock_F", true);
  }
  progress(ctx, "streamed: direct Fock");
  for (const char* nm : {"st_phiW", "st_phiV", "G"}) ctx->release(nm);
}

}  // namespace uw12g
