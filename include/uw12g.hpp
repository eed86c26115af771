// uw12g.hpp -- C++ mirror of the libuw12 public API (zwill22/libuw12,
// libuw12.hpp:12-171) over the C ABI of include/uw12g.h.
//
// Same names, argument order, layouts and error behaviour as the reference:
// a caller that includes this header instead of libuw12.hpp and links
// libuw12g.so keeps its code unchanged, and the work runs on one MI355X.
// Failures throw std::runtime_error with the library's message, as the
// reference's setup/evaluation functions do.
#ifndef UW12G_HPP
#define UW12G_HPP

#include <cstdlib>
#include <cstdio>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "uw12g.h"

namespace uw12 {

namespace detail {

inline void check(int rc) {
  if (rc != 0) throw std::runtime_error(std::string("uw12g: ") + uw12g_last_error());
}

// Process-wide device context (device from $UW12G_DEVICE, default 0).
inline uw12g_context* context() {
  static std::shared_ptr<uw12g_context> ctx = [] {
    const char* env = std::getenv("UW12G_DEVICE");
    uw12g_context* c = nullptr;
    check(uw12g_context_create(env ? std::atoi(env) : 0, &c));
    return std::shared_ptr<uw12g_context>(c, uw12g_context_destroy);
  }();
  return ctx.get();
}

}  // namespace detail

/// Device-resident DF integrals of one operator (base_integrals.hpp:46-293).
class BaseIntegrals {
 public:
  BaseIntegrals() = default;
  BaseIntegrals(uw12g_base_integrals* h, size_t n_ao, size_t n_df, size_t n_ri)
      : h_(h, uw12g_base_integrals_destroy), n_ao_(n_ao), n_df_(n_df), n_ri_(n_ri) {}
  const uw12g_base_integrals* handle() const { return h_.get(); }
  size_t get_number_ao() const { return n_ao_; }
  size_t get_number_df() const { return n_df_; }
  size_t get_number_ri() const { return n_ri_; }

 private:
  std::shared_ptr<uw12g_base_integrals> h_;
  size_t n_ao_ = 0, n_df_ = 0, n_ri_ = 0;
};

/// S^+ blocks of the AO + RI space (ri_utils.hpp:20-74).
class ABSProjectors {
 public:
  ABSProjectors() = default;
  explicit ABSProjectors(uw12g_abs_projectors* h) : h_(h, uw12g_abs_projectors_destroy) {}
  const uw12g_abs_projectors* handle() const { return h_.get(); }

 private:
  std::shared_ptr<uw12g_abs_projectors> h_;
};

/// Orbital coefficients, n_ao x (n_orb n_spin), alpha block then beta (libuw12.hpp:82-92).
struct Orbitals {
  std::vector<double> C;
  size_t n_ao = 0, n_orb = 0, n_spin = 0;
  size_t size() const { return n_spin; }
};

/// Occupation numbers, alpha then beta (libuw12.hpp:94-113).
struct Occupations {
  std::vector<double> occ;
  std::vector<size_t> n_occ;
  size_t size() const { return n_occ.size(); }
};

inline BaseIntegrals setup_base_integrals(const double* X3, const double* X2, const double* X3_ri, size_t n_ao,
                                          size_t n_df, size_t n_ri, bool /*copy_data*/ = false) {
  uw12g_base_integrals* h = nullptr;
  detail::check(uw12g_setup_base_integrals(detail::context(), X3, X2, X3_ri, n_ao, n_df, n_ri, &h));
  return BaseIntegrals(h, n_ao, n_df, n_ri);
}

inline BaseIntegrals setup_base_integrals(const double* X3, const double* X2, size_t n_ao, size_t n_df,
                                          bool /*copy_data*/ = false) {
  uw12g_base_integrals* h = nullptr;
  detail::check(uw12g_setup_base_integrals(detail::context(), X3, X2, nullptr, n_ao, n_df, 0, &h));
  return BaseIntegrals(h, n_ao, n_df, 0);
}

inline ABSProjectors setup_abs_projectors(const double* S, size_t n_ao, size_t n_ri) {
  uw12g_abs_projectors* h = nullptr;
  detail::check(uw12g_setup_abs_projectors(detail::context(), S, n_ao, n_ri, &h));
  return ABSProjectors(h);
}

inline Orbitals setup_orbitals(const double* C, size_t n_ao, size_t n_orb, size_t n_spin) {
  if (n_spin != 1 && n_spin != 2) throw std::runtime_error("Number of spin channels must be 1 or 2");
  Orbitals o;
  o.C.assign(C, C + n_ao * n_orb * n_spin);
  o.n_ao = n_ao;
  o.n_orb = n_orb;
  o.n_spin = n_spin;
  return o;
}

inline Occupations setup_occupations(const double* occ, size_t n_occ) {
  Occupations o;
  o.occ.assign(occ, occ + n_occ);
  o.n_occ = {n_occ};
  return o;
}

inline Occupations setup_occupations(const double* occ, size_t n_occ_alpha, size_t n_occ_beta) {
  Occupations o;
  o.occ.assign(occ, occ + n_occ_alpha + n_occ_beta);
  o.n_occ = {n_occ_alpha, n_occ_beta};
  return o;
}

namespace detail {
inline void check_channels(const Orbitals& orb, const Occupations& occ, const std::vector<size_t>& n_active) {
  if (occ.size() != orb.n_spin)
    throw std::runtime_error("Different number of spin channels for orbitals and occupations");
  if (n_active.size() != orb.n_spin)
    throw std::runtime_error("Different number of spin channels in n_active and orbitals");
}

// print_level > 1: the library's progress lines for the duration of one call
struct ProgressScope {
  bool on;
  explicit ProgressScope(size_t print_level) : on(print_level > 1) {
    if (on) check(uw12g_context_set_verbosity(context(), 1));
  }
  ~ProgressScope() {
    if (on) (void)uw12g_context_set_verbosity(context(), 0);
  }
};
}  // namespace detail

inline double uw12_energy(const BaseIntegrals& W, const BaseIntegrals& V, const BaseIntegrals& WV,
                          const ABSProjectors& abs_projectors, const Orbitals& orbitals, const Occupations& occ,
                          const std::vector<size_t>& n_active, double scale_opp_spin, double scale_same_spin,
                          size_t print_level = 0) {
  detail::check_channels(orbitals, occ, n_active);
  detail::ProgressScope progress(print_level);
  double e = 0.0;
  detail::check(uw12g_energy(detail::context(), W.handle(), V.handle(), WV.handle(), abs_projectors.handle(),
                             orbitals.C.data(), orbitals.n_ao, orbitals.n_orb, orbitals.n_spin, occ.occ.data(),
                             occ.n_occ.data(), n_active.data(), scale_opp_spin, scale_same_spin, &e));
  if (print_level > 0) std::printf("Total UW12 energy %20.12f\n", e);
  return e;
}

inline double uw12_fock(double* fock, const BaseIntegrals& W, const BaseIntegrals& V, const BaseIntegrals& WV,
                        const ABSProjectors& abs_projectors, const Orbitals& orbitals, const Occupations& occ,
                        const std::vector<size_t>& n_active, double scale_opp_spin, double scale_same_spin,
                        size_t print_level = 0) {
  detail::check_channels(orbitals, occ, n_active);
  detail::ProgressScope progress(print_level);
  double e = 0.0;
  detail::check(uw12g_fock(detail::context(), fock, W.handle(), V.handle(), WV.handle(), abs_projectors.handle(),
                           orbitals.C.data(), orbitals.n_ao, orbitals.n_orb, orbitals.n_spin, occ.occ.data(),
                           occ.n_occ.data(), n_active.data(), scale_opp_spin, scale_same_spin, &e));
  if (print_level > 0) std::printf("Total UW12 energy %20.12f\n", e);
  return e;
}

}  // namespace uw12

#endif  // UW12G_HPP
