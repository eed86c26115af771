// C++ caller of the libuw12-style API (include/uw12g.hpp), written the way
// the reference's own example (example/example.cpp:109-219) drives libuw12:
// set up W, V, WV integrals and the ABS+ projectors, then call uw12_energy and
// uw12_fock and check that both give the same energy.
//
// Usage: uw12_cpp_driver <dir> [static]
//   static: keep the integral and projector handles in static storage, which is
//   destroyed after the header's function-local context at exit
//   <dir>/dims.txt : n_ao n_df n_ri n_orb n_spin n_occ_a n_occ_b n_act_a n_act_b k0 k1
//   <dir>/{W3,W2,W3ri,V3,V2,V3ri,WV3,WV2,S,C,occ}.bin : raw float64, column-major
// Writes <dir>/energy.txt (uw12_energy, uw12_fock energies) and <dir>/fock.bin.
#include <cmath>
#include <cstdio>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include "uw12g.hpp"

static std::vector<double> load(const std::string& path, size_t n) {
  std::vector<double> v(n);
  std::ifstream f(path, std::ios::binary);
  if (!f.read(reinterpret_cast<char*>(v.data()), std::streamsize(n * sizeof(double))))
    throw std::runtime_error("cannot read " + path);
  return v;
}

// Handles in static storage: constructed before the library's context (a
// function-local static created on first use), hence destroyed after it.
static uw12::BaseIntegrals g_W, g_V, g_WV;
static uw12::ABSProjectors g_abs;

int main(int argc, char** argv) {
  if (argc != 2 && !(argc == 3 && std::string(argv[2]) == "static")) {
    std::cerr << "usage: uw12_cpp_driver <dir> [static]\n";
    return 2;
  }
  const std::string dir = argv[1];
  const bool keep_static = argc == 3;
  try {
    std::ifstream d(dir + "/dims.txt");
    size_t n_ao, n_df, n_ri, n_orb, n_spin, noa, nob, naa, nab;
    double k0, k1;
    if (!(d >> n_ao >> n_df >> n_ri >> n_orb >> n_spin >> noa >> nob >> naa >> nab >> k0 >> k1))
      throw std::runtime_error("bad dims.txt");
    const size_t P = n_ao * (n_ao + 1) / 2, M = n_ao + n_ri;
    auto W3 = load(dir + "/W3.bin", P * n_df), W2 = load(dir + "/W2.bin", n_df * n_df);
    auto W3ri = load(dir + "/W3ri.bin", n_ao * n_ri * n_df);
    auto V3 = load(dir + "/V3.bin", P * n_df), V2 = load(dir + "/V2.bin", n_df * n_df);
    auto V3ri = load(dir + "/V3ri.bin", n_ao * n_ri * n_df);
    auto WV3 = load(dir + "/WV3.bin", P * n_df), WV2 = load(dir + "/WV2.bin", n_df * n_df);
    auto S = load(dir + "/S.bin", M * M);
    auto C = load(dir + "/C.bin", n_ao * n_orb * n_spin);
    auto occv = load(dir + "/occ.bin", noa + nob);

    const auto W = uw12::setup_base_integrals(W3.data(), W2.data(), W3ri.data(), n_ao, n_df, n_ri);
    const auto V = uw12::setup_base_integrals(V3.data(), V2.data(), V3ri.data(), n_ao, n_df, n_ri);
    const auto WV = uw12::setup_base_integrals(WV3.data(), WV2.data(), n_ao, n_df);
    const auto abs = uw12::setup_abs_projectors(S.data(), n_ao, n_ri);
    const auto orbitals = uw12::setup_orbitals(C.data(), n_ao, n_orb, n_spin);
    const auto occ = n_spin == 1 ? uw12::setup_occupations(occv.data(), noa)
                                 : uw12::setup_occupations(occv.data(), noa, nob);
    const std::vector<size_t> n_active = n_spin == 1 ? std::vector<size_t>{naa} : std::vector<size_t>{naa, nab};
    if (keep_static) {
      g_W = W;
      g_V = V;
      g_WV = WV;
      g_abs = abs;
    }

    const double energy = uw12::uw12_energy(W, V, WV, abs, orbitals, occ, n_active, k0, k1);
    std::vector<double> fock(n_ao * n_ao * n_spin);
    const double energy2 = uw12::uw12_fock(fock.data(), W, V, WV, abs, orbitals, occ, n_active, k0, k1, 1);
    if (std::fabs(energy - energy2) > 1e-12 * std::fabs(energy)) {
      std::cerr << "energy and Fock-path energy differ\n";
      return 1;
    }
    std::ofstream e(dir + "/energy.txt");
    e.precision(17);
    e << energy << " " << energy2 << "\n";
    std::ofstream f(dir + "/fock.bin", std::ios::binary);
    f.write(reinterpret_cast<const char*>(fock.data()), std::streamsize(fock.size() * sizeof(double)));

    // the reference throws on inconsistent spin channels (tests/test_uw12.cpp:96-112)
    bool threw = false;
    try {
      uw12::uw12_energy(W, V, WV, abs, orbitals, occ, {naa, nab, 0}, k0, k1);
    } catch (const std::runtime_error&) {
      threw = true;
    }
    if (!threw) {
      std::cerr << "expected an exception for inconsistent n_active\n";
      return 1;
    }
    return 0;
  } catch (const std::exception& ex) {
    std::cerr << "error: " << ex.what() << "\n";
    return 1;
  }
}
