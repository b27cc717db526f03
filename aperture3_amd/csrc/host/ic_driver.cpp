// ic_driver.cpp -- the scenario of the reference's tests/test_ic.cu:24-69 on
// the C++ host mirror: inverse-Compton tables for a broken power law of
// soft photons (alpha 1.25, delta 1.1, peak at e_min = 1e-5, 1e-10 .. 0.1,
// n0 = 1.5e24 / ic_path with ic_path = 0.01), then the energies of photons
// scattered by 100 electrons at gamma = 10, one per line (%.17g), followed
// by the triplet rate table of the same spectrum.
//
//   ic_driver [n_gamma n_ep seed]
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "aperture_host.hpp"

using namespace Aperture;

int main(int argc, char **argv) {
  SimParams params;
  params.e_min = 1.0e-5f;
  if (argc > 2) {
    params.n_gamma = std::atoi(argv[1]);
    params.n_ep = std::atoi(argv[2]);
  }
  const uint64_t seed = argc > 3 ? std::strtoull(argv[3], nullptr, 10) : 1234;
  const double ic_path = 0.01;
  try {
    inverse_compton rt_ic(params);
    Spectra::broken_power_law rt_ne(1.25, 1.1, params.e_min, 1.0e-10, 0.1);
    rt_ic.init(rt_ne, rt_ne.emin(), rt_ne.emax(), 1.50e24 / ic_path);
    std::vector<double> ptc_E(100, 10.0);
    std::vector<double> ph_E = rt_ic.gen_photon_e(ptc_E, seed);
    for (double e : ph_E) std::printf("%.17g\n", e);
    triplet_pairs tpp(params);
    tpp.init(rt_ne, rt_ne.emin(), rt_ne.emax(), 1.50e24 / ic_path);
    for (double r : tpp.rate()) std::printf("%.17g\n", r);
  } catch (const std::exception &e) {
    std::fprintf(stderr, "ic_driver: %s\n", e.what());
    return 1;
  }
  return 0;
}
