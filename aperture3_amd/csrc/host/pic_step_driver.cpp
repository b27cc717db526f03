// pic_step_driver.cpp -- a lab main() written against the reference-shaped
// host classes (the loop of labs/box/src/main.cpp:69-92): load particles,
// set fields, then update_particles / update_fields / sort_particles per
// step.  Writes the final particles and fields as raw binary so a test can
// compare them with an independent run.
//
//   pic_step_driver <out.bin> <dim> <N> <ppc> <steps> <precision> <theta>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "aperture_host.hpp"

using namespace Aperture;

int main(int argc, char **argv) {
  if (argc < 8) {
    std::fprintf(stderr,
                 "usage: %s out.bin dim N ppc steps precision theta\n",
                 argv[0]);
    return 2;
  }
  const char *out = argv[1];
  int dim = std::atoi(argv[2]), N = std::atoi(argv[3]);
  int ppc = std::atoi(argv[4]), steps = std::atoi(argv[5]);
  int prec = std::atoi(argv[6]);
  double theta = std::atof(argv[7]);
  try {
    SimParams p;
    p.precision = prec;
    p.num_species = 2;
    p.q_e = 1.0;
    p.data_interval = 1000000;
    p.sort_interval = 20;
    p.random_seed = 43210;
    uint64_t cells = 1;
    for (int d = 0; d < dim; d++) {
      p.N[d] = N;
      p.guard[d] = 2;
      p.size[d] = 1.0;
      p.periodic_boundary[d] = true;
      cells *= (uint64_t)N;
    }
    p.max_ptc_number = cells * 2 * ppc + 64;
    p.delta_t = 0.5 / N;
    sim_environment env(p);
    sim_data data(env);
    field_solver solver(env);
    ptc_updater pusher(env);

    data.init_bg_B_field(0, [](Scalar, Scalar, Scalar) { return 1.0; });
    data.E.initialize(0, [](Scalar x, Scalar, Scalar) {
      return 1.0e-3 * std::sin(2.0 * M_PI * 4.0 * x);
    });
    data.fill_multiplicity(1.0, ppc, theta, theta, 0.0, 0.0);

    for (int step = 0; step < steps; step++) {
      pusher.update_particles(data, p.delta_t, step);
      solver.update_fields(data, p.delta_t, step);
      if (step % p.sort_interval == 0 && step != 0) data.sort_particles();
    }
    uint64_t n = data.num_particles();
    std::vector<uint32_t> cell(n), flag(n);
    std::vector<double> x1(n), p1(n), p2(n), p3(n);
    apt_particles v{};
    v.cell = cell.data();
    v.x1 = x1.data();
    v.p1 = p1.data();
    v.p2 = p2.data();
    v.p3 = p3.data();
    v.flag = flag.data();
    uint64_t got = 0;
    check(data.ctx, apt_ptc_get(data.ctx, &v, n, &got), "apt_ptc_get");
    std::vector<double> e1 = data.E.get(0);
    std::FILE *f = std::fopen(out, "wb");
    if (!f) return 3;
    std::fwrite(&n, 8, 1, f);
    std::fwrite(cell.data(), 4, n, f);
    std::fwrite(x1.data(), 8, n, f);
    std::fwrite(p1.data(), 8, n, f);
    std::fwrite(p2.data(), 8, n, f);
    std::fwrite(p3.data(), 8, n, f);
    uint64_t ne = e1.size();
    std::fwrite(&ne, 8, 1, f);
    std::fwrite(e1.data(), 8, ne, f);
    std::fclose(f);
    std::printf("pic_step_driver: %llu particles after %d steps\n",
                (unsigned long long)n, steps);
  } catch (const std::exception &e) {
    std::fprintf(stderr, "pic_step_driver: %s\n", e.what());
    return 1;
  }
  return 0;
}
