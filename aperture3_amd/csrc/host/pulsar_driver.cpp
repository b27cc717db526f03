// pulsar_driver.cpp -- a lab main() in the shape of labs/pulsar/src/main.cpp
// (restart from a snapshot, output every data_interval, injection, particle
// and field update, pair creation, sort, snapshot every snapshot_interval),
// written against the reference-shaped host classes and data_exporter.  A
// 2D log-spherical box, FP64 form; the state after the last step is saved
// to <data_dir>/final.h5 so a test can compare a restarted run with an
// uninterrupted one.
//
//   pulsar_driver <data_dir> <max_steps> [restart_file]
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>

#include "aperture_host.hpp"
#include "data_exporter.hpp"

using namespace Aperture;

int main(int argc, char **argv) {
  if (argc < 3) {
    std::fprintf(stderr, "usage: %s data_dir max_steps [restart_file]\n", argv[0]);
    return 2;
  }
  try {
    SimParams p;
    p.data_dir = argv[1];
    p.max_steps = (size_t)std::atol(argv[2]);
    if (argc > 3) {
      p.is_restart = true;
      p.restart_file = argv[3];
    }
    p.precision = APT_PREC_FP64;
    p.coord_system = "LogSpherical";
    p.N[0] = 64;
    p.N[1] = 64;
    p.guard[0] = p.guard[1] = 3;
    p.size[0] = std::log(30.0);
    p.size[1] = M_PI;
    p.num_species = 3;
    p.max_ptc_number = 400000;
    p.max_photon_number = 200000;
    p.data_interval = 4;
    p.snapshot_interval = 6;
    p.sort_interval = 5;
    p.downsample = 2;
    p.omega = 0.1;
    p.B0 = 1.0e3;
    p.damping_length = 8;
    p.damping_coef = 0.01;
    p.rt_model = APT_RT_EXP_PATH;
    p.gamma_thr = 5.0;
    p.photon_path = 0.1;
    p.e_frac = 0.2;
    p.r_cutoff = 30.0;
    p.delta_t = 0.3 * p.size[0] / p.N[0];

    uint32_t start_step = 0, step = start_step;
    Scalar start_time = 0.0;
    sim_environment env(p);
    sim_data data(env);
    data_exporter exporter(env, start_step);
    if (p.is_restart) {
      exporter.load_snapshot(p.restart_file, data, start_step, start_time);
      step = start_step + 1;
    } else {
      const Scalar B0 = p.B0;
      data.init_bg_B_field(0, [B0](Scalar x1, Scalar x2, Scalar) {
        Scalar r = std::exp(x1);
        return 2.0 * B0 * std::cos(x2) / (r * r * r);
      });
      data.init_bg_B_field(1, [B0](Scalar x1, Scalar x2, Scalar) {
        Scalar r = std::exp(x1);
        return B0 * std::sin(x2) / (r * r * r);
      });
      data.fill_multiplicity(1.0, 2);
    }
    field_solver_logsph field_solver(env);
    ptc_updater_logsph ptc_updater(env);
    radiative_transfer rad(env);

    const double dt = p.delta_t;
    for (; step <= p.max_steps; step++) {
      // time from the step number alone, so that a restarted run takes
      // the same times as an uninterrupted one
      const double time = step * dt;
      if (step % p.data_interval == 0) {
        exporter.write_output(data, step, time);
        data.photon_produced.initialize();
        data.pair_produced.initialize();
        data.EdotB.initialize();
      }
      ptc_updater.inject_ptc(data, 1, 0.01, 0.0, 0.0, 1.0, p.omega);
      ptc_updater.update_particles(data, dt, step);
      field_solver.update_fields(data, dt, time);
      field_solver.apply_boundary(data, p.omega, time);
      rad.emit_photons(data);
      rad.produce_pairs(data);
      if (step % p.sort_interval == 0 && step != 0) data.sort_particles();
      if (step % p.snapshot_interval == 0 && step > 0)
        exporter.save_snapshot(p.data_dir + "/snapshot.h5", data, step, time);
    }
    exporter.save_snapshot(p.data_dir + "/final.h5", data, (uint32_t)p.max_steps,
                           p.max_steps * dt);
    std::printf("pulsar_driver: %llu particles after step %zu\n",
                (unsigned long long)data.num_particles(), p.max_steps);
  } catch (const std::exception &e) {
    std::fprintf(stderr, "pulsar_driver: %s\n", e.what());
    return 1;
  }
  return 0;
}
