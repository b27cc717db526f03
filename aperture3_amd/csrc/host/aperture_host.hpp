// aperture_host.hpp -- C++ host classes with the reference's interfaces
// (src/algorithms/ptc_updater.h:11-24, src/algorithms/field_solver.h:11-25,
// src/sim_data.h:14-70, src/sim_environment.h), so a lab main() written
// against the reference compiles against these.  They own no numerics: every
// method forwards to the C ABI in include/aperture_amd.h, and a failing call
// throws std::runtime_error with apt_last_error() (the reference throws from
// CudaSafeCall, cudaUtility.h:62-74).
#pragma once
#include <cstdint>
#include <functional>
#include <stdexcept>
#include <string>
#include <vector>

#include "aperture_amd.h"

namespace Aperture {

typedef double Scalar;

// The SimParams fields this path reads (sim_params.h:17-123).
struct SimParams {
  Scalar delta_t = 0.01;
  Scalar q_e = 1.0;
  uint64_t max_ptc_number = 100;
  Scalar ion_mass = 1.0;
  int num_species = 2;
  bool periodic_boundary[3] = {false, false, false};
  int current_smoothing = 0;
  size_t max_steps = 1000;
  size_t data_interval = 100;
  int random_seed = 43210;
  int sort_interval = 20;
  int N[3] = {1, 1, 1};
  int guard[3] = {0, 0, 0};
  Scalar lower[3] = {0.0, 0.0, 0.0};
  Scalar size[3] = {1.0, 1.0, 1.0};
  // not in the reference: numeric form and mixed-form tile
  int precision = APT_PREC_FP64;
  int tile[3] = {0, 0, 0};
  int tile_halo = 0;
  int tile_threads = 0;
};

void check(apt_ctx *ctx, apt_status st, const char *what);

// Replaces sim_environment(int*, char***) (sim_environment_impl.hpp:17-65)
// on one GPU: no MPI, the configuration comes in as a struct.
class sim_environment {
 public:
  explicit sim_environment(const SimParams &params) : m_params(params) {}
  const SimParams &params() const { return m_params; }
  SimParams &params() { return m_params; }
  int dim() const;
  // Quadmesh::pos(i, n, stagger) (quadmesh.h:119-150)
  Scalar pos(int i, int n, int stagger) const;

 private:
  SimParams m_params;
};

// Field handle with field<N,T>::initialize(component, f) (fields.h:111-143)
class field_ref {
 public:
  field_ref(struct sim_data *data, int which) : m_data(data), m_which(which) {}
  void initialize(int component,
                  const std::function<Scalar(Scalar, Scalar, Scalar)> &f);
  void initialize();  // zero every component
  std::vector<Scalar> get(int component) const;
  void set(int component, const std::vector<Scalar> &v);

 private:
  struct sim_data *m_data;
  int m_which;
};

struct sim_data {
  explicit sim_data(sim_environment &env);
  ~sim_data();
  sim_data(const sim_data &) = delete;
  sim_data &operator=(const sim_data &) = delete;

  // sim_data::fill_multiplicity (cuda/sim_data.cu:268-293); this build
  // also takes Maxwell-Juttner temperatures and x-drifts per species.
  void fill_multiplicity(Scalar weight, int multiplicity,
                         Scalar theta_e = 0.0, Scalar theta_p = 0.0,
                         Scalar drift_e = 0.0, Scalar drift_p = 0.0);
  // sim_data.h:23-36: same function into the background and the field
  void init_bg_B_field(int component,
                       const std::function<Scalar(Scalar, Scalar, Scalar)> &f);
  void init_bg_E_field(int component,
                       const std::function<Scalar(Scalar, Scalar, Scalar)> &f);
  void sort_particles();
  uint64_t num_particles() const;

  sim_environment &env;
  apt_ctx *ctx = nullptr;
  field_ref E, B, Ebg, Bbg, J, divE, divB;
};

class ptc_updater {
 public:
  explicit ptc_updater(sim_environment &env) : m_env(env) {}
  void update_particles(sim_data &data, double dt, uint32_t step = 0);
  void apply_boundary(sim_data &data, double dt, uint32_t step = 0);
  void smooth_current(sim_data &data, uint32_t step = 0);

 protected:
  sim_environment &m_env;
};

class field_solver {
 public:
  explicit field_solver(sim_environment &env) : m_env(env) {}
  void update_fields(sim_data &data, double dt, uint32_t step = 0);
  void apply_outflow_boundary(sim_data &data, double time = 0.0);
  void compute_divs(sim_data &data);
  void update_e_field(sim_data &data, double dt);
  void update_b_field(sim_data &data, double dt);

 private:
  sim_environment &m_env;
};

}  // namespace Aperture
