// data_exporter.hpp -- the reference's data_exporter (src/utils/
// data_exporter.h:16-90) over the C ABI of libaperture_io.so
// (include/aperture_io.h).  Header-only: a lab main() that includes it links
// -laperture_io -laperture_amd.  Every failing call throws
// std::runtime_error with apt_io_last_error().
#pragma once
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <filesystem>
#include <fstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "aperture_host.hpp"
#include "aperture_io.h"

namespace Aperture {

class data_exporter {
 public:
  // data_exporter(sim_environment&, uint32_t&) (data_exporter.cpp:80-117):
  // makes the output directory, copies the configuration file, writes
  // grid.h5.
  data_exporter(sim_environment &env, uint32_t &timestep) : m_env(env) {
    (void)timestep;
    const SimParams &p = env.params();
    m_dir = p.data_dir;
    if (m_dir.empty() || m_dir.back() != '/') m_dir.push_back('/');
    std::filesystem::create_directories(m_dir);
    for (int a = 0; a < 3; a++) {
      m_grid.N[a] = p.N[a];
      m_grid.guard[a] = p.guard[a];
      m_grid.lower[a] = p.lower[a];
      m_grid.size[a] = p.size[a];
      m_grid.periodic[a] = p.periodic_boundary[a] ? 1 : 0;
      m_out_ext[a] = a < env.dim() ? p.N[a] / p.downsample : 1;
    }
    copy_config_file();
    write_grid();
  }

  void write_grid() {
    const int coord =
        m_env.params().coord_system == "LogSpherical" ? APT_COORD_LOGSPH : APT_COORD_CARTESIAN;
    io(apt_io_write_grid(&m_grid, coord, dir(), m_env.params().downsample), "write_grid");
  }

  void copy_config_file() {
    const std::string &src = m_env.params().conf_file;
    if (!src.empty() && std::filesystem::exists(src))
      std::filesystem::copy_file(src, m_dir + "config.toml",
                                 std::filesystem::copy_options::overwrite_existing);
  }

  // write_output (data_exporter.cpp:639-665): fld.%05d.h5, data.xmf,
  // ptc.%05d.h5, then the output number advances
  void write_output(sim_data &data, uint32_t timestep, double time) {
    write_field_output(data, timestep, time);
    m_times.push_back(time);
    m_nums.push_back(m_output_num);
    write_xmf();
    write_ptc_output(data, timestep, time);
    m_output_num += 1;
  }

  void write_field_output(sim_data &data, uint32_t, double) {
    const SimParams &p = m_env.params();
    io(apt_io_write_fields(data.ctx, &m_grid, p.num_species, dir(), m_output_num,
                           p.downsample),
       "write_field_output");
  }

  // tracked particles: at most min(capacity, MAX_TRACKED) - 1 of each set
  // (particle_base_impl.cu:243-245, constant_defs.h:34)
  void write_ptc_output(sim_data &data, uint32_t, double) {
    const SimParams &p = m_env.params();
    const uint64_t max_tracked = 1000000;
    auto cap = [&](uint64_t n) { return std::max<uint64_t>(2, std::min(n, max_tracked)); };
    io(apt_io_write_tracked(data.ctx, &m_grid, p.num_species, dir(), m_output_num,
                            cap(p.max_ptc_number), cap(p.max_photon_number)),
       "write_ptc_output");
  }

  void save_snapshot(const std::string &filename, sim_data &data, uint32_t step, Scalar time) {
    const SimParams &p = m_env.params();
    io(apt_io_save_snapshot(data.ctx, filename.c_str(), step, time, (int)p.data_interval,
                            m_output_num, (uint64_t)p.random_seed),
       "save_snapshot");
  }

  // load_snapshot (data_exporter.cpp:490-593): a missing file is reported
  // and the run proceeds without it, as in the reference
  void load_snapshot(const std::string &filename, sim_data &data, uint32_t &step,
                     Scalar &time) {
    if (!std::filesystem::exists(filename)) {
      std::fprintf(stderr, "Can't find restart file, proceeding without loading it!\n");
      return;
    }
    uint32_t s = 0;
    double t = 0.0;
    int num = 0;
    io(apt_io_load_snapshot(data.ctx, &m_grid, filename.c_str(), &s, &t, &num),
       "load_snapshot");
    step = s;
    time = t;
    m_output_num = num;
    prepare_xmf_restart(step, (int)m_env.params().data_interval, (float)time);
  }

  // keeps the outputs of data.xmf up to the one at `time`
  // (data_exporter.cpp:595-636)
  void prepare_xmf_restart(uint32_t, int, float time) {
    m_times.clear();
    m_nums.clear();
    std::ifstream in(m_dir + "data.xmf");
    std::string line;
    const std::string t_tag = "  <Time Type=\"Single\" Value=\"";
    bool in_step = false, found = false, have_t = false, have_num = false;
    double t = 0.0;
    while (std::getline(in, line)) {
      if (line == "<Grid Name=\"quadmesh\" Type=\"Uniform\">") {
        in_step = true;
        have_t = have_num = false;
      } else if (in_step && line.compare(0, t_tag.size(), t_tag) == 0) {
        t = std::stod(line.substr(t_tag.size()));
        have_t = true;
        if (std::fabs((float)t - time) < 1.0e-4f) found = true;
      } else if (in_step && have_t && !have_num) {
        // the first attribute names this output's field file
        const size_t k = line.find("fld.");
        if (k != std::string::npos) {
          m_times.push_back(t);
          m_nums.push_back(std::stoi(line.substr(k + 4, 5)));
          have_num = true;
        }
      }
      if (line == "</Grid>" && in_step) {
        in_step = false;
        if (found) break;
      }
    }
    write_xmf();
  }

  int output_num() const { return m_output_num; }

 private:
  void write_xmf() {
    io(apt_io_write_xmf(dir(), m_env.dim(), m_out_ext, m_times.data(), m_nums.data(),
                        (int)m_times.size(), m_env.params().num_species),
       "write_xmf");
  }
  const char *dir() const { return m_dir.c_str(); }
  static void io(apt_status st, const char *what) {
    if (st != APT_OK)
      throw std::runtime_error(std::string(what) + ": " + apt_io_last_error());
  }

  sim_environment &m_env;
  std::string m_dir;
  apt_grid m_grid{};
  int m_out_ext[3] = {1, 1, 1};
  int m_output_num = 0;
  std::vector<double> m_times;
  std::vector<int> m_nums;
};

}  // namespace Aperture
