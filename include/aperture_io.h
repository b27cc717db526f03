/* aperture_io.h -- HDF5 output and snapshots of a PIC context.
 *
 * The on-disk formats of the reference's data_exporter
 * (src/utils/data_exporter.cpp): the snapshot/restart file
 * (save_snapshot / load_snapshot, :399-593), the downsampled field output
 * fld.%05d.h5 with the default field set (user_write_field_output,
 * src/utils/user_data_output.hpp:11-48), the grid file grid.h5
 * (write_grid, :120-216) and the XDMF index data.xmf (:227-388).  One
 * process, one GPU: the parallel-HDF5 offsets of the reference are all 0.
 *
 * Built into libaperture_io.so (links the HDF5 C library); every entry
 * point returns an apt_status and leaves a message for apt_io_last_error().
 */
#ifndef APERTURE_IO_H
#define APERTURE_IO_H

#include "aperture_amd.h"

#ifdef __cplusplus
extern "C" {
#endif

const char *apt_io_last_error(void);

/* data_exporter::save_snapshot (data_exporter.cpp:399-488): E, B and the
 * backgrounds with guard cells ("Ex".."Bz", "E0x".."B0z"), "ptc_num",
 * "ph_num", "step", "time", "data_interval", "num_ranks", "output_num",
 * and every particle / photon array ("ptc_x1", ..., "ph_path_left", ...).
 * The reference stores its XORWOW states ("rand_states"); this build's
 * random streams are counter-based, so their state is the seed and the
 * epochs: "rng_seed", "rt_epoch", "inject_epoch", "ptc_id_next",
 * "ph_id_next".  Values are written in FP64 (exact for both forms). */
apt_status apt_io_save_snapshot(apt_ctx *ctx, const char *filename, uint32_t step,
                                double time, int data_interval, int output_num,
                                uint64_t rng_seed);

/* data_exporter::load_snapshot (:490-593): restores fields, particles,
 * photons and the random-stream counters, erases particles and photons in
 * guard cells (:582-583); returns step, time and output_num.  `grid` is
 * the context's grid and must match the file. */
apt_status apt_io_load_snapshot(apt_ctx *ctx, const apt_grid *grid, const char *filename,
                                uint32_t *step, double *time, int *output_num);

/* data_exporter::write_grid (:120-216): grid.h5 with x1 (, x2, x3) at the
 * downsampled nodes, FP32; log-spherical grids give x1 = r sin(theta),
 * x2 = r cos(theta). */
apt_status apt_io_write_grid(const apt_grid *grid, int coord, const char *dir,
                             int downsample);

/* write_field_output with the default user_write_field_output: fld.%05d.h5
 * holding E1..E3, B1..B3, J1..J3, Rho_e, Rho_p (, Rho_i), divE, divB,
 * photon_produced, pair_produced, each downsampled to FP32 by
 * multi_array::downsample (cuda/core/multi_array.cu:18-68, 296-326). */
apt_status apt_io_write_fields(apt_ctx *ctx, const apt_grid *grid, int num_species,
                               const char *dir, int output_num, int downsample);

/* data_exporter::write_ptc_output (:667-758): ptc.%05d.h5 with the tracked
 * particles (flag bit ParticleFlag::tracked, cell != MAX_CELL) grouped by
 * species as "<type>_id", "<type>_x1".."<type>_x3" (global position,
 * FP32), "<type>_p1".."<type>_p3", and the tracked photons under the type
 * name of index num_species.  At most max_tracked - 1 particles (all
 * species together) and max_tracked_ph - 1 photons, in slot order (the
 * reference collects them with an atomic counter, in no fixed order, and
 * caps at min(capacity, MAX_TRACKED) - 1, particle_base_impl.hpp:84). */
apt_status apt_io_write_tracked(apt_ctx *ctx, const apt_grid *grid, int num_species,
                                const char *dir, int output_num, uint64_t max_tracked,
                                uint64_t max_tracked_ph);

/* data.xmf listing n_outputs outputs at times[] whose field files are
 * fld.<output_nums[o]>.h5 (NULL: 0..n_outputs-1) (write_xmf_head,
 * write_xmf_step_header/close, write_xmf_tail, :227-388; times printed
 * with "%g" as fmt 5.2 prints a double): the whole index is rewritten,
 * which equals the reference's appended file. */
apt_status apt_io_write_xmf(const char *dir, int dim, const int out_ext[3],
                            const double *times, const int *output_nums, int n_outputs,
                            int num_species);

/* H5File::read_vector (hdf_wrapper_impl.hpp): a whole dataset as doubles,
 * e.g. the field-line profile of coef.h5 for apt_1dgr_set_profile
 * (grid_1dgr.cpp:60-66).  With out == NULL only *n is set. */
apt_status apt_io_read_dataset(const char *filename, const char *name, double *out,
                               uint64_t cap, uint64_t *n);

/* multi_array::downsample on a host array: sample every d-th node from
 * `guard` and interpolate from stagger `st` to the node stagger 0b111 with
 * the reference's neighbour offsets.  out holds N/d values per axis. */
apt_status apt_io_downsample(const double *f, const int dims[3], const int guard[3],
                             int dim, int st, int d, float *out);

#ifdef __cplusplus
}
#endif
#endif /* APERTURE_IO_H */
