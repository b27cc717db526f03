// apt_kernels.hpp -- device pointer bundles and host launchers of every
// kernel in the library (the role data_ptrs plays in the reference,
// cuda/data_ptrs.h:11-29, but passed per launch and split by precision).
#pragma once
#include "apt_common.hpp"

namespace apt {

enum CounterIdx {
  CTR_NAN = 0,        // NaN momentum seen by the push
  CTR_MULTICELL = 1,  // a particle moved more than one cell in a step
  CTR_OVERFLOW = 2,   // a particle could not be stored (capacity)
  CTR_OUTSIDE = 3,    // a particle left the deposit window of its tile
  CTR_NUM = 8
};

// ---- plain FP64 form ------------------------------------------------
struct PtcF64 {
  uint32_t *cell;
  double *x1, *x2, *x3;
  double *p1, *p2, *p3;
  double *E;
  double *w;
  uint32_t *flag;
  uint64_t *id;
};

struct FieldsF64 {
  const double *E[3];
  const double *B[3];
};

struct AccF64 {
  long long *J[3][2];
  long long *R[4][2];
};

hipError_t launch_push_deposit_f64(const Mesh &m, const Species &sp,
                                   double dt, int H, int rho,
                                   const PtcF64 &p, const FieldsF64 &f,
                                   const AccF64 &a,
                                   unsigned long long *counters, uint64_t n,
                                   hipStream_t st);
hipError_t launch_finalize_f64(long long *hi, long long *lo, double *out,
                               uint64_t n, int H, hipStream_t st);

// ---- mixed form (FP32 push, int32 cell + FP32 offset, int64 J) -----
struct PtcF32 {
  uint32_t *cell;
  float *x1, *x2, *x3;
  float *u1, *u2, *u3;
  float *w;
  uint32_t *flag;
};

struct FieldsF32 {
  const float *E[3];
  const float *B[3];
};

// Tiling of the mixed deposit: particles sorted by tile; tile t owns
// particles [tile_off[t], tile_off[t+1]).  Particles past tile_off[ntiles]
// (appended since the last sort) go through the untiled path.
struct TileInfo {
  int T[3];       // tile extent in cells
  int nt[3];      // tiles per dimension
  int halo;       // cells of LDS halo around a tile
  int threads;    // threads per tile workgroup
  const uint32_t *tile_off;
  uint32_t ntiles;
  uint64_t tiled_end;  // tile_off[ntiles], known on the host
};

// int64 accumulators of the mixed form: J[c] always, R[s] when rho != 0
struct AccMixed {
  long long *J[3];
  long long *R[4];
  int rho;
};

hipError_t launch_push_deposit_mixed(const Mesh &m, const Species &sp,
                                     float dt, int S, const PtcF32 &p,
                                     const FieldsF32 &f, const AccMixed &a,
                                     const TileInfo &tiles,
                                     unsigned long long *counters,
                                     uint64_t n, hipStream_t st);
hipError_t launch_finalize_mixed(long long *acc, float *out, uint64_t n,
                                 int S, hipStream_t st);

// ---- fields (T = double: FP64 form, T = float: mixed form) ----------
template <typename T>
struct FieldSet {
  T *E[3], *B[3];
  const T *Ebg[3], *Bbg[3];
  const T *J[3];
  T *divE, *divB;
};

template <typename T>
hipError_t launch_update_e(const Mesh &m, const FieldSet<T> &f, double dt,
                           hipStream_t st);
template <typename T>
hipError_t launch_update_b(const Mesh &m, const FieldSet<T> &f, double dt,
                           hipStream_t st);
template <typename T>
hipError_t launch_divs(const Mesh &m, const FieldSet<T> &f, hipStream_t st);
template <typename T>
hipError_t launch_guard_copy(const Mesh &m, T *a, hipStream_t st);
template <typename T>
hipError_t launch_filter(const Mesh &m, T *j, T *tmp, int fix3d,
                         hipStream_t st);
template <typename T>
hipError_t launch_fill(T *a, T v, uint64_t n, hipStream_t st);

// ---- loading, sort, diagnostics -------------------------------------
struct LoadParams {
  int ppc;
  double theta[2];
  double drift[2];
  double weight;
  uint64_t seed;
  uint64_t base;  // index of the first new particle
};
hipError_t launch_load_uniform_f64(const Mesh &m, const LoadParams &lp,
                                   const PtcF64 &p,
                                   unsigned long long *counters,
                                   hipStream_t st);
hipError_t launch_load_uniform_f32(const Mesh &m, const LoadParams &lp,
                                   const PtcF32 &p,
                                   unsigned long long *counters,
                                   hipStream_t st);

hipError_t launch_erase_guard(const Mesh &m, uint32_t *cell, uint64_t n,
                              hipStream_t st);
hipError_t launch_energy_histogram_f64(const double *E, const uint32_t *flag,
                                       uint64_t n, uint32_t *hist, int nbins,
                                       double E_max, int flag_bit,
                                       hipStream_t st);

}  // namespace apt
