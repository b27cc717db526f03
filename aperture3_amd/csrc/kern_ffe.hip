// kern_ffe.hip -- 3D log-spherical (ln r, theta, phi) force-free / vacuum
// field solver, ffe_solver_logsph (ffe_solver_logsph.cu):
//   rk_push: dB = -dt curl E (ffe_logsph_rk_update_b, :329-369) and
//            dE = dt curl(B - B0) (ffe_logsph_rk_update_e, :371-437; the
//            solver's J stays zero, :995-1005), integral form with the 2D
//            Grid_LogSph coefficients shared by every phi plane;
//   rk_update: e = c1 en + c2 e + c3 de, de = e; b = c1 bn + c2 b + c3 db
//            (:499-538), SSP-RK3 stages (1,0,1), (3/4,1/4,1/4), (1/3,2/3,2/3);
//   apply_boundary: star (:701-745), absorbing layer (:801-827), axis
//            (:747-799);
//   clean_epar (:540-612) and check_eGTb (:614-699).
// The kernels cover the bulk plus two guard layers (SHIFT_GHOST), one
// thread per cell, x fastest.  T = double is bit-identical to the CPU
// restatement; T = float keeps the coefficients in FP64 and casts them.
#include <type_traits>

#include "apt_common.hpp"
#include "apt_kernels.hpp"
#include "apt_math.hpp"

namespace apt {

constexpr int FFE_SHIFT = 2;

__host__ __device__ inline double ffe_pos(const Mesh &m, int d, int n, int st) {
  return m.lower[d] + m.delta[d] * ((double)(n - m.guard[d]) + (st ? 1.0 : 0.5));
}

// the shifted range [guard - 2, dims - guard + 2) in each dimension
struct FfeRange {
  int lo[3], ext[3];
  uint64_t count;
};
static FfeRange ffe_range(const Mesh &m) {
  FfeRange r;
  r.count = 1;
  for (int d = 0; d < 3; d++) {
    r.lo[d] = m.guard[d] - FFE_SHIFT;
    r.ext[d] = m.dims[d] - 2 * m.guard[d] + 2 * FFE_SHIFT;
    r.count *= (uint64_t)r.ext[d];
  }
  return r;
}
__device__ inline void ffe_ijk(const FfeRange &r, uint64_t t, int &i, int &j,
                               int &k) {
  i = r.lo[0] + (int)(t % r.ext[0]);
  uint64_t q = t / r.ext[0];
  j = r.lo[1] + (int)(q % r.ext[1]);
  k = r.lo[2] + (int)(q / r.ext[1]);
}

#define CF(k, o) ((T)cf.a[k][o])

template <typename T>
__global__ void k_ffe_rk_push(Mesh m, FfeRange r, LsCoef cf, FieldSet<T> f,
                              T *dE0, T *dE1, T *dE2, T *dB0, T *dB1, T *dB2,
                              T dt) {
  uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= r.count) return;
  int i, j, k;
  ffe_ijk(r, t, i, j, k);
  const int d0 = m.dims[0];
  const uint64_t dj = (uint64_t)d0, dk = (uint64_t)d0 * m.dims[1];
  const uint64_t ij = (uint64_t)i + (uint64_t)j * d0;
  const uint64_t ijk = ij + (uint64_t)k * dk;
  const T *e1 = f.E[0], *e2 = f.E[1], *e3 = f.E[2];
  const T *b1 = f.B[0], *b2 = f.B[1], *b3 = f.B[2];
  const T *g1 = f.Bbg[0], *g2 = f.Bbg[1], *g3 = f.Bbg[2];
  // ffe_logsph_rk_update_b
  dB0[ijk] = dt * -(e3[ijk] * CF(LS_L3E, ij) - e3[ijk - dj] * CF(LS_L3E, ij - dj) -
                    (e2[ijk] - e2[ijk - dk]) * CF(LS_L2E, ij)) /
             CF(LS_A1B, ij);
  // A2_b vanishes on the theta = 0, pi node rows: increment 0 there (the
  // reference divides 0 by 0; the axis condition zeroes B2 on those rows)
  dB1[ijk] = cf.a[LS_A2B][ij] == 0.0
                 ? (T)0
                 : dt * -((e1[ijk] - e1[ijk - dk]) * CF(LS_L1E, ij) -
                          e3[ijk] * CF(LS_L3E, ij) + e3[ijk - 1] * CF(LS_L3E, ij - 1)) /
                       CF(LS_A2B, ij);
  dB2[ijk] = dt * -(e2[ijk] * CF(LS_L2E, ij) - e2[ijk - 1] * CF(LS_L2E, ij - 1) -
                    e1[ijk] * CF(LS_L1E, ij) + e1[ijk - dj] * CF(LS_L1E, ij - dj)) /
             CF(LS_A3B, ij);
  // ffe_logsph_rk_update_e, J = 0
  dE0[ijk] = dt *
             ((b3[ijk + dj] - g3[ijk + dj]) * CF(LS_L3B, ij + dj) -
              (b3[ijk] - g3[ijk]) * CF(LS_L3B, ij) -
              (b2[ijk + dk] - g2[ijk + dk]) * CF(LS_L2B, ij) +
              (b2[ijk] - g2[ijk]) * CF(LS_L2B, ij)) /
             CF(LS_A1E, ij);
  dE1[ijk] = dt *
             ((b1[ijk + dk] - g1[ijk + dk] - b1[ijk] + g1[ijk]) * CF(LS_L1B, ij) -
              (b3[ijk + 1] - g3[ijk + 1]) * CF(LS_L3B, ij + 1) +
              (b3[ijk] - g3[ijk]) * CF(LS_L3B, ij)) /
             CF(LS_A2E, ij);
  dE2[ijk] = dt *
             ((b2[ijk + 1] - g2[ijk + 1]) * CF(LS_L2B, ij + 1) -
              (b2[ijk] - g2[ijk]) * CF(LS_L2B, ij) -
              (b1[ijk + dj] - g1[ijk + dj]) * CF(LS_L1B, ij + dj) +
              (b1[ijk] - g1[ijk]) * CF(LS_L1B, ij)) /
             CF(LS_A3E, ij);
}
#undef CF

template <typename T>
__global__ void k_ffe_rk_update(Mesh m, FfeRange r, FieldSet<T> f, FfeBuf<T> b,
                                T c1, T c2, T c3) {
  uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= r.count) return;
  int i, j, k;
  ffe_ijk(r, t, i, j, k);
  const uint64_t ijk = m.idx(i, j, k);
#pragma unroll
  for (int q = 0; q < 3; q++) {
    T e = c1 * b.En[q][ijk] + c2 * f.E[q][ijk] + c3 * b.dE[q][ijk];
    f.E[q][ijk] = e;
    b.dE[q][ijk] = e;
  }
#pragma unroll
  for (int q = 0; q < 3; q++)
    f.B[q][ijk] = c1 * b.Bn[q][ijk] + c2 * f.B[q][ijk] + c3 * b.dB[q][ijk];
}

// star_field_b1 / b2 (ffe_solver_logsph.cu:56-69), B0() = 1000
__device__ inline double ffe_star_b1(double r, double theta) {
  double s, c;
  apt_sincos(theta, s, c);
  return 2.0 * 1000.0 * c / (r * r * r);
}
__device__ inline double ffe_star_b2(double r, double theta) {
  return 1000.0 * apt_sin(theta) / (r * r * r);
}

template <typename T>
__global__ void k_ffe_stellar(Mesh m, FieldSet<T> f, T *G1, T *G2, T *G3,
                              double omega) {
  const int nj = m.dims[1] - m.guard[1] - (m.guard[1] - 1);
  const int nk = m.dims[2] - m.guard[2] - (m.guard[2] - 1);
  const int g0 = m.guard[0];
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nj * nk * g0) return;
  const int i = t % g0;
  const int j = m.guard[1] - 1 + (t / g0) % nj;
  const int k = m.guard[2] - 1 + (t / g0) / nj;
  const double theta_s = ffe_pos(m, 1, j, 1), theta = ffe_pos(m, 1, j, 0);
  const double r_s = apt_exp(ffe_pos(m, 0, i, 1));
  const double r = apt_exp(ffe_pos(m, 0, i, 0));
  const uint64_t o = m.idx(i, j, k);
  const double sb1 = ffe_star_b1(r_s, theta), sb2 = ffe_star_b2(r, theta_s);
  f.B[0][o] = (T)sb1;
  G1[o] = (T)sb1;
  f.E[2][o] = (T)0;
  f.E[1][o] = (T)(-omega * apt_sin(theta) * sb1 * r_s);
  f.E[0][o] = (T)(omega * apt_sin(theta_s) * sb2 * r);
  f.B[1][o] = (T)sb2;
  G2[o] = (T)sb2;
  f.B[2][o] = (T)0;
  G3[o] = (T)0;
}

template <typename T>
__global__ void k_ffe_outflow(Mesh m, int L, double coef, FieldSet<T> f) {
  const int nj = m.dims[1] - m.guard[1] - (m.guard[1] - 1);
  const int nk = m.dims[2] - m.guard[2] - (m.guard[2] - 1);
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nj * nk * L) return;
  const int i = t % L;
  const int j = m.guard[1] - 1 + (t / L) % nj;
  const int k = m.guard[2] - 1 + (t / L) / nj;
  const uint64_t o = m.idx(m.dims[0] - L + i, j, k);
  const double x = (double)i / L;
  const T lambda = (T)(1.0 - coef * (x * x));
  f.E[0][o] *= lambda;
  f.E[1][o] *= lambda;
  f.E[2][o] *= lambda;
  f.B[2][o] *= lambda;
}

template <typename T>
__global__ void k_ffe_axis(Mesh m, FieldSet<T> f, T *G2, int upper) {
  const int ni = m.dims[0] - 2 * m.guard[0];
  const int nk = m.dims[2] - 2 * m.guard[2];
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ni * nk) return;
  const int i = m.guard[0] + t % ni, k = m.guard[2] + t / ni;
  T *E1 = f.E[0], *E2 = f.E[1], *E3 = f.E[2], *B1 = f.B[0], *B2 = f.B[1],
    *B3 = f.B[2];
  if (!upper) {
    const int g = m.guard[1];
    const uint64_t a = m.idx(i, g - 1, k), b = m.idx(i, g, k);
    E1[a] = (T)0;
    E3[a] = (T)0;
    E2[a] = -E2[b];
    B3[a] = (T)0;
    B2[a] = (T)0;
    G2[a] = (T)0;
    B1[a] = B1[b];
  } else {
    const int jl = m.dims[1] - m.guard[1];
    const uint64_t a = m.idx(i, jl - 1, k), b = m.idx(i, jl, k);
    E1[a] = (T)0;
    E3[a] = (T)0;
    E2[b] = -E2[a];
    B3[a] = (T)0;
    G2[a] = (T)0;
    B2[a] = (T)0;
    B1[b] = B1[a];
  }
}

// interpolate(f, idx, in, out, dim0, dim1) (interpolation.cuh:211-237)
template <typename T>
__device__ inline T ffe_interp(const T *__restrict__ a, uint64_t idx, int in,
                               int out, int64_t D0, int64_t D01) {
  const int64_t di_m = ((in & 1) == (out & 1)) ? 0 : -1 - (out & 1);
  const int64_t di_p = ((in & 1) == (out & 1)) ? 0 : -(out & 1);
  const int64_t dj_m = (((in >> 1) & 1) == ((out >> 1) & 1)) ? 0 : -1 - ((out >> 1) & 1);
  const int64_t dj_p = (((in >> 1) & 1) == ((out >> 1) & 1)) ? 0 : -((out >> 1) & 1);
  const int64_t dk_m = (((in >> 2) & 1) == ((out >> 2) & 1)) ? 0 : -1 - ((out >> 2) & 1);
  const int64_t dk_p = (((in >> 2) & 1) == ((out >> 2) & 1)) ? 0 : -((out >> 2) & 1);
  const int64_t x = (int64_t)idx;
  T f11 = (T)0.5 * (a[x + di_p + dj_p * D0 + dk_m * D01] + a[x + di_p + dj_p * D0 + dk_p * D01]);
  T f10 = (T)0.5 * (a[x + di_p + dj_m * D0 + dk_m * D01] + a[x + di_p + dj_m * D0 + dk_p * D01]);
  T f01 = (T)0.5 * (a[x + di_m + dj_p * D0 + dk_m * D01] + a[x + di_m + dj_p * D0 + dk_p * D01]);
  T f00 = (T)0.5 * (a[x + di_m + dj_m * D0 + dk_m * D01] + a[x + di_m + dj_m * D0 + dk_p * D01]);
  T f1 = (T)0.5 * (f11 + f10);
  T f0 = (T)0.5 * (f01 + f00);
  return (T)0.5 * (f1 + f0);
}

constexpr int FFE_ST_E[3] = {0x6, 0x5, 0x3}, FFE_ST_B[3] = {0x1, 0x2, 0x4};

// mode 0: clean_epar, dE := E - (E.B) B / (B^2 + TINY)
// mode 1: check_eGTb, dE := s dE with s from |E|, |B| (E read only)
template <typename T>
__global__ void k_ffe_fix(Mesh m, FfeRange r, FieldSet<T> f, FfeBuf<T> b,
                          int mode) {
  uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= r.count) return;
  int i, j, k;
  ffe_ijk(r, t, i, j, k);
  const uint64_t ijk = m.idx(i, j, k);
  const int64_t D0 = m.dims[0], D01 = (int64_t)m.dims[0] * m.dims[1];
  const T tiny = (T)(double)1e-7f;
  for (int q = 0; q < 3; q++) {
    const int so = FFE_ST_E[q];
    T ex = ffe_interp(f.E[0], ijk, FFE_ST_E[0], so, D0, D01);
    T ey = ffe_interp(f.E[1], ijk, FFE_ST_E[1], so, D0, D01);
    T ez = ffe_interp(f.E[2], ijk, FFE_ST_E[2], so, D0, D01);
    T bx = ffe_interp(f.B[0], ijk, FFE_ST_B[0], so, D0, D01);
    T by = ffe_interp(f.B[1], ijk, FFE_ST_B[1], so, D0, D01);
    T bz = ffe_interp(f.B[2], ijk, FFE_ST_B[2], so, D0, D01);
    if (mode == 0) {
      T bq = q == 0 ? bx : (q == 1 ? by : bz);
      b.dE[q][ijk] = f.E[q][ijk] - (ex * bx + ey * by + ez * bz) * bq /
                                       (bx * bx + by * by + bz * bz + tiny);
    } else {
      T emag = ex * ex + ey * ey + ez * ez + tiny;
      T bmag = bx * bx + by * by + bz * bz + tiny;
      T s = emag > bmag ? sqrt(bmag / emag) : (T)1;
      b.dE[q][ijk] = s * b.dE[q][ijk];
    }
  }
}

template <typename T>
__global__ void k_ffe_copy_e(Mesh m, FfeRange r, FieldSet<T> f, FfeBuf<T> b) {
  uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= r.count) return;
  int i, j, k;
  ffe_ijk(r, t, i, j, k);
  const uint64_t ijk = m.idx(i, j, k);
  for (int q = 0; q < 3; q++) f.E[q][ijk] = b.dE[q][ijk];
}

// ---------------------------------------------------------------------
static inline unsigned ffe_blocks(uint64_t n) {
  return (unsigned)((n + 255) / 256);
}

template <typename T>
hipError_t launch_ffe_boundary(const Mesh &m, int damping_length,
                               double damping_coef, const FieldSet<T> &f,
                               double omega, hipStream_t st) {
  T *G1 = const_cast<T *>(f.Bbg[0]), *G2 = const_cast<T *>(f.Bbg[1]),
    *G3 = const_cast<T *>(f.Bbg[2]);
  const int nj = m.dims[1] - 2 * m.guard[1] + 1, nk = m.dims[2] - 2 * m.guard[2] + 1;
  if (m.guard[0] > 0)
    k_ffe_stellar<T><<<ffe_blocks((uint64_t)nj * nk * m.guard[0]), 256, 0, st>>>(
        m, f, G1, G2, G3, omega);
  if (damping_length > 0)
    k_ffe_outflow<T><<<ffe_blocks((uint64_t)nj * nk * damping_length), 256, 0, st>>>(
        m, damping_length, damping_coef, f);
  const uint64_t na = (uint64_t)(m.dims[0] - 2 * m.guard[0]) * (m.dims[2] - 2 * m.guard[2]);
  k_ffe_axis<T><<<ffe_blocks(na), 256, 0, st>>>(m, f, G2, 0);
  k_ffe_axis<T><<<ffe_blocks(na), 256, 0, st>>>(m, f, G2, 1);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_ffe_update(const Mesh &m, const LsCoef &cf, int damping_length,
                             double damping_coef, const FieldSet<T> &f,
                             const FfeBuf<T> &b, double dt, double omega,
                             double time, hipStream_t st) {
  const FfeRange r = ffe_range(m);
  const size_t bytes = m.size() * sizeof(T);
  hipError_t e;
  for (int q = 0; q < 3; q++) {
    if ((e = hipMemcpyAsync(b.En[q], f.E[q], bytes, hipMemcpyDeviceToDevice, st)) != hipSuccess)
      return e;
    if ((e = hipMemcpyAsync(b.Bn[q], f.B[q], bytes, hipMemcpyDeviceToDevice, st)) != hipSuccess)
      return e;
    if ((e = launch_fill<T>(b.dE[q], (T)0, m.size(), st)) != hipSuccess) return e;
    if ((e = launch_fill<T>(b.dB[q], (T)0, m.size(), st)) != hipSuccess) return e;
  }
  const T cs[3][3] = {{(T)1.0, (T)0.0, (T)1.0},
                      {(T)0.75, (T)0.25, (T)0.25},
                      {(T)(1.0 / 3.0), (T)(2.0 / 3.0), (T)(2.0 / 3.0)}};
  const double ts[3] = {time + dt, time + 0.5 * dt, time + dt};
  for (int s = 0; s < 3; s++) {
    k_ffe_rk_push<T><<<ffe_blocks(r.count), 256, 0, st>>>(
        m, r, cf, f, b.dE[0], b.dE[1], b.dE[2], b.dB[0], b.dB[1], b.dB[2], (T)dt);
    k_ffe_rk_update<T><<<ffe_blocks(r.count), 256, 0, st>>>(m, r, f, b, cs[s][0],
                                                             cs[s][1], cs[s][2]);
    (void)ts[s];
    if ((e = launch_ffe_boundary<T>(m, damping_length, damping_coef, f, omega,
                                    st)) != hipSuccess)
      return e;
    for (int q = 0; q < 3; q++) {
      if ((e = launch_guard_copy<T>(m, f.E[q], st)) != hipSuccess) return e;
      if ((e = launch_guard_copy<T>(m, f.B[q], st)) != hipSuccess) return e;
    }
  }
  return hipGetLastError();
}

// The midpoint interpolation reaches two cells down in each dimension;
// the reference's range (guard - 2) reads before the array when guard = 3,
// so clean_epar / check_eGTb start at max(guard - 2, 2).
static FfeRange ffe_fix_range(const Mesh &m) {
  FfeRange r;
  r.count = 1;
  for (int d = 0; d < 3; d++) {
    int lo = m.guard[d] - FFE_SHIFT;
    if (lo < 2) lo = 2;
    r.lo[d] = lo;
    r.ext[d] = m.dims[d] - m.guard[d] + FFE_SHIFT - lo;
    r.count *= (uint64_t)r.ext[d];
  }
  return r;
}

template <typename T>
hipError_t launch_ffe_clean(const Mesh &m, const FieldSet<T> &f,
                            const FfeBuf<T> &b, int mode, hipStream_t st) {
  const FfeRange r = ffe_fix_range(m);
  k_ffe_fix<T><<<ffe_blocks(r.count), 256, 0, st>>>(m, r, f, b, mode);
  if (mode == 1)
    k_ffe_copy_e<T><<<ffe_blocks(r.count), 256, 0, st>>>(m, r, f, b);
  return hipGetLastError();
}

#define INST(T)                                                                \
  template hipError_t launch_ffe_update<T>(const Mesh &, const LsCoef &, int,  \
                                           double, const FieldSet<T> &,        \
                                           const FfeBuf<T> &, double, double,  \
                                           double, hipStream_t);               \
  template hipError_t launch_ffe_boundary<T>(const Mesh &, int, double,        \
                                             const FieldSet<T> &, double,      \
                                             hipStream_t);                     \
  template hipError_t launch_ffe_clean<T>(const Mesh &, const FieldSet<T> &,   \
                                          const FfeBuf<T> &, int, hipStream_t);
INST(double)
INST(float)
#undef INST

}  // namespace apt
