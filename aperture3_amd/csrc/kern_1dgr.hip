// kern_1dgr.hip -- the 1D general-relativistic gap along one Kerr field
// line (the labs/1dgr problem): the particle and photon pushers of
// ptc_updater_1dgr (src/cuda/algorithms/ptc_updater_1dgr.cu:136-406), its
// K-weighted current filter (:408-427) and the E update of
// field_solver_1dgr (src/cuda/algorithms/field_solver_1dgr.cu:14-38).
//
// FP64, one thread per particle (the 1D problem has at most a few million
// particles and is latency bound).  exp/sin/cos are the build's own
// (apt_math.hpp) so the CPU restatement agrees bit for bit; the current and
// charge are summed in the two-limb int64 fixed point of the Cartesian FP64
// form instead of the reference's double atomicAdd.
#include "apt_common.hpp"
#include "apt_f64.hpp"
#include "apt_kernels.hpp"
#include "apt_math.hpp"

namespace apt {

namespace {

__device__ inline double sq(double x) { return x * x; }

__device__ inline double lerp_c(const double *v, int c, double x1, double nx1) {
  return v[c] * x1 + v[c - 1] * nx1;
}

__device__ inline double cic(double dx) {
  const double a = 1.0 - fabs(dx);
  return a > 0.0 ? a : 0.0;
}

__device__ inline bool bulk1(const Mesh &m, uint32_t c) {
  return (int)c >= m.guard[0] && (int)c < m.dims[0] - m.guard[0];
}

// push and move of one live particle (the body of update_ptc_1dgr before
// its deposit); returns what the deposit needs
__device__ __forceinline__ void gr1d_push_one(const Mesh &m, const Gr1dCoef &cf,
                                              const Species &sp, const PtcF64 &p,
                                              const double *__restrict__ e3, double dt,
                                              uint32_t step, int data_interval, double a,
                                              double omega, uint64_t idx, uint32_t c32,
                                              bool &dep, int &c_out, int &dc_out, int &s_out,
                                              double &x1_out, double &nx1_out,
                                              double &weight_out) {
  const int c = (int)c32;
  const double rp = 1.0 + sqrt(1.0 - a * a);
  const double rm = 1.0 - sqrt(1.0 - a * a);
  const double *cD1 = cf.a[GR_D1], *cD2 = cf.a[GR_D2], *cD3 = cf.a[GR_D3];
  const double *cal = cf.a[GR_ALPHA];
  double p1 = p.p1[idx];
  const double x1 = p.x1[idx];
  double u0 = p.E[idx];
  double nx1 = 1.0 - x1;
  const uint32_t flag = p.flag[idx];
  const int s = ptc_species(flag);
  const double xi = m.lower[0] + m.delta[0] * ((double)(c - m.guard[0]) + x1);
  const double exp_xi = apt_exp(xi * (rp - rm));
  const double r = (rp - rm * exp_xi) / (1.0 - exp_xi);

  double alpha = lerp_c(cal, c, x1, nx1);
  double D1 = lerp_c(cD1, c, x1, nx1);
  double D2 = lerp_c(cD2, c, x1, nx1);
  double D3 = lerp_c(cD3, c, x1, nx1);
  const double Dxi = x1 < 0.5 ? e3[c] * (0.5 + x1) + e3[c - 1] * (0.5 - x1)
                              : e3[c] * (1.5 - x1) + e3[c + 1] * (x1 - 0.5);
  const double agrr = lerp_c(cf.a[GR_AGRR], c, x1, nx1);
  const double Er = agrr * Dxi;

  double vr = (p1 / u0 - D1) / D2;
  const double id0 = m.inv_delta[0];
  const double da2dr = (sq(cal[c]) - sq(cal[c - 1])) * id0;
  const double dD1dr = (cD1[c] - cD1[c - 1]) * id0;
  const double dD2dr = (cD2[c] - cD2[c - 1]) * id0;
  const double dD3dr = (cD3[c] - cD3[c - 1]) * id0;
  const double dDelta_dr = 2.0 * r - 2.0;

  const double gr_term = -0.5 * u0 *
                             (da2dr - (dD3dr + 2.0 * (dD1dr - D1 * dDelta_dr) * vr +
                                       (dD2dr - 2.0 * D2 * dDelta_dr) * vr * vr)) *
                             dt +
                         dDelta_dr * p1 * vr * dt;
  p1 += gr_term;
  const double E_term = Er * sp.q[s] / sp.m[s] * dt;
  p1 += E_term;
  u0 = sqrt((D2 + p1 * p1) / (D2 * (alpha * alpha - D3) + D1 * D1));
  p.p1[idx] = p1;
  vr = (p1 / u0 - D1) / D2;

  // u_0 at infinity of tracked particles on output steps, kept in p2
  if (data_interval > 0 && step % (uint32_t)data_interval == 0 && (flag & FLAG_TRACKED)) {
    const double theta = lerp_c(cf.a[GR_THETA], c, x1, nx1);
    double st, ct;
    apt_sincos(theta, st, ct);
    const double Sigma = r * r + a * a * sq(ct);
    const double Delta = r * r - 2.0 * r + a * a;
    const double g_00 = -(1.0 - 2.0 * r / Sigma);
    const double g_03 = -4.0 * r * a * sq(st) / Sigma;
    const double B31 = lerp_c(cf.a[GR_B3B1], c, x1, nx1);
    p.p2[idx] = (g_00 + g_03 * omega) * u0 + g_03 * B31 * Delta * u0 * vr;
  }

  double new_x1 = x1 + vr * dt * id0;
  const int dc1 = (int)floor(new_x1);
  new_x1 -= (double)dc1;
  p.cell[idx] = (uint32_t)(c + dc1);
  p.x1[idx] = new_x1;

  nx1 = 1.0 - new_x1;
  alpha = lerp_c(cal, c + dc1, new_x1, nx1);
  D1 = lerp_c(cD1, c + dc1, new_x1, nx1);
  D2 = lerp_c(cD2, c + dc1, new_x1, nx1);
  D3 = lerp_c(cD3, c + dc1, new_x1, nx1);
  u0 = sqrt((D2 + p1 * p1) / (D2 * (alpha * alpha - D3) + D1 * D1));
  p.E[idx] = u0;

  c_out = c;
  dc_out = dc1;
  s_out = s;
  x1_out = x1;
  nx1_out = new_x1;
  dep = !(flag & FLAG_IGNORE_CURRENT);
  if (dep) weight_out = -sp.q[s] * p.w[idx];
}

// update_ptc_1dgr (ptc_updater_1dgr.cu:136-299)
__global__ void k_1dgr_push(Mesh m, Gr1dCoef cf, Species sp, PtcF64 p,
                            const double *__restrict__ e3, AccF64 acc, double dt,
                            uint32_t step, int data_interval, double a, double omega,
                            int H, uint64_t n) {
  const uint64_t idx = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  // every lane of the wave reaches the deposit (wave_acc_add); lanes without
  // a live particle take part with nothing to add
  bool live = idx < n;
  const uint32_t c32 = live ? p.cell[idx] : MAX_CELL;
  live = live && c32 != MAX_CELL;
  if (live && !bulk1(m, c32)) {
    p.cell[idx] = MAX_CELL;
    live = false;
  }
  bool dep = false;
  int c = 0, dc1 = 0, s = 0;
  double x1 = 0.0, new_x1 = 0.0, weight = 0.0;
  if (live) gr1d_push_one(m, cf, sp, p, e3, dt, step, data_interval, a, omega, idx, c32, dep,
                          c, dc1, s, x1, new_x1, weight);
  const double *K1j = cf.a[GR_K1J], *K1 = cf.a[GR_K1];
  const int i_0 = dc1 == -1 ? -2 : -1;
  const int i_1 = dc1 == 1 ? 1 : 0;
  double djx = 0.0;
  for (int i = -2; i <= 1; i++) {
    const bool in = dep && i >= i_0 && i <= i_1;
    double vj = 0.0, vr = 0.0;
    uint64_t o1 = 0, o = 0;
    if (in) {
      const double sx0 = cic(-x1 + i + 1);
      const double sx1 = cic(-new_x1 + (i + 1 - dc1));
      djx += sx1 - sx0;
      o1 = (uint64_t)(i + c + 1);
      o = (uint64_t)(i + c);
      vj = weight * djx / K1j[i + c + 1] * m.delta[0] / dt;
      vr = -weight * sx1 / K1[i + c];
    }
    wave_acc_add(acc.J[0][0], acc.J[0][1], o1, o1, vj, H, in);
    wave_acc_add(acc.R[s][0], acc.R[s][1], o, o | ((uint64_t)s << 56), vr, H, in);
  }
}



// update_photon_1dgr (ptc_updater_1dgr.cu:301-406), with the reference's
// fixed a = 0.99
__global__ void k_1dgr_photons(Mesh m, Gr1dCoef cf, PhoF64 ph, double dt, uint32_t step,
                               int data_interval, uint64_t n) {
  const uint64_t idx = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= n) return;
  const uint32_t c32 = ph.cell[idx];
  if (c32 == MAX_CELL) return;
  if (!bulk1(m, c32)) {
    ph.cell[idx] = MAX_CELL;
    return;
  }
  const int c = (int)c32;
  const double a = 0.99;
  const double rp = 1.0 + sqrt(1.0 - a * a);
  const double rm = 1.0 - sqrt(1.0 - a * a);
  const double *cal = cf.a[GR_ALPHA], *cgrr = cf.a[GR_GRR], *cgff = cf.a[GR_GFF];
  const double *cbeta = cf.a[GR_BETA];
  double p1 = ph.p1[idx];
  const double p3 = ph.p3[idx], x1 = ph.x1[idx];
  const double nx1 = 1.0 - x1;
  const double xi = m.lower[0] + m.delta[0] * ((double)(c - m.guard[0]) + x1);
  const double exp_xi = apt_exp(xi * (rp - rm));
  const double r = (rp - rm * exp_xi) / (1.0 - exp_xi);
  const double alpha = lerp_c(cal, c, x1, nx1);
  const double gamma11 = lerp_c(cgrr, c, x1, nx1);
  const double gamma33 = lerp_c(cgff, c, x1, nx1);
  const double dDelta_dr = 2.0 * r - 2.0;
  const double Delta_sqr = sq(r * r - 2.0 * r + a * a);
  double u0 = sqrt(gamma11 * p1 * p1 / Delta_sqr + gamma33 * p3 * p3) / alpha;
  const double id0 = m.inv_delta[0];
  const double da2dr = (sq(cal[c]) - sq(cal[c - 1])) * id0;
  const double dg11 = (cgrr[c] - cgrr[c - 1]) * id0;
  const double dg33 = (cgff[c] - cgff[c - 1]) * id0;
  const double dbeta = (cbeta[c] - cbeta[c - 1]) * id0;
  const double gr_term = (-0.5 * u0 * da2dr -
                          (p1 * p1 * dg11 / Delta_sqr + p3 * p3 * dg33) * 0.5 / u0 +
                          p3 * dbeta + gamma11 * p1 * p1 * dDelta_dr / (Delta_sqr * u0)) *
                         dt;
  p1 += gr_term;
  u0 = sqrt(gamma11 * p1 * p1 / Delta_sqr + gamma33 * p3 * p3) / alpha;
  ph.p1[idx] = p1;
  ph.E[idx] = u0;
  const double vr = gamma11 * p1 / (Delta_sqr * u0);
  if (data_interval > 0 && step % (uint32_t)data_interval == 0 &&
      (ph.flag[idx] & FLAG_TRACKED)) {
    const double theta = lerp_c(cf.a[GR_THETA], c, x1, nx1);
    double st, ct;
    apt_sincos(theta, st, ct);
    const double Sigma = r * r + a * a * sq(ct);
    const double Delta = r * r - 2.0 * r + a * a;
    const double g_00 = -(1.0 + 2.0 * r * (r * r + a * a) / Sigma / Delta);
    const double g_03 = -2.0 * r * a / Sigma / Delta;
    ph.p2[idx] = (u0 - g_03 * ph.p3[idx]) / g_00;
  }
  double new_x1 = x1 + vr * dt * id0;
  const int dc1 = (int)floor(new_x1);
  new_x1 -= (double)dc1;
  ph.cell[idx] = (uint32_t)(c + dc1);
  ph.x1[idx] = new_x1;
}

// filter_current1d (ptc_updater_1dgr.cu:408-427); the first and last array
// cells are copied (the reference reads one past each end there)
__global__ void k_1dgr_filter(Mesh m, const double *__restrict__ j, double *__restrict__ tmp,
                              const double *__restrict__ k1) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int n = m.dims[0];
  if (i >= n) return;
  if (i == 0 || i == n - 1) {
    tmp[i] = j[i];
    return;
  }
  const int dp = (i == n - m.guard[0] - 1) ? 0 : 1;
  const int dm = (i == m.guard[0]) ? 0 : 1;
  double v = 0.5 * j[i] * k1[i];
  v += 0.25 * j[i + dp] * k1[i + dp];
  v += 0.25 * j[i - dm] * k1[i - dm];
  v /= k1[i];
  tmp[i] = v;
}

// update_e_1dgr (field_solver_1dgr.cu:14-38)
__global__ void k_1dgr_update_e(Mesh m, const double *__restrict__ j0, double *e1,
                                const double *__restrict__ j1, double dt) {
  const int i = m.guard[0] + 1 + blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m.dims[0] - m.guard[0] - 5) return;
  e1[i] += (j0[i] - j1[i]) * dt;
}

inline unsigned blocks(uint64_t n) { return (unsigned)((n + 255) / 256); }

}  // namespace

hipError_t launch_1dgr_push(const Mesh &m, const Gr1dCoef &cf, const Species &sp,
                            const PtcF64 &p, const double *e3, const AccF64 &acc,
                            double dt, uint32_t step, int data_interval, double a,
                            double omega, int H, uint64_t n, hipStream_t st) {
  if (n == 0) return hipSuccess;
  k_1dgr_push<<<blocks(n), 256, 0, st>>>(m, cf, sp, p, e3, acc, dt, step, data_interval,
                                         a, omega, H, n);
  return hipGetLastError();
}

hipError_t launch_1dgr_photons(const Mesh &m, const Gr1dCoef &cf, const PhoF64 &ph,
                               double dt, uint32_t step, int data_interval, uint64_t n,
                               hipStream_t st) {
  if (n == 0) return hipSuccess;
  k_1dgr_photons<<<blocks(n), 256, 0, st>>>(m, cf, ph, dt, step, data_interval, n);
  return hipGetLastError();
}

hipError_t launch_1dgr_filter(const Mesh &m, const double *j, double *tmp, const double *k1,
                              hipStream_t st) {
  k_1dgr_filter<<<blocks((uint64_t)m.dims[0]), 256, 0, st>>>(m, j, tmp, k1);
  return hipGetLastError();
}

hipError_t launch_1dgr_update_e(const Mesh &m, const Gr1dCoef &cf, double *e1,
                                const double *j1, double dt, hipStream_t st) {
  const int n = m.dims[0] - 2 * m.guard[0] - 6;
  if (n <= 0) return hipSuccess;
  k_1dgr_update_e<<<blocks((uint64_t)n), 256, 0, st>>>(m, cf.a[GR_J0], e1, j1, dt);
  return hipGetLastError();
}

}  // namespace apt
