/*
 * gr1d_oracle.c -- CPU restatement of the 1D general-relativistic gap
 * (Kerr field line): Grid_1dGR::compute_coef (src/grids/grid_1dgr.cpp:
 * 21-190), ptc_updater_1dgr (src/cuda/algorithms/ptc_updater_1dgr.cu:
 * 16-530) and field_solver_1dgr (src/cuda/algorithms/field_solver_1dgr.cu:
 * 14-58).  TEST INFRASTRUCTURE ONLY: the checker for the HIP path, never
 * linked into the product.
 *
 * FP64 throughout (the reference's USE_DOUBLE build), with the reference's
 * float inputs kept as floats: the profile arrays of coef.h5 are read as
 * float (grid_1dgr.cpp:60-66) and the parameters a, omega and B0 are float
 * (sim_params.h:58-63).  exp/sin/cos are the build's own (orc_math.c) so
 * the GPU reproduces them bit for bit; the current is summed in the same
 * two-limb fixed point as the Cartesian oracle.
 */
#include <math.h>
#include <string.h>

#include "orc_internal.h"

#define GR_D1 0
#define GR_D2 1
#define GR_D3 2
#define GR_ALPHA 3
#define GR_THETA 4
#define GR_K1 5
#define GR_K1J 6
#define GR_J0 7
#define GR_AGRR 8
#define GR_AGRF 9
#define GR_RHO0 10
#define GR_GRR 11
#define GR_GFF 12
#define GR_BETA 13
#define GR_B3B1 14

static double sq(double x) { return x * x; }

/* r(xi) of the tortoise-like coordinate: r = (r+ - r- e^{xi (r+ - r-)}) /
 * (1 - e^{xi (r+ - r-)}) (grid_1dgr.cpp:89-93) */
static double r_of_xi(double xi, double a) {
  double rp = 1.0 + sqrt(1.0 - a * a);
  double rm = 1.0 - sqrt(1.0 - a * a);
  double e = orc_exp(xi * (rp - rm));
  return (rp - rm * e) / (1.0 - e);
}

/* Grid_1dGR::compute_coef (grid_1dgr.cpp:21-175).  Profile arrays are the
 * coef.h5 datasets r, Bu1, Bu3, rho, ju1, th, dpsidth as float; coef[15]
 * each dims[0] doubles, zero-filled first (theta and agrf are never set by
 * the reference and stay 0).  Returns 0, or -1 when a node's r lies past
 * the profile. */
int orc_gr1d_coef(const orc_mesh *m, const float *vr, const float *vB1, const float *vB3,
                  const float *vrho, const float *vjr, const float *vth,
                  const float *vdpsi, uint64_t nprof, float a_f, float omega_f,
                  float B0_f, double *const *coef) {
  const double a = a_f, omega = omega_f, B0 = B0_f;
  const int n_size = m->dims[0];
  for (int k = 0; k < 15; k++) memset(coef[k], 0, sizeof(double) * (size_t)n_size);
  uint64_t n_data = 0, n_mid = 0;
  for (int n = 0; n < n_size; n++) {
    double xi = m->lower[0] + (double)(n - m->guard[0] + 1) * m->delta[0];
    double r = r_of_xi(xi, a);
    while (n_data + 1 < nprof && vr[n_data + 1] < r) n_data += 1;
    if (n_data + 1 >= nprof) return -1;
    double x = (r - vr[n_data]) / (vr[n_data + 1] - vr[n_data]);
    double B1 = vB1[n_data] * (1.0 - x) + vB1[n_data + 1] * x;
    double B3 = vB3[n_data] * (1.0 - x) + vB3[n_data + 1] * x;
    double theta = vth[n_data] * (1.0 - x) + vth[n_data + 1] * x;
    double b31 = B3 / B1;
    coef[GR_B3B1][n] = b31;
    coef[GR_RHO0][n] = B0 * (vrho[n_data] * (1.0 - x) + vrho[n_data + 1] * x);
    double st, ct;
    orc_sincos(theta, &st, &ct);
    double Sigma = r * r + a * a * sq(ct);
    double Delta = r * r - 2.0 * r + a * a;
    double A = sq(r * r + a * a) - Delta * a * a * sq(st);
    double w = 2.0 * a * r / A;
    double g11 = Sigma / Delta;
    double g22 = Sigma;
    double g33 = A * sq(st) / Sigma;
    double sqrt_gamma = sqrt(g11 * g22 * g33);
    coef[GR_K1][n] = Delta * sqrt_gamma /
                     (vdpsi[n_data] * (1.0 - x) + vdpsi[n_data + 1] * x);
    coef[GR_D1][n] = (b31 * g33 * (omega - w)) * Delta;
    coef[GR_D2][n] = (g11 + b31 * b31 * g33) * Delta * Delta;
    coef[GR_D3][n] = g33 * sq(omega - w);
    coef[GR_ALPHA][n] = sqrt(Delta * Sigma / A);
    coef[GR_GRR][n] = 1.0 / g11;
    coef[GR_GFF][n] = 1.0 / g33;
    coef[GR_BETA][n] = -2.0 * a * r / A;

    double xi_mid = m->lower[0] + ((double)(n - m->guard[0]) + 0.5) * m->delta[0];
    double r_mid = r_of_xi(xi_mid, a);
    while (n_mid + 1 < nprof && vr[n_mid + 1] < r_mid) n_mid += 1;
    if (n_mid + 1 >= nprof) return -1;
    /* the reference interpolates the mid-point quantities with the node
     * r, not r_mid (grid_1dgr.cpp:146) */
    double y = (r - vr[n_mid]) / (vr[n_mid + 1] - vr[n_mid]);
    coef[GR_J0][n] = B0 * (vjr[n_mid] * (1.0 - y) + vjr[n_mid + 1] * y) / Delta;
    double theta_mid = vth[n_mid] * (1.0 - y) + vth[n_mid + 1] * y;
    orc_sincos(theta_mid, &st, &ct);
    Sigma = r_mid * r_mid + a * a * sq(ct);
    Delta = r_mid * r_mid - 2.0 * r_mid + a * a;
    A = sq(r_mid * r_mid + a * a) - Delta * a * a * sq(st);
    g11 = Sigma / Delta;
    g22 = Sigma;
    g33 = A * sq(st) / Sigma;
    sqrt_gamma = sqrt(g11 * g22 * g33);
    coef[GR_K1J][n] = Delta * sqrt_gamma /
                      (vdpsi[n_mid] * (1.0 - y) + vdpsi[n_mid + 1] * y);
    coef[GR_AGRR][n] = Delta * Delta * sqrt(Delta * Sigma / A) * g11;
  }
  return 0;
}

/* prepare_initial_condition (ptc_updater_1dgr.cu:16-72): per bulk cell,
 * `mult` e-/e+ pairs at x1 = 1, p1 = 0, u0 from the metric, weights B0/mult
 * plus the part of rho0 K1 of each sign; slot = cell * 2 mult + 2 n (+1),
 * so the first guard * 2 mult slots stay empty and the pairs of the last
 * guard cells fall past number = N * 2 mult (kept as in the reference).
 * Tracked flag with probability track_percent from the Philox stream
 * (purpose 8); every particle gets an id = its slot.  Returns number. */
uint64_t orc_gr1d_initial(const orc_mesh *m, double *const *coef, orc_ptc *p, int mult,
                          float B0_f, float track_f, uint64_t seed, uint64_t cap) {
  const double B0 = B0_f, track = track_f;
  const uint64_t number = (uint64_t)(m->dims[0] - 2 * m->guard[0]) * 2 * (uint64_t)mult;
  for (uint64_t i = 0; i < number && i < cap; i++) {
    p->cell[i] = ORC_MAX_CELL;
    p->id[i] = i;
  }
  for (int cell = m->guard[0]; cell < m->dims[0] - m->guard[0]; cell++)
    for (int n = 0; n < mult; n++) {
      uint64_t idx = (uint64_t)cell * (uint64_t)mult * 2 + (uint64_t)n * 2;
      if (idx + 1 >= number) continue;
      double D1 = coef[GR_D1][cell], D2 = coef[GR_D2][cell], D3 = coef[GR_D3][cell];
      double alpha = coef[GR_ALPHA][cell], K1 = coef[GR_K1][cell];
      double rho0 = coef[GR_RHO0][cell];
      double u0 = sqrt(D2 / (D2 * (alpha * alpha - D3) + D1 * D1));
      double rk = rho0 * K1;
      double u[2];
      orc_uniform2(seed, idx, 0, 8, u);
      for (int s = 0; s < 2; s++) {
        uint64_t k = idx + (uint64_t)s;
        p->x1[k] = 1.0;
        p->x2[k] = 0.0;
        p->x3[k] = 0.0;
        p->p1[k] = 0.0;
        p->p2[k] = 0.0;
        p->p3[k] = 0.0;
        p->E[k] = u0;
        p->cell[k] = (uint32_t)cell;
        p->w[k] = s == 0 ? 1.0 * B0 / mult + fabs(fmin(rk, 0.0)) / mult
                         : 1.0 * B0 / mult + fmax(rk, 0.0) / mult;
        p->flag[k] = ((uint32_t)s << 29) | (u[s] < track ? 2u : 0u);
        p->id[k] = k;
      }
    }
  p->number = number;
  return number;
}

/* interpolation of a node array at x1 in cell c: c-1 .. c */
static double lerp_c(const double *v, int c, double x1, double nx1) {
  return v[c] * x1 + v[c - 1] * nx1;
}

/* update_ptc_1dgr (ptc_updater_1dgr.cu:136-299): Er from the filtered
 * E3, the GR momentum update of p1 along the field line, the move, and the
 * current J1 / charge Rho[sp] deposit divided by K1_j / K1.  p2 keeps
 * u_0 at infinity for tracked particles on output steps.  e3 is the
 * filtered E field the reference keeps in E.data(2). */
void orc_gr1d_push(const orc_mesh *m, double *const *coef, orc_ptc *p,
                   const orc_species *sp, const double *e3, orc_fields *f, double dt,
                   uint32_t step, int data_interval, float a_f, float omega_f) {
  const double a = a_f, omega = omega_f;
  const int H = sp->jexp;
  const double rp = 1.0 + sqrt(1.0 - a * a);
  const double rm = 1.0 - sqrt(1.0 - a * a);
  const double *cD1 = coef[GR_D1], *cD2 = coef[GR_D2], *cD3 = coef[GR_D3];
  const double *cal = coef[GR_ALPHA], *cth = coef[GR_THETA], *cK1 = coef[GR_K1];
  const double *cK1j = coef[GR_K1J], *cagrr = coef[GR_AGRR], *cB31 = coef[GR_B3B1];
  for (uint64_t idx = 0; idx < p->number; idx++) {
    uint32_t c32 = p->cell[idx];
    if (c32 == ORC_MAX_CELL) continue;
    if (!is_in_bulk(m, c32)) {
      p->cell[idx] = ORC_MAX_CELL;
      continue;
    }
    const int c = (int)c32;
    double p1 = p->p1[idx];
    double x1 = p->x1[idx];
    double u0 = p->E[idx];
    double nx1 = 1.0 - x1;
    uint32_t flag = p->flag[idx];
    int s = (int)(flag >> 29);
    double xi = m->lower[0] + m->delta[0] * ((double)(c - m->guard[0]) + x1);
    double exp_xi = orc_exp(xi * (rp - rm));
    double r = (rp - rm * exp_xi) / (1.0 - exp_xi);

    double alpha = lerp_c(cal, c, x1, nx1);
    double D1 = lerp_c(cD1, c, x1, nx1);
    double D2 = lerp_c(cD2, c, x1, nx1);
    double D3 = lerp_c(cD3, c, x1, nx1);
    double Dxi = x1 < 0.5 ? e3[c] * (0.5 + x1) + e3[c - 1] * (0.5 - x1)
                          : e3[c] * (1.5 - x1) + e3[c + 1] * (x1 - 0.5);
    double agrr = lerp_c(cagrr, c, x1, nx1);
    double Er = agrr * Dxi;

    double vr = (p1 / u0 - D1) / D2;
    double da2dr = (sq(cal[c]) - sq(cal[c - 1])) * m->inv_delta[0];
    double dD1dr = (cD1[c] - cD1[c - 1]) * m->inv_delta[0];
    double dD2dr = (cD2[c] - cD2[c - 1]) * m->inv_delta[0];
    double dD3dr = (cD3[c] - cD3[c - 1]) * m->inv_delta[0];
    double dDelta_dr = 2.0 * r - 2.0;

    double gr_term = -0.5 * u0 *
                         (da2dr - (dD3dr + 2.0 * (dD1dr - D1 * dDelta_dr) * vr +
                                   (dD2dr - 2.0 * D2 * dDelta_dr) * vr * vr)) *
                         dt +
                     dDelta_dr * p1 * vr * dt;
    p1 += gr_term;
    double E_term = Er * sp->q[s] / sp->m[s] * dt;
    p1 += E_term;
    u0 = sqrt((D2 + p1 * p1) / (D2 * (alpha * alpha - D3) + D1 * D1));
    p->p1[idx] = p1;
    vr = (p1 / u0 - D1) / D2;

    if (data_interval > 0 && step % (uint32_t)data_interval == 0 && (flag & 2u)) {
      double theta = lerp_c(cth, c, x1, nx1);
      double st, ct;
      orc_sincos(theta, &st, &ct);
      double Sigma = r * r + a * a * sq(ct);
      double Delta = r * r - 2.0 * r + a * a;
      double g_00 = -(1.0 - 2.0 * r / Sigma);
      double g_03 = -4.0 * r * a * sq(st) / Sigma;
      double B31 = lerp_c(cB31, c, x1, nx1);
      p->p2[idx] = (g_00 + g_03 * omega) * u0 + g_03 * B31 * Delta * u0 * vr;
    }

    double new_x1 = x1 + vr * dt * m->inv_delta[0];
    int dc1 = (int)floor(new_x1);
    new_x1 -= (double)dc1;
    p->cell[idx] = (uint32_t)(c + dc1);
    p->x1[idx] = new_x1;

    nx1 = 1.0 - new_x1;
    alpha = lerp_c(cal, c + dc1, new_x1, nx1);
    D1 = lerp_c(cD1, c + dc1, new_x1, nx1);
    D2 = lerp_c(cD2, c + dc1, new_x1, nx1);
    D3 = lerp_c(cD3, c + dc1, new_x1, nx1);
    u0 = sqrt((D2 + p1 * p1) / (D2 * (alpha * alpha - D3) + D1 * D1));
    p->E[idx] = u0;

    if (!(flag & FLAG_IGNORE_CURRENT)) {
      double weight = -sp->q[s] * p->w[idx];
      int i_0 = dc1 == -1 ? -2 : -1;
      int i_1 = dc1 == 1 ? 1 : 0;
      double djx = 0.0;
      for (int i = i_0; i <= i_1; i++) {
        double sx0 = cic(-x1 + i + 1);
        double sx1 = cic(-new_x1 + (i + 1 - dc1));
        djx += sx1 - sx0;
        int64_t h, l;
        orc_fx_split(weight * djx / cK1j[i + c + 1] * m->delta[0] / dt, H, &h, &l);
        f->Jacc[0][0][i + c + 1] += h;
        f->Jacc[0][1][i + c + 1] += l;
        orc_fx_split(-weight * sx1 / cK1[i + c], H, &h, &l);
        f->Racc[s][0][i + c] += h;
        f->Racc[s][1][i + c] += l;
      }
    }
  }
}

/* update_photon_1dgr (ptc_updater_1dgr.cu:301-406): the photon's p1 along
 * the field line with the reference's fixed a = 0.99; p2 keeps u_0 at
 * infinity for tracked photons on output steps. */
void orc_gr1d_photons(const orc_mesh *m, double *const *coef, orc_photons *ph, double dt,
                      uint32_t step, int data_interval) {
  const double a = 0.99;
  const double rp = 1.0 + sqrt(1.0 - a * a);
  const double rm = 1.0 - sqrt(1.0 - a * a);
  const double *cal = coef[GR_ALPHA], *cgrr = coef[GR_GRR], *cgff = coef[GR_GFF];
  const double *cbeta = coef[GR_BETA], *cth = coef[GR_THETA];
  for (uint64_t idx = 0; idx < ph->number; idx++) {
    uint32_t c32 = ph->cell[idx];
    if (c32 == ORC_MAX_CELL) continue;
    if (!is_in_bulk(m, c32)) {
      ph->cell[idx] = ORC_MAX_CELL;
      continue;
    }
    const int c = (int)c32;
    double p1 = ph->p1[idx], p3 = ph->p3[idx], x1 = ph->x1[idx];
    double u0;
    double nx1 = 1.0 - x1;
    double xi = m->lower[0] + m->delta[0] * ((double)(c - m->guard[0]) + x1);
    double exp_xi = orc_exp(xi * (rp - rm));
    double r = (rp - rm * exp_xi) / (1.0 - exp_xi);
    double alpha = lerp_c(cal, c, x1, nx1);
    double gamma11 = lerp_c(cgrr, c, x1, nx1);
    double gamma33 = lerp_c(cgff, c, x1, nx1);
    double dDelta_dr = 2.0 * r - 2.0;
    double Delta_sqr = sq(r * r - 2.0 * r + a * a);
    u0 = sqrt(gamma11 * p1 * p1 / Delta_sqr + gamma33 * p3 * p3) / alpha;
    double da2dr = (sq(cal[c]) - sq(cal[c - 1])) * m->inv_delta[0];
    double dg11 = (cgrr[c] - cgrr[c - 1]) * m->inv_delta[0];
    double dg33 = (cgff[c] - cgff[c - 1]) * m->inv_delta[0];
    double dbeta = (cbeta[c] - cbeta[c - 1]) * m->inv_delta[0];
    double gr_term = (-0.5 * u0 * da2dr -
                      (p1 * p1 * dg11 / Delta_sqr + p3 * p3 * dg33) * 0.5 / u0 +
                      p3 * dbeta + gamma11 * p1 * p1 * dDelta_dr / (Delta_sqr * u0)) *
                     dt;
    p1 += gr_term;
    u0 = sqrt(gamma11 * p1 * p1 / Delta_sqr + gamma33 * p3 * p3) / alpha;
    ph->p1[idx] = p1;
    ph->E[idx] = u0;
    double vr = gamma11 * p1 / (Delta_sqr * u0);
    if (data_interval > 0 && step % (uint32_t)data_interval == 0 && (ph->flag[idx] & 2u)) {
      double theta = lerp_c(cth, c, x1, nx1);
      double st, ct;
      orc_sincos(theta, &st, &ct);
      double Sigma = r * r + a * a * sq(ct);
      double Delta = r * r - 2.0 * r + a * a;
      double g_00 = -(1.0 + 2.0 * r * (r * r + a * a) / Sigma / Delta);
      double g_03 = -2.0 * r * a / Sigma / Delta;
      ph->p2[idx] = (u0 - g_03 * ph->p3[idx]) / g_00;
    }
    double new_x1 = x1 + vr * dt * m->inv_delta[0];
    int dc1 = (int)floor(new_x1);
    new_x1 -= (double)dc1;
    ph->cell[idx] = (uint32_t)(c + dc1);
    ph->x1[idx] = new_x1;
  }
}

/* filter_current1d (ptc_updater_1dgr.cu:408-427): a K-weighted [1 2 1]/4
 * pass; at the bulk edges (single rank: both are boundaries) the neighbour
 * outside is replaced by the cell itself.  The reference also filters the
 * first and last array cells, reading one past each end; here those two
 * are copied unchanged. */
void orc_gr1d_filter(const orc_mesh *m, const double *j, double *tmp, const double *k1) {
  const int n = m->dims[0];
  for (int i = 0; i < n; i++) {
    if (i == 0 || i == n - 1) {
      tmp[i] = j[i];
      continue;
    }
    int dp = (i == n - m->guard[0] - 1) ? 0 : 1;
    int dm = (i == m->guard[0]) ? 0 : 1;
    double v = 0.5 * j[i] * k1[i];
    v += 0.25 * j[i + dp] * k1[i + dp];
    v += 0.25 * j[i - dm] * k1[i - dm];
    v /= k1[i];
    tmp[i] = v;
  }
}

/* update_e_1dgr (field_solver_1dgr.cu:14-38): E1 += (j0 - J1) dt on
 * guard + 1 <= i < dims - guard - 5 */
void orc_gr1d_field_update(const orc_mesh *m, double *const *coef, double *e1,
                           const double *j1, double dt) {
  for (int i = m->guard[0] + 1; i < m->dims[0] - m->guard[0] - 5; i++)
    e1[i] += (coef[GR_J0][i] - j1[i]) * dt;
}
