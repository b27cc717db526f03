/*
 * ffe_oracle.c -- CPU FP64 restatement of Aperture's 3D log-spherical
 * force-free / vacuum field solver (ffe_solver_logsph).  TEST
 * INFRASTRUCTURE ONLY (the checker, never the product): see pic_oracle.h.
 *
 * Reference: src/cuda/algorithms/ffe_solver_logsph.cu
 *   rk_push (:969-1012): dB = -dt curl E (ffe_logsph_rk_update_b, :329-369),
 *     dE = dt curl(B - B0) - dt J (ffe_logsph_rk_update_e, :371-437), J = 0
 *     (the compute_j call is disabled, :995-1005);
 *   rk_update (:1014-1027, kernel :499-538): e = c1 en + c2 e + c3 de,
 *     de = e; b = c1 bn + c2 b + c3 db;
 *   update_fields (:880-930): SSP-RK3 with (1,0,1), (3/4,1/4,1/4),
 *     (1/3,2/3,2/3), the boundaries after each stage at t+dt, t+dt/2, t+dt;
 *   apply_boundary (:932-966): star (:701-745), absorbing layer (:801-827),
 *     axis (:747-799);
 *   clean_epar (:540-612), check_eGTb (:614-699) with the staggered
 *     midpoint interpolation of interpolation.cuh:211-237.
 * The kernels cover the bulk plus SHIFT_GHOST = 2 guard layers, so the
 * grid needs guard >= 3.  The metric coefficients are the 2D Grid_LogSph
 * arrays (dims0 x dims1), shared by every phi plane.  The stage time step
 * is the dt argument (the reference reads dev_params.delta_t, equal to it
 * in its lab).  exp/sin/cos on the boundary are orc_math.c's.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "orc_internal.h"

#define FFE_TINY ((double)1e-7f)
#define FFE_SHIFT 2

static inline double fpos(const orc_mesh *m, int d, int n, int st) {
  return m->lower[d] + m->delta[d] * ((double)(n - m->guard[d]) + (st ? 1.0 : 0.5));
}

/* star_field_b1/b2/b3 (ffe_solver_logsph.cu:56-74), B0() = 1000 */
static inline double star_b1(double r, double theta) {
  return 2.0 * 1000.0 * orc_cos(theta) / (r * r * r);
}
static inline double star_b2(double r, double theta) {
  return 1000.0 * orc_sin(theta) / (r * r * r);
}

#define IDX(i, j, k) ((size_t)(i) + (size_t)(j) * d0 + (size_t)(k) * d0 * d1)
#define IJ(i, j) ((size_t)(i) + (size_t)(j) * d0)

static void ffe_rk_push(const orc_mesh *m, const orc_logsph_coef *c,
                        orc_fields *f, double *dE[3], double *dB[3],
                        double dt) {
  const int d0 = m->dims[0], d1 = m->dims[1];
  const double *e1 = f->E[0], *e2 = f->E[1], *e3 = f->E[2];
  const double *b1 = f->B[0], *b2 = f->B[1], *b3 = f->B[2];
  const double *g1 = f->Bbg[0], *g2 = f->Bbg[1], *g3 = f->Bbg[2];
  const double *l1e = c->a[ORC_L1E], *l2e = c->a[ORC_L2E], *l3e = c->a[ORC_L3E];
  const double *l1b = c->a[ORC_L1B], *l2b = c->a[ORC_L2B], *l3b = c->a[ORC_L3B];
  const double *A1e = c->a[ORC_A1E], *A2e = c->a[ORC_A2E], *A3e = c->a[ORC_A3E];
  const double *A1b = c->a[ORC_A1B], *A2b = c->a[ORC_A2B], *A3b = c->a[ORC_A3B];
  const size_t dj = (size_t)d0, dk = (size_t)d0 * d1;
  /* ffe_logsph_rk_update_b */
  for (int k = m->guard[2] - FFE_SHIFT; k < m->dims[2] - m->guard[2] + FFE_SHIFT; k++)
    for (int j = m->guard[1] - FFE_SHIFT; j < m->dims[1] - m->guard[1] + FFE_SHIFT; j++)
      for (int i = m->guard[0] - FFE_SHIFT; i < m->dims[0] - m->guard[0] + FFE_SHIFT; i++) {
        size_t ijk = IDX(i, j, k), ij = IJ(i, j);
        dB[0][ijk] = dt * -(e3[ijk] * l3e[ij] - e3[ijk - dj] * l3e[ij - dj] -
                            (e2[ijk] - e2[ijk - dk]) * l2e[ij]) /
                     A1b[ij];
        /* A2_b vanishes on the theta = 0, pi node rows; the reference divides
         * 0 by it there (NaN that reaches the outer guard corners), here the
         * increment is 0 -- the axis condition zeroes B2 on those rows. */
        dB[1][ijk] = A2b[ij] == 0.0
                         ? 0.0
                         : dt * -((e1[ijk] - e1[ijk - dk]) * l1e[ij] -
                                  e3[ijk] * l3e[ij] + e3[ijk - 1] * l3e[ij - 1]) /
                               A2b[ij];
        dB[2][ijk] = dt * -(e2[ijk] * l2e[ij] - e2[ijk - 1] * l2e[ij - 1] -
                            e1[ijk] * l1e[ij] + e1[ijk - dj] * l1e[ij - dj]) /
                     A3b[ij];
      }
  /* ffe_logsph_rk_update_e (J = 0) */
  for (int k = m->guard[2] - FFE_SHIFT; k < m->dims[2] - m->guard[2] + FFE_SHIFT; k++)
    for (int j = m->guard[1] - FFE_SHIFT; j < m->dims[1] - m->guard[1] + FFE_SHIFT; j++)
      for (int i = m->guard[0] - FFE_SHIFT; i < m->dims[0] - m->guard[0] + FFE_SHIFT; i++) {
        size_t ijk = IDX(i, j, k), ij = IJ(i, j);
        dE[0][ijk] = dt *
                         ((b3[ijk + dj] - g3[ijk + dj]) * l3b[ij + dj] -
                          (b3[ijk] - g3[ijk]) * l3b[ij] -
                          (b2[ijk + dk] - g2[ijk + dk]) * l2b[ij] +
                          (b2[ijk] - g2[ijk]) * l2b[ij]) /
                         A1e[ij];
        dE[1][ijk] = dt *
                         ((b1[ijk + dk] - g1[ijk + dk] - b1[ijk] + g1[ijk]) * l1b[ij] -
                          (b3[ijk + 1] - g3[ijk + 1]) * l3b[ij + 1] +
                          (b3[ijk] - g3[ijk]) * l3b[ij]) /
                         A2e[ij];
        dE[2][ijk] = dt *
                         ((b2[ijk + 1] - g2[ijk + 1]) * l2b[ij + 1] -
                          (b2[ijk] - g2[ijk]) * l2b[ij] -
                          (b1[ijk + dj] - g1[ijk + dj]) * l1b[ij + dj] +
                          (b1[ijk] - g1[ijk]) * l1b[ij]) /
                         A3e[ij];
      }
}

static void ffe_rk_update(const orc_mesh *m, orc_fields *f, double *En[3],
                          double *Bn[3], double *dE[3], double *dB[3],
                          double c1, double c2, double c3) {
  const int d0 = m->dims[0], d1 = m->dims[1];
  for (int k = m->guard[2] - FFE_SHIFT; k < m->dims[2] - m->guard[2] + FFE_SHIFT; k++)
    for (int j = m->guard[1] - FFE_SHIFT; j < m->dims[1] - m->guard[1] + FFE_SHIFT; j++)
      for (int i = m->guard[0] - FFE_SHIFT; i < m->dims[0] - m->guard[0] + FFE_SHIFT; i++) {
        size_t ijk = IDX(i, j, k);
        for (int q = 0; q < 3; q++) {
          f->E[q][ijk] = c1 * En[q][ijk] + c2 * f->E[q][ijk] + c3 * dE[q][ijk];
          dE[q][ijk] = f->E[q][ijk];
        }
        for (int q = 0; q < 3; q++)
          f->B[q][ijk] = c1 * Bn[q][ijk] + c2 * f->B[q][ijk] + c3 * dB[q][ijk];
      }
}

/* periodic guard copy along the periodic dimensions (send_field_guard_cells
 * on one rank) */
static void ffe_guard(const orc_mesh *m, orc_fields *f) {
  for (int q = 0; q < 3; q++) {
    orc_guard_copy(m, f->E[q]);
    orc_guard_copy(m, f->B[q]);
  }
}

void orc_ffe_boundary(const orc_mesh *m, const orc_ffe_params *fp,
                      orc_fields *f, double omega, double t) {
  const int d0 = m->dims[0], d1 = m->dims[1];
  double *E1 = f->E[0], *E2 = f->E[1], *E3 = f->E[2];
  double *B1 = f->B[0], *B2 = f->B[1], *B3 = f->B[2];
  double *G1 = f->Bbg[0], *G2 = f->Bbg[1], *G3 = f->Bbg[2];
  /* ffe_logsph_stellar_boundary */
  for (int k = m->guard[2] - 1; k < m->dims[2] - m->guard[2]; k++)
    for (int j = m->guard[1] - 1; j < m->dims[1] - m->guard[1]; j++) {
      double theta_s = fpos(m, 1, j, 1), theta = fpos(m, 1, j, 0);
      for (int i = 0; i < m->guard[0]; i++) {
        double r_s = orc_exp(fpos(m, 0, i, 1));
        double r = orc_exp(fpos(m, 0, i, 0));
        size_t o = IDX(i, j, k);
        B1[o] = star_b1(r_s, theta);
        G1[o] = star_b1(r_s, theta);
        E3[o] = 0.0;
        E2[o] = -omega * orc_sin(theta) * star_b1(r_s, theta) * r_s;
        E1[o] = omega * orc_sin(theta_s) * star_b2(r, theta_s) * r;
        B2[o] = star_b2(r, theta_s);
        G2[o] = star_b2(r, theta_s);
        B3[o] = 0.0;
        G3[o] = 0.0;
      }
    }
  (void)t;
  /* ffe_logsph_outflow_boundary */
  const int L = fp->damping_length;
  for (int k = m->guard[2] - 1; k < m->dims[2] - m->guard[2]; k++)
    for (int j = m->guard[1] - 1; j < m->dims[1] - m->guard[1]; j++)
      for (int i = 0; i < L; i++) {
        size_t o = IDX(d0 - L + i, j, k);
        double x = (double)i / L;
        double lambda = 1.0 - fp->damping_coef * (x * x);
        E1[o] *= lambda;
        E2[o] *= lambda;
        E3[o] *= lambda;
        B3[o] *= lambda;
      }
  /* ffe_logsph_axis_boundary_lower / upper */
  const int g = m->guard[1], jl = m->dims[1] - m->guard[1];
  for (int k = m->guard[2]; k < m->dims[2] - m->guard[2]; k++)
    for (int i = m->guard[0]; i < m->dims[0] - m->guard[0]; i++) {
      E1[IDX(i, g - 1, k)] = 0.0;
      E3[IDX(i, g - 1, k)] = 0.0;
      E2[IDX(i, g - 1, k)] = -E2[IDX(i, g, k)];
      B3[IDX(i, g - 1, k)] = 0.0;
      B2[IDX(i, g - 1, k)] = 0.0;
      G2[IDX(i, g - 1, k)] = 0.0;
      B1[IDX(i, g - 1, k)] = B1[IDX(i, g, k)];
    }
  for (int k = m->guard[2]; k < m->dims[2] - m->guard[2]; k++)
    for (int i = m->guard[0]; i < m->dims[0] - m->guard[0]; i++) {
      E1[IDX(i, jl - 1, k)] = 0.0;
      E3[IDX(i, jl - 1, k)] = 0.0;
      E2[IDX(i, jl, k)] = -E2[IDX(i, jl - 1, k)];
      B3[IDX(i, jl - 1, k)] = 0.0;
      G2[IDX(i, jl - 1, k)] = 0.0;
      B2[IDX(i, jl - 1, k)] = 0.0;
      B1[IDX(i, jl, k)] = B1[IDX(i, jl - 1, k)];
    }
}

/* scratch: 12 arrays of ncell doubles (En, Bn, dE, dB) */
void orc_ffe_update(const orc_mesh *m, const orc_logsph_coef *c,
                    const orc_ffe_params *fp, orc_fields *f, double dt,
                    double omega, double time, double *scratch) {
  const size_t n = (size_t)m->dims[0] * m->dims[1] * m->dims[2];
  double *En[3], *Bn[3], *dE[3], *dB[3];
  for (int q = 0; q < 3; q++) {
    En[q] = scratch + (size_t)q * n;
    Bn[q] = scratch + (size_t)(3 + q) * n;
    dE[q] = scratch + (size_t)(6 + q) * n;
    dB[q] = scratch + (size_t)(9 + q) * n;
    memcpy(En[q], f->E[q], n * sizeof(double));
    memcpy(Bn[q], f->B[q], n * sizeof(double));
    memset(dE[q], 0, n * sizeof(double));
    memset(dB[q], 0, n * sizeof(double));
  }
  const double cs[3][3] = {{1.0, 0.0, 1.0}, {0.75, 0.25, 0.25},
                           {1.0 / 3.0, 2.0 / 3.0, 2.0 / 3.0}};
  const double ts[3] = {time + dt, time + 0.5 * dt, time + dt};
  for (int s = 0; s < 3; s++) {
    ffe_rk_push(m, c, f, dE, dB, dt);
    ffe_rk_update(m, f, En, Bn, dE, dB, cs[s][0], cs[s][1], cs[s][2]);
    orc_ffe_boundary(m, fp, f, omega, ts[s]);
    ffe_guard(m, f);
  }
}

/* one rk_push: dE, dB (each 3 x ncell) from the current fields */
void orc_ffe_rk_push(const orc_mesh *m, const orc_logsph_coef *c,
                     orc_fields *f, double *dE0, double *dE1, double *dE2,
                     double *dB0, double *dB1, double *dB2, double dt) {
  double *dE[3] = {dE0, dE1, dE2}, *dB[3] = {dB0, dB1, dB2};
  ffe_rk_push(m, c, f, dE, dB, dt);
}

/* interpolate(f, idx, in, out, dim0, dim1) (interpolation.cuh:211-237) */
static double stag_interp(const double *a, size_t idx, int in, int out,
                          int d0, int d1) {
  long di_m = ((in & 1) == (out & 1)) ? 0 : -1 - (out & 1);
  long di_p = ((in & 1) == (out & 1)) ? 0 : -(out & 1);
  long dj_m = (((in >> 1) & 1) == ((out >> 1) & 1)) ? 0 : -1 - ((out >> 1) & 1);
  long dj_p = (((in >> 1) & 1) == ((out >> 1) & 1)) ? 0 : -((out >> 1) & 1);
  long dk_m = (((in >> 2) & 1) == ((out >> 2) & 1)) ? 0 : -1 - ((out >> 2) & 1);
  long dk_p = (((in >> 2) & 1) == ((out >> 2) & 1)) ? 0 : -((out >> 2) & 1);
  long D0 = d0, D01 = (long)d0 * d1;
  double f11 = 0.5 * (a[idx + di_p + dj_p * D0 + dk_m * D01] +
                      a[idx + di_p + dj_p * D0 + dk_p * D01]);
  double f10 = 0.5 * (a[idx + di_p + dj_m * D0 + dk_m * D01] +
                      a[idx + di_p + dj_m * D0 + dk_p * D01]);
  double f01 = 0.5 * (a[idx + di_m + dj_p * D0 + dk_m * D01] +
                      a[idx + di_m + dj_p * D0 + dk_p * D01]);
  double f00 = 0.5 * (a[idx + di_m + dj_m * D0 + dk_m * D01] +
                      a[idx + di_m + dj_m * D0 + dk_p * D01]);
  double f1 = 0.5 * (f11 + f10);
  double f0 = 0.5 * (f01 + f00);
  return 0.5 * (f1 + f0);
}

static const int ST_E[3] = {0x6, 0x5, 0x3}, ST_B[3] = {0x1, 0x2, 0x4};

/* The midpoint interpolation reaches two cells down in each dimension;
 * the reference's range (guard - 2) reads before the array when guard = 3,
 * so clean_epar / check_eGTb start at max(guard - 2, 2). */
#define FIX_LO(m, d) ((m)->guard[d] - FFE_SHIFT > 2 ? (m)->guard[d] - FFE_SHIFT : 2)

/* ffe_clean_epar (:540-612): dE := E - (E.B) B / (B^2 + TINY), each
 * component at its own location (dE: 3 x ncell). */
void orc_ffe_clean_epar(const orc_mesh *m, orc_fields *f, double *dE) {
  const int d0 = m->dims[0], d1 = m->dims[1];
  const size_t n = (size_t)m->dims[0] * m->dims[1] * m->dims[2];
  for (int k = FIX_LO(m, 2); k < m->dims[2] - m->guard[2] + FFE_SHIFT; k++)
    for (int j = FIX_LO(m, 1); j < m->dims[1] - m->guard[1] + FFE_SHIFT; j++)
      for (int i = FIX_LO(m, 0); i < m->dims[0] - m->guard[0] + FFE_SHIFT; i++) {
        size_t ijk = IDX(i, j, k);
        for (int q = 0; q < 3; q++) {
          double ex = stag_interp(f->E[0], ijk, ST_E[0], ST_E[q], d0, d1);
          double ey = stag_interp(f->E[1], ijk, ST_E[1], ST_E[q], d0, d1);
          double ez = stag_interp(f->E[2], ijk, ST_E[2], ST_E[q], d0, d1);
          double bx = stag_interp(f->B[0], ijk, ST_B[0], ST_E[q], d0, d1);
          double by = stag_interp(f->B[1], ijk, ST_B[1], ST_E[q], d0, d1);
          double bz = stag_interp(f->B[2], ijk, ST_B[2], ST_E[q], d0, d1);
          double bq = q == 0 ? bx : (q == 1 ? by : bz);
          dE[(size_t)q * n + ijk] =
              f->E[q][ijk] - (ex * bx + ey * by + ez * bz) * bq /
                                 (bx * bx + by * by + bz * bz + FFE_TINY);
        }
      }
}

/* ffe_check_eGTb (:614-699): E := s dE with s = sqrt(|B|^2/|E|^2) where
 * |E| > |B| (both interpolated to the component's location), else 1.  The
 * magnitudes are taken from E before any component is rewritten (the
 * reference reads neighbours other threads may already have rewritten);
 * dE is overwritten with the result. */
void orc_ffe_check_egtb(const orc_mesh *m, orc_fields *f, double *dE) {
  const int d0 = m->dims[0], d1 = m->dims[1];
  const size_t n = (size_t)m->dims[0] * m->dims[1] * m->dims[2];
  for (int k = FIX_LO(m, 2); k < m->dims[2] - m->guard[2] + FFE_SHIFT; k++)
    for (int j = FIX_LO(m, 1); j < m->dims[1] - m->guard[1] + FFE_SHIFT; j++)
      for (int i = FIX_LO(m, 0); i < m->dims[0] - m->guard[0] + FFE_SHIFT; i++) {
        size_t ijk = IDX(i, j, k);
        for (int q = 0; q < 3; q++) {
          double ex = stag_interp(f->E[0], ijk, ST_E[0], ST_E[q], d0, d1);
          double ey = stag_interp(f->E[1], ijk, ST_E[1], ST_E[q], d0, d1);
          double ez = stag_interp(f->E[2], ijk, ST_E[2], ST_E[q], d0, d1);
          double emag = ex * ex + ey * ey + ez * ez + FFE_TINY;
          double bx = stag_interp(f->B[0], ijk, ST_B[0], ST_E[q], d0, d1);
          double by = stag_interp(f->B[1], ijk, ST_B[1], ST_E[q], d0, d1);
          double bz = stag_interp(f->B[2], ijk, ST_B[2], ST_E[q], d0, d1);
          double bmag = bx * bx + by * by + bz * bz + FFE_TINY;
          double temp = emag > bmag ? sqrt(bmag / emag) : 1.0;
          dE[(size_t)q * n + ijk] = temp * dE[(size_t)q * n + ijk];
        }
      }
  for (int k = FIX_LO(m, 2); k < m->dims[2] - m->guard[2] + FFE_SHIFT; k++)
    for (int j = FIX_LO(m, 1); j < m->dims[1] - m->guard[1] + FFE_SHIFT; j++)
      for (int i = FIX_LO(m, 0); i < m->dims[0] - m->guard[0] + FFE_SHIFT; i++) {
        size_t ijk = IDX(i, j, k);
        for (int q = 0; q < 3; q++) f->E[q][ijk] = dE[(size_t)q * n + ijk];
      }
}
#undef IDX
#undef IJ
