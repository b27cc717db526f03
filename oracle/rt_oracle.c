/*
 * rt_oracle.c -- CPU FP64 restatement of Aperture's pair-creation Monte
 * Carlo and photon transport.  TEST INFRASTRUCTURE ONLY (the checker, never
 * the product): see pic_oracle.h.
 *
 * Reference files restated (relative to the Aperture3 tree):
 *   src/cuda/radiation/radiative_transfer.cu       emit_photons /
 *       produce_pairs: decide, compact, append (:28-177, :198-301)
 *   src/cuda/radiation/user_radiative_functions.cu default hooks (:13-125)
 *   src/cuda/algorithms/ptc_updater.cu             move_photons_cart_{1,2,3}d
 *                                                  (:627-741)
 *   src/cuda/algorithms/ptc_updater_logsph.cu      move_photons (:58-132)
 *
 * Determinism (the reference uses per-thread curand states and per-block
 * atomic slot counters, so its event order depends on the launch): here
 * the k-th emitting particle in slot order writes photon slot number + k,
 * the k-th converting photon writes particle slots number + 2k, 2k+1, and
 * every random draw is Philox4x32-10 keyed on (seed, slot, call epoch), so
 * the event set is a pure function of the inputs.  The HIP product uses the
 * same keys and the same prefix-sum order.
 *
 * Models:
 *   0 -- the reference's default hooks: emit when gamma > gamma_thr, photon
 *        energy 2.5 + 2u(E_secondary - 1), free path lph (1/2 + u/2) with
 *        lph = min(10, (1/sin theta - 1) photon_path), no photon within
 *        0.005 of the axis or beyond r_cutoff; convert when path_left <= 0.
 *   1 -- fixed energy fraction with an exponential free path: emit when
 *        gamma > gamma_thr, photon energy e_frac gamma, free path
 *        -photon_path ln u (mean photon_path), dropped beyond r_cutoff.
 * Both convert with produce_pair (user_radiative_functions.cu:70-125):
 * each member gets ratio = sqrt(1/4 - 1/E^2) of the photon momentum.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "orc_internal.h"

#define RT_PI 3.14159265358979323846
#define FLAG_SECONDARY (1u << 7)

static inline double rt_pos(const orc_mesh *m, int d, int n, double x) {
  if (d < m->dim) return m->lower[d] + m->delta[d] * ((double)(n - m->guard[d]) + x);
  return x * m->delta[d];
}

static inline void split_cell(const orc_mesh *m, uint32_t c, int *c1, int *c2,
                              int *c3) {
  *c1 = (int)(c % (uint32_t)m->dims[0]);
  *c2 = (int)((c / (uint32_t)m->dims[0]) % (uint32_t)m->dims[1]);
  *c3 = (int)(c / ((uint32_t)m->dims[0] * (uint32_t)m->dims[1]));
}

/* count_photon_produced + produce_photons (radiative_transfer.cu:28-96)
 * with check_emit_photon / emit_photon.  Returns the number of emitting
 * particles (= photon slots reserved); *overflow counts emitters dropped
 * for lack of photon capacity.  `produced` (may be NULL) gains every
 * emitter's weight at its cell, before any capacity or path test
 * (:48-57). */
uint64_t orc_emit_photons(const orc_mesh *m, const orc_rt_params *rt,
                          orc_ptc *p, orc_photons *ph, uint64_t seed,
                          uint32_t epoch, uint64_t id_base,
                          uint64_t *overflow, double *produced) {
  uint64_t base = ph->number, k = 0;
  for (uint64_t idx = 0; idx < p->number; idx++) {
    uint32_t c = p->cell[idx];
    if (c == ORC_MAX_CELL || !is_in_bulk(m, c)) continue;
    uint32_t flag = p->flag[idx];
    if ((flag >> 29) == 2) continue; /* ions do not radiate */
    double gamma = p->E[idx];
    if (!(gamma > rt->gamma_thr)) continue;
    if (produced) produced[c] += p->w[idx];
    uint64_t slot = base + k;
    if (slot >= rt->max_photon) {
      if (overflow) (*overflow)++;
      continue;
    }
    k++;
    ph->cell[slot] = ORC_MAX_CELL;
    ph->id[slot] = id_base + (slot - base);
    double u[2];
    orc_uniform2(seed, idx, epoch, 5, u);
    double p1 = p->p1[idx], p2 = p->p2[idx], p3 = p->p3[idx];
    double pi = sqrt(gamma * gamma - 1.0);
    double Eph;
    if (rt->model == 0)
      Eph = 2.5 + u[0] * (rt->E_secondary - 1.0) * 2.0;
    else
      Eph = rt->e_frac * gamma;
    double ge = gamma - Eph;
    double pf2 = ge * ge - 1.0;
    double pf = sqrt(pf2 > 0.0 ? pf2 : 0.0);
    p->p1[idx] = p1 * pf / pi;
    p->p2[idx] = p2 * pf / pi;
    p->p3[idx] = p3 * pf / pi;
    p->E[idx] = ge;
    double path;
    if (rt->model == 0) {
      int c1, c2, c3;
      split_cell(m, c, &c1, &c2, &c3);
      double theta = rt_pos(m, 1, c2, p->x2[idx]);
      double lph = (1.0 / orc_sin(theta) - 1.0) * rt->photon_path;
      if (lph > 10.0) lph = 10.0;
      if (theta < 0.005 || theta > RT_PI - 0.005) continue;
      path = lph * (0.5 + 0.5 * u[1]);
    } else {
      path = -rt->photon_path * orc_log(u[1]);
    }
    if (path > rt->r_cutoff) continue;
    ph->x1[slot] = p->x1[idx];
    ph->x2[slot] = p->x2[idx];
    ph->x3[slot] = p->x3[idx];
    ph->p1[slot] = Eph * p1 / pi;
    ph->p2[slot] = Eph * p2 / pi;
    ph->p3[slot] = Eph * p3 / pi;
    ph->E[slot] = Eph;
    ph->w[slot] = p->w[idx];
    ph->path[slot] = path;
    ph->flag[slot] = 0;
    ph->cell[slot] = c;
  }
  ph->number = base + k;
  return k;
}

/* count_pairs_produced + produce_pairs (radiative_transfer.cu:99-177) with
 * check_produce_pair / produce_pair.  Returns the number of pairs made;
 * photons past the particle capacity are removed and counted.  `produced`
 * (may be NULL) gains every converting photon's weight at its cell
 * (:120-130). */
uint64_t orc_produce_pairs(const orc_mesh *m, const orc_rt_params *rt,
                           orc_ptc *p, orc_photons *ph, uint64_t id_base,
                           uint64_t *overflow, double *produced) {
  uint64_t base = p->number, k = 0;
  for (uint64_t t = 0; t < ph->number; t++) {
    uint32_t c = ph->cell[t];
    if (c == ORC_MAX_CELL || !is_in_bulk(m, c)) continue;
    if (!(ph->path[t] <= 0.0)) continue;
    if (produced) produced[c] += ph->w[t];
    uint64_t off = base + 2 * k;
    if (off + 1 >= rt->max_ptc) {
      if (overflow) (*overflow)++;
      ph->cell[t] = ORC_MAX_CELL;
      continue;
    }
    k++;
    double p1 = ph->p1[t], p2 = ph->p2[t], p3 = ph->p3[t];
    double E_ph2 = p1 * p1 + p2 * p2 + p3 * p3;
    if (E_ph2 <= 4.01) E_ph2 = 4.01;
    double ratio = sqrt(0.25 - 1.0 / E_ph2);
    double gamma = sqrt(1.0 + ratio * ratio * E_ph2);
    for (int s = 0; s < 2; s++) {
      uint64_t o = off + s;
      p->x1[o] = ph->x1[t];
      p->x2[o] = ph->x2[t];
      p->x3[o] = ph->x3[t];
      p->p1[o] = ratio * p1;
      p->p2[o] = ratio * p2;
      p->p3[o] = ratio * p3;
      p->E[o] = gamma;
      p->w[o] = ph->w[t];
      p->cell[o] = c;
      p->flag[o] = FLAG_SECONDARY | ((uint32_t)s << 29);
      p->id[o] = id_base + 2 * (k - 1) + s;
    }
    ph->cell[t] = ORC_MAX_CELL;
  }
  p->number = base + 2 * k;
  return k;
}

/* move_photons_cart_{1,2,3}d (ptc_updater.cu:627-741) with the periodic
 * wrap that send_particles does for photons. */
static void move_photon_cart(const orc_mesh *m, orc_photons *ph, uint64_t t,
                             double dt) {
  uint32_t c = ph->cell[t];
  double p1 = ph->p1[t], p2 = ph->p2[t], p3 = ph->p3[t];
  double E = sqrt(p1 * p1 + p2 * p2 + p3 * p3);
  double x1 = ph->x1[t], x2 = ph->x2[t], x3 = ph->x3[t];
  int c1, c2, c3;
  split_cell(m, c, &c1, &c2, &c3);
  int dc[3] = {0, 0, 0};
  double nx[3];
  if (m->dim == 1) {
    nx[0] = x1 + p1 / E * dt * m->inv_delta[0];
    dc[0] = (int)floor(nx[0]);
    nx[0] -= (double)dc[0];
    nx[1] = x2 + p2 * dt / E;
    nx[2] = x3 + p3 * dt / E;
  } else {
    double v1 = p1 / E, v2 = p2 / E, v3 = p3 / E;
    nx[0] = x1 + v1 * dt * m->inv_delta[0];
    nx[1] = x2 + v2 * dt * m->inv_delta[1];
    dc[0] = (int)floor(nx[0]);
    dc[1] = (int)floor(nx[1]);
    nx[0] -= (double)dc[0];
    nx[1] -= (double)dc[1];
    if (m->dim == 2) {
      nx[2] = x3 + v3 * dt;
    } else {
      nx[2] = x3 + v3 * dt * m->inv_delta[2];
      dc[2] = (int)floor(nx[2]);
      nx[2] -= (double)dc[2];
    }
  }
  if (dc[0] < -1 || dc[0] > 1 || dc[1] < -1 || dc[1] > 1 || dc[2] < -1 ||
      dc[2] > 1) {
    ph->cell[t] = ORC_MAX_CELL;
    return;
  }
  int n[3] = {c1 + dc[0], c2 + dc[1], c3 + dc[2]};
  for (int d = 0; d < m->dim; d++) {
    if (!m->periodic[d]) continue;
    int g = m->guard[d], N = m->dims[d] - 2 * g;
    if (n[d] < g) n[d] += N;
    else if (n[d] >= N + g) n[d] -= N;
  }
  ph->cell[t] = (uint32_t)(n[0] + n[1] * m->dims[0] +
                           n[2] * m->dims[0] * m->dims[1]);
  ph->x1[t] = nx[0];
  ph->x2[t] = nx[1];
  ph->x3[t] = nx[2];
  ph->path[t] -= dt;
}

/* move_photons (ptc_updater_logsph.cu:58-132).  As in the reference the
 * cell index is taken before the axis reflection, so a photon that crosses
 * the axis lands in the guard row and is removed by the guard clear. */
static void move_photon_logsph(const orc_mesh *m, double r_cutoff,
                               orc_photons *ph, uint64_t t, double dt) {
  uint32_t c = ph->cell[t];
  const int d0 = m->dims[0];
  int c1 = (int)(c % (uint32_t)d0), c2 = (int)(c / (uint32_t)d0);
  double v1 = ph->p1[t], v2 = ph->p2[t], v3 = ph->p3[t];
  double E = sqrt(v1 * v1 + v2 * v2 + v3 * v3);
  v1 = v1 / E;
  v2 = v2 / E;
  v3 = v3 / E;
  double old_x1 = ph->x1[t], old_x2 = ph->x2[t], old_x3 = ph->x3[t];
  double r1 = rt_pos(m, 0, c1, old_x1);
  double exp_r1 = orc_exp(r1);
  if (exp_r1 > r_cutoff || exp_r1 < 1.02) {
    ph->cell[t] = ORC_MAX_CELL;
    return;
  }
  double r2 = rt_pos(m, 1, c2, old_x2);
  double s2, c2v, s3, c3;
  orc_sincos(r2, &s2, &c2v);
  orc_sincos(old_x3, &s3, &c3);
  double x = exp_r1 * s2 * c3;
  double y = exp_r1 * s2 * s3;
  double z = exp_r1 * c2v;
  {
    double a = v1, b = v2, cc = v3;
    v1 = a * s2 * c3 + b * c2v * c3 - cc * s3;
    v2 = a * s2 * s3 + b * c2v * s3 + cc * c3;
    v3 = a * c2v - b * s2;
  }
  x += v1 * dt;
  y += v2 * dt;
  z += v3 * dt;
  double r1p = sqrt(x * x + y * y + z * z);
  double r2p = orc_acos(z / r1p);
  r1p = orc_log(r1p);
  double r3p = orc_atan2(y, x);
  {
    double ps2, pc2, ps3, pc3;
    orc_sincos(r2p, &ps2, &pc2);
    orc_sincos(r3p, &ps3, &pc3);
    double a = v1, b = v2, cc = v3;
    v1 = a * ps2 * pc3 + b * ps2 * ps3 + cc * pc2;
    v2 = a * pc2 * pc3 + b * pc2 * ps3 - cc * ps2;
    v3 = -a * ps3 + b * pc3;
  }
  ph->p1[t] = v1 * E;
  ph->p2[t] = v2 * E;
  ph->p3[t] = v3 * E;
  double new_x1 = old_x1 + (r1p - r1) / m->delta[0];
  double new_x2 = old_x2 + (r2p - r2) / m->delta[1];
  int dc1 = (int)floor(new_x1);
  int dc2 = (int)floor(new_x2);
  if (dc1 < -1 || dc1 > 1 || dc2 < -1 || dc2 > 1) {
    ph->cell[t] = ORC_MAX_CELL;
    return;
  }
  ph->cell[t] = (uint32_t)((c1 + dc1) + (c2 + dc2) * d0);
  new_x1 -= (double)dc1;
  new_x2 -= (double)dc2;
  if (rt_pos(m, 1, c2 + dc2, new_x2) < 0.0) {
    new_x2 = 1.0 - new_x2;
  } else if (rt_pos(m, 1, c2 + dc2, new_x2) >= RT_PI) {
    new_x2 = 1.0 - new_x2;
  }
  ph->x1[t] = new_x1;
  ph->x2[t] = new_x2;
  ph->x3[t] = r3p;
  ph->path[t] -= dt;
}

void orc_move_photons(const orc_mesh *m, int logsph, double r_cutoff,
                      orc_photons *ph, double dt) {
  for (uint64_t t = 0; t < ph->number; t++) {
    if (ph->cell[t] == ORC_MAX_CELL) continue;
    if (logsph)
      move_photon_logsph(m, r_cutoff, ph, t, dt);
    else
      move_photon_cart(m, ph, t, dt);
  }
}

/* photons.clear_guard_cells (ptc_updater_logsph.cu:819) */
void orc_clear_photon_guard_cells(const orc_mesh *m, orc_photons *ph) {
  for (uint64_t t = 0; t < ph->number; t++)
    if (ph->cell[t] != ORC_MAX_CELL && !is_in_bulk(m, ph->cell[t]))
      ph->cell[t] = ORC_MAX_CELL;
}

/* photons.sort_by_cell (sim_data_impl.hpp:80, particle_base_impl.cu:435-464):
 * stable counting sort by cell, empty slots last and dropped (the guard
 * clear removes out-of-bulk photons beforehand).  scratch: number uint64. */
uint64_t orc_sort_photons(const orc_mesh *m, orc_photons *ph,
                          uint64_t *scratch) {
  uint64_t n = ph->number, kept = 0;
  for (uint64_t t = 0; t < n; t++)
    if (ph->cell[t] != ORC_MAX_CELL) kept++;
  /* order: stable by cell */
  size_t nc = (size_t)m->dims[0] * m->dims[1] * m->dims[2];
  uint64_t *cnt = (uint64_t *)calloc(nc + 1, sizeof(uint64_t));
  for (uint64_t t = 0; t < n; t++) {
    uint32_t c = ph->cell[t];
    size_t key = c != ORC_MAX_CELL ? c : nc;
    cnt[key]++;
  }
  uint64_t acc = 0;
  for (size_t k = 0; k <= nc; k++) {
    uint64_t v = cnt[k];
    cnt[k] = acc;
    acc += v;
  }
  for (uint64_t t = 0; t < n; t++) {
    uint32_t c = ph->cell[t];
    size_t key = c != ORC_MAX_CELL ? c : nc;
    scratch[cnt[key]++] = t;
  }
  free(cnt);
#define PGATHER(T, a)                                          \
  do {                                                         \
    T *tmp_ = (T *)malloc(sizeof(T) * (n ? n : 1));            \
    for (uint64_t q = 0; q < n; q++) tmp_[q] = (a)[scratch[q]]; \
    memcpy((a), tmp_, sizeof(T) * n);                          \
    free(tmp_);                                                \
  } while (0)
  PGATHER(uint32_t, ph->cell);
  PGATHER(double, ph->x1);
  PGATHER(double, ph->x2);
  PGATHER(double, ph->x3);
  PGATHER(double, ph->p1);
  PGATHER(double, ph->p2);
  PGATHER(double, ph->p3);
  PGATHER(double, ph->E);
  PGATHER(double, ph->w);
  PGATHER(double, ph->path);
  PGATHER(uint32_t, ph->flag);
  PGATHER(uint64_t, ph->id);
#undef PGATHER
  for (uint64_t t = kept; t < n; t++) ph->cell[t] = ORC_MAX_CELL;
  ph->number = kept;
  return kept;
}

/* Wall injection of the batched 1D ensemble (no reference counterpart; the
 * workload's open left wall): n_pairs co-located e-/e+ pairs in the first
 * bulk cell, offset 1 - u (Philox purpose 7, key (epoch << 16) | n),
 * Maxwell-Juttner momenta at temperature theta (key 2^63 | (pair << 1) | s)
 * with p1 turned inward.  Pairs that do not fit below `cap` are dropped and
 * counted.  Returns the new particle count. */
uint64_t orc_inject_left(const orc_mesh *m, orc_ptc *p, int n_pairs,
                         double theta, double weight, uint64_t seed,
                         uint32_t epoch, uint64_t id_base, uint64_t cap,
                         uint64_t *overflow) {
  uint64_t base = p->number, made = 0;
  for (int n = 0; n < n_pairs; n++) {
    if (base + 2 * made + 2 > cap) {
      if (overflow) *overflow += (uint64_t)(n_pairs - n);
      break;
    }
    uint64_t pair = ((uint64_t)epoch << 16) | (uint64_t)n;
    double a[2];
    orc_uniform2(seed, pair, 0, 7, a);
    for (int s = 0; s < 2; s++) {
      uint64_t o = base + 2 * made + s;
      double u[3];
      orc_sample_mj(theta, 0.0, seed, (1ull << 63) | (pair << 1) | (uint64_t)s, u);
      p->cell[o] = (uint32_t)m->guard[0];
      p->x1[o] = 1.0 - a[0];
      p->x2[o] = 0.0;
      p->x3[o] = 0.0;
      p->p1[o] = fabs(u[0]);
      p->p2[o] = u[1];
      p->p3[o] = u[2];
      p->E[o] = sqrt(1.0 + p->p1[o] * p->p1[o] + u[1] * u[1] + u[2] * u[2]);
      p->w[o] = weight;
      p->flag[o] = (uint32_t)s << 29;
      p->id[o] = id_base + 2 * made + s;
    }
    made++;
  }
  p->number = base + 2 * made;
  return p->number;
}
