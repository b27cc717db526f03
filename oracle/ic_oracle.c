/*
 * ic_oracle.c -- CPU restatement of the reference's inverse-Compton tables
 * and sampling (src/cuda/radiation/rt_ic.cu:22-547; soft-photon spectra of
 * src/radiation/spectra.h:10-95) and of its triplet pair-production tables
 * (src/cuda/radiation/rt_tpp.cu:21-294).  TEST INFRASTRUCTURE ONLY: the
 * checker for the HIP path, never linked into the product.
 *
 * The tables are the reference's host loops (rt_ic.cu:279-421) with this
 * build's exp/log (orc_math.c) and pow(x, y) = exp(y ln x), so the GPU
 * tables agree bit for bit.  beta(gamma) is 1/sqrt(1 - 1/gamma^2) as in
 * the reference (rt_ic.cu:22-25).
 */
#include <math.h>
#include <stdint.h>

#include "orc_internal.h"

#define IC_MIN_GAMMA ((double)1.01f)
#define IC_MAX_GAMMA 1.0e15
#define IC_MIN_EP 1.0e-10
#define IC_N_MU 100
#define IC_N_E 800
#define IC_RE_SQUARE 7.91402e-26
#define IC_PI 3.14159265358979323846

static double sqr(double x) { return x * x; }
static double ic_pow(double x, double y) { return orc_exp(y * orc_log(x)); }

/* spectra.h: kind 0 hard power law (delta, emin, emax), 1 soft power law
 * (alpha, emin, emax), 2 broken power law (alpha, delta, epeak, emin,
 * emax), 3 black body (kT), 4 mono-energetic (e0, de) */
double orc_ic_spectrum(int kind, const double *p, double e) {
  switch (kind) {
    case 0: return (e < p[2] && e > p[1]) ? ic_pow(e / p[2], p[0]) / e : 0.0;
    case 1: return (e < p[2] && e > p[1]) ? ic_pow(e / p[1], -p[0]) / e : 0.0;
    case 2:
      if (e < p[2] && e > p[3]) return ic_pow(e / p[2], p[1]) / e;
      if (e > p[2] && e < p[4]) return ic_pow(e / p[2], -p[0]) / e;
      return 0.0;
    case 3: return 1.75464e30 * e * e / (orc_exp(e / p[0]) - 1.0);
    default: return (e < p[0] + p[1] && e > p[0] - p[1]) ? 1.0 / (2.0 * p[1]) : 0.0;
  }
}

static double ic_beta(double gamma) { return 1.0 / sqrt(1.0 - 1.0 / sqr(gamma)); }

double orc_sigma_ic(double x) {
  if (x < 1.0e-3) return 1.0f - 2.0f * x + 26.0f * x * x / 5.0f;
  double l = orc_log(1.0 + 2.0 * x);
  return 0.75 * ((1.0 + x) * (2.0 * x * (1.0 + x) / (1.0 + 2.0 * x) - l) / (x * x * x) +
                 0.5 * l / x - (1.0 + 3.0 * x) / sqr(1.0 + 2.0 * x));
}

double orc_sigma_gg(double b) {
  return (1.0 - sqr(b)) *
         ((3.0 - b * b * b * b) * orc_log((1.0 + b) / (1.0 - b)) - 2.0 * b * (2.0 - b * b));
}

static double sigma_lab(double q, double ge) {
  return 2.0 * q * orc_log(q) + (1.0 + 2.0 * q) * (1.0 - q) +
         0.5 * sqr(ge * q) * (1.0 - q) / (1.0 + q * ge);
}

/* the grids of inverse_compton::inverse_compton (rt_ic.cu:227-262) */
void orc_ic_grids(int ng, int nep, double *gammas, double *ep, double *log_ep, double *dg,
                  double *dep, double *dlep) {
  *dep = 1.0 / ((double)nep - 1.0);
  for (int n = 0; n < nep; n++) ep[n] = *dep * n;
  *dlep = (-orc_log(IC_MIN_EP)) / ((double)nep - 1.0);
  for (int n = 0; n < nep; n++) log_ep[n] = orc_exp(*dlep * n + orc_log(IC_MIN_EP));
  *dg = (orc_log(IC_MAX_GAMMA) - orc_log(IC_MIN_GAMMA)) / ((double)ng - 1.0);
  for (int n = 0; n < ng; n++) gammas[n] = orc_exp(orc_log(IC_MIN_GAMMA) + *dg * (double)n);
}

/* inverse_compton::init (rt_ic.cu:279-421): ic_rate, gg_rate and the two
 * cumulative spectra dnde (on ep) and dnde_th (on log_ep), ng x nep */
void orc_ic_init(int kind, const double *p, double emin, double emax, double n0, int ng,
                 int nep, const double *gammas, const double *ep, const double *log_ep,
                 double *ic_rate, double *gg_rate, double *dnde, double *dnde_th) {
  const double dmu = 2.0 / (IC_N_MU - 1.0);
  const double de = (orc_log(emax) - orc_log(emin)) / (IC_N_E - 1.0);
  for (int n = 0; n < ng; n++) {
    double gamma = gammas[n], result = 0.0;
    for (int i_mu = 0; i_mu < IC_N_MU; i_mu++) {
      double mu = i_mu * dmu - 1.0;
      for (int i_e = 0; i_e < IC_N_E; i_e++) {
        double e = orc_exp(orc_log(emin) + i_e * de);
        double x = gamma * e * (1.0 - mu * ic_beta(gamma));
        double sigma = orc_sigma_ic(x);
        result += 0.5f * orc_ic_spectrum(kind, p, e) * sigma * (1.0f - ic_beta(gamma) * mu) * e;
      }
    }
    ic_rate[n] = result * dmu * de * (IC_RE_SQUARE * n0) * 8.0 * IC_PI / 3.0;
  }
  for (int n = 0; n < ng; n++) {
    double eph = gammas[n];
    if (eph < 2.0) {
      gg_rate[n] = 0.0;
      continue;
    }
    double result = 0.0;
    for (int i_mu = 0; i_mu < IC_N_MU; i_mu++) {
      double mu = i_mu * dmu - 1.0;
      for (int i_e = 0; i_e < IC_N_E; i_e++) {
        double e = orc_exp(orc_log(emin) + i_e * de);
        double s = eph * e * (1.0 - mu) * 0.5;
        if (s <= 1.0) continue;
        double b = sqrt(1.0 - 1.0 / s);
        if (b == 1.0) continue;
        result += orc_sigma_gg(b) * (1.0 - mu) * orc_ic_spectrum(kind, p, e) * e;
      }
    }
    gg_rate[n] = 0.25 * result * dmu * de * IC_PI * (n0 * IC_RE_SQUARE);
  }
  for (int pass = 0; pass < 2; pass++) {
    double *table = pass ? dnde_th : dnde;
    const double *e1s = pass ? log_ep : ep;
    for (int n = 0; n < ng; n++) {
      double gamma = gammas[n];
      double *row = table + (size_t)n * nep;
      for (int i = 0; i < nep; i++) {
        double e1 = e1s[i], result = 0.0;
        for (int i_e = 0; i_e < IC_N_E; i_e++) {
          double e = orc_exp(orc_log(emin) + i_e * de);
          double nv = orc_ic_spectrum(kind, p, e);
          double ge = gamma * e * 4.0;
          double q = e1 / (ge * (1.0 - e1));
          if (e1 < ge / (1.0 + ge) && e1 > e / gamma) result += nv * sigma_lab(q, ge) / gamma;
        }
        row[i] = pass ? result * de * e1 : result * de;
      }
      for (int i = 1; i < nep; i++) row[i] += row[i - 1];
      double last = row[nep - 1];
      for (int i = 0; i < nep; i++) row[i] /= last;
    }
  }
}

static int ic_search(float u, const double *a, int size, double *l, double *h) {
  int lo = 0, hi = size - 1, tmp = 0;
  double v = 0.0;
  while (lo < hi) {
    tmp = (lo + hi) / 2;
    v = a[tmp];
    if (v < u) {
      lo = tmp + 1;
    } else if (v > u) {
      hi = tmp - 1;
    } else {
      hi = tmp;
      *l = *h = v;
      return hi;
    }
  }
  if (v < u) {
    *l = v;
    *h = lo < size ? a[lo] : v;
    return tmp;
  }
  *h = v;
  *l = hi >= 0 ? a[hi] : v;
  return hi > 0 ? hi : 0;
}

/* find_n_gamma (rt_ic.cu:74-79); a NaN gamma is taken as the first point */
static int find_n_gamma(double gamma, int ng, double dg) {
  if (gamma < IC_MIN_GAMMA) return 0;
  if (gamma > IC_MAX_GAMMA) return ng - 1;
  double r = (orc_log(gamma) - orc_log(IC_MIN_GAMMA)) / dg;
  if (!(r >= 0.0)) return 0;
  int gn = (int)r;
  return gn > ng - 1 ? ng - 1 : gn;
}

/* find_ic_rate / find_gg_rate (rt_ic.cu:156-171), the upper neighbour of
 * the last grid point being the point itself */
static double find_rate(int ng, double dg, const double *gammas, const double *rate, double g,
                        int gn) {
  int gu = gn + 1 < ng ? gn + 1 : gn;
  if (g > IC_MAX_GAMMA) return rate[gn];
  double x = (orc_log(g) - orc_log(gammas[gn])) / dg;
  return rate[gn] * (1.0f - x) + rate[gu] * x;
}

/* Kernels::gen_photon_e (rt_ic.cu:127-153) for the draw u */
static double gen_photon_e(int ng, int nep, double dg, double dep, double dlep,
                           const double *dnde, const double *dnde_th, double g, float u) {
    double l = 0.0, h = 0.0, bb, result;
    int gn = find_n_gamma(g, ng, dg);
    int b = ic_search(u, dnde + (size_t)gn * nep, nep, &l, &h);
  if (b < 2) {
      b = ic_search(u, dnde_th + (size_t)gn * nep, nep, &l, &h);
      bb = (l == h ? b : (u - l) / (h - l) + b);
      result = orc_exp(dlep * bb + orc_log(IC_MIN_EP));
    } else {
      bb = (l == h ? b : (u - l) / (h - l) + b);
      result = dep * bb;
    }
    double v = result * g;
    return fmax(fmin(v, g - 1.01), 0.0);
}

/* accept 0: gen_photon_e with u the first uniform of the Philox block
 * (seed, k, epoch, purpose 9); accept 1: gen_photon_energies
 * (rt_ic.cu:198-212), the first uniform accepting when below find_ic_rate
 * and the second drawing the energy, else 0 */
void orc_ic_gen_photon_e(int ng, int nep, double dg, double dep, double dlep,
                         const double *gammas, const double *ic_rate, const double *dnde,
                         const double *dnde_th, const double *gamma, double *eph, uint64_t n,
                         uint64_t seed, uint32_t epoch, int accept) {
  for (uint64_t k = 0; k < n; k++) {
    double uu[2];
    orc_uniform2(seed, k, epoch, 9, uu);
    double g = gamma[k];
    if (!accept) {
      eph[k] = gen_photon_e(ng, nep, dg, dep, dlep, dnde, dnde_th, g, (float)uu[0]);
    } else {
      int gn = find_n_gamma(g, ng, dg);
      eph[k] = (float)uu[0] < find_rate(ng, dg, gammas, ic_rate, g, gn)
                   ? gen_photon_e(ng, nep, dg, dep, dlep, dnde, dnde_th, g, (float)uu[1])
                   : 0.0;
    }
  }
}

/* gen_rand_gammas (rt_ic.cu:214-225), u the first uniform of the Philox
 * block (seed, k, epoch, purpose 10) */
void orc_ic_random_gamma(double *gammas, uint64_t n, uint64_t seed, uint32_t epoch) {
  for (uint64_t k = 0; k < n; k++) {
    double uu[2];
    orc_uniform2(seed, k, epoch, 10, uu);
    float u = (float)uu[0];
    gammas[k] = orc_exp(orc_log(1.01) + u * (orc_log(1e10) - orc_log(1.01)));
  }
}

void orc_ic_rates(int ng, double dg, const double *gammas, const double *ic_rate,
                  const double *gg_rate, const double *gamma, double *ic, double *gg,
                  uint64_t n) {
  for (uint64_t k = 0; k < n; k++) {
    double g = gamma[k];
    int gn = find_n_gamma(g, ng, dg);
    ic[k] = find_rate(ng, dg, gammas, ic_rate, g, gn);
    gg[k] = g < 2.0 ? 0.0 : find_rate(ng, dg, gammas, gg_rate, g, gn);
  }
}

/* ---- triplet pair production (rt_tpp.cu:21-294) ---------------------- */

/* total_f (rt_tpp.cu:21-43) */
double orc_tpp_total_f(double s) {
  if (s < 4.0) return 0.0;
  if (s < 4.6) {
    double s4 = s - 4.0;
    return 0.001 * s4 * s4 *
           (5.6 + 20.4 * s4 - 10.9 * s4 * s4 - 3.6 * s4 * s4 * s4 + 7.4 * s4 * s4 * s4 * s4);
  }
  if (s < 6.0) return 0.582814 - 0.29842 * s + 0.04354 * s * s - 0.0012977 * s * s * s;
  if (s < 14.0)
    return (3.1247 - 1.3394 * s + 0.14612 * s * s) / (1 + 0.4648 * s + 0.016683 * s * s);
  double l2s = orc_log(2.0 * s);
  return l2s * 28.0 / 9.0 - 218.0 / 27.0 +
         (-(4.0 / 3.0) * l2s * l2s * l2s + 3.863 * l2s * l2s - 11.0 * l2s + 27.9) / s;
}

/* approx_Em (rt_tpp.cu:45-60) */
double orc_tpp_approx_em(double s) {
  if (s <= 4.01) return 0.0;
  if (s <= 100.0)
    return orc_exp((orc_log(1.337) + orc_log(5.2495 / 1.337) * orc_log(s / 4.01) /
                                         orc_log(100.0 / 4.01))) /
           s;
  double l2s = orc_log(2.0 * s);
  return 0.187 * l2s * l2s / s;
}

/* E_plus_min / E_plus_max (rt_tpp.cu:62-74): NaN below s = 4 */
void orc_tpp_e_plus(double gamma, double s, double *emin, double *emax) {
  double r = sqrt((gamma * gamma - 1.0f) * s * (s - 4.0f));
  *emin = (gamma * (s - 1.0f) - r) / (1.0f + 2.0f * s);
  *emax = (gamma * (s - 1.0f) + r) / (1.0f + 2.0f * s);
}

/* triplet_pairs::init (rt_tpp.cu:188-275): rate and Em per gamma, then the
 * n_gamma x n_gamma pair-energy CDF; the gamma grid is orc_ic_grids' */
void orc_tpp_init(int kind, const double *p, double emin, double emax, double n0, int ng,
                  const double *gammas, double *rate, double *em, double *dnde) {
  const double dmu = 2.0 / (IC_N_MU - 1.0);
  const double de = (orc_log(emax) - orc_log(emin)) / (IC_N_E - 1.0);
  for (int n = 0; n < ng; n++) {
    double gamma = gammas[n], result = 0.0, emean = 0.0;
    for (int i_mu = 0; i_mu < IC_N_MU; i_mu++) {
      double mu = i_mu * dmu - 1.0;
      for (int i_e = 0; i_e < IC_N_E; i_e++) {
        double e = orc_exp(orc_log(emin) + i_e * de);
        double b = ic_beta(gamma);
        double s = gamma * e * (1.0 - b * mu);
        double nv = orc_ic_spectrum(kind, p, e);
        result += 0.5f * nv * orc_tpp_total_f(s) * (1.0 - b * mu) * e;
        emean += 0.5f * nv * orc_tpp_total_f(s) * orc_tpp_approx_em(s) * (1.0 - b * mu) * e;
      }
    }
    rate[n] = result * dmu * de * n0 * IC_RE_SQUARE / 137.0;
    em[n] = result > 0.0 ? emean / result : 0.0;
  }
  for (int n = 0; n < ng; n++) {
    double gamma = gammas[n];
    double *row = dnde + (size_t)n * ng;
    for (int i = 0; i < ng; i++) {
      double e_p = gammas[i];
      if (e_p > gamma) {
        row[i] = 0.0;
        continue;
      }
      double result = 0.0;
      for (int i_e = 0; i_e < IC_N_E; i_e++) {
        double e = orc_exp(orc_log(emin) + i_e * de);
        double nv = orc_ic_spectrum(kind, p, e);
        double s = gamma * e, lo, hi;
        orc_tpp_e_plus(gamma, s, &lo, &hi);
        if (e_p < lo || e_p > hi) continue;
        double norm = 4.0 * (ic_pow(hi, 0.25) - ic_pow(lo, 0.25));
        result += nv * rate[n] * em[n] * ic_pow(e_p, -1.75) * e / norm;
      }
      row[i] = result * de * dmu;
    }
    for (int i = 1; i < ng; i++) row[i] += row[i - 1];
    double last = row[ng - 1];
    for (int i = 0; i < ng; i++) row[i] /= last;
  }
}

/* gen_tpp_Ep (rt_tpp.cu:132-143), u the first uniform of the Philox block
 * (seed, k, epoch, purpose 11) */
void orc_tpp_gen_ep(int ng, double dg, const double *dnde, const double *gamma, double *ep,
                    uint64_t n, uint64_t seed, uint32_t epoch) {
  for (uint64_t k = 0; k < n; k++) {
    double uu[2], l = 0.0, h = 0.0;
    orc_uniform2(seed, k, epoch, 11, uu);
    float u = (float)uu[0];
    double g = gamma[k];
    int gn = find_n_gamma(g, ng, dg);
    int b = ic_search(u, dnde + (size_t)gn * ng, ng, &l, &h);
    double bb = (l == h ? b : (u - l) / (h - l) + b);
    double result = orc_exp(orc_log(IC_MIN_GAMMA) + dg * bb);
    ep[k] = fmin(result, g - 1.01);
  }
}

/* find_tpp_rate / find_tpp_Em (rt_tpp.cu:114-130) */
void orc_tpp_rates(int ng, double dg, const double *gammas, const double *rate,
                   const double *em, const double *gamma, double *r, double *m, uint64_t n) {
  for (uint64_t k = 0; k < n; k++) {
    double g = gamma[k];
    int gn = find_n_gamma(g, ng, dg);
    r[k] = find_rate(ng, dg, gammas, rate, g, gn);
    m[k] = find_rate(ng, dg, gammas, em, g, gn);
  }
}

/* ---- analytic emission models (inverse_compton_power_law.cuh,
 * inverse_compton_dummy.cuh) ------------------------------------------- */

/* std::log of a float in the reference: this build's log, rounded to float */
static float logf_(float x) { return (float)orc_log((double)x); }

/* compute_A1 / compute_A2 (inverse_compton_power_law.cuh:11-26) */
static double pl_A1(const orc_ic_model *m, double er) {
  return 1.0f / (er * (0.5f + 1.0f / m->alpha -
                       (1.0f / (m->alpha * (m->alpha + 1.0f))) * ic_pow(er / m->es, m->alpha)));
}
static double pl_A2(const orc_ic_model *m, double er, double et) {
  return 1.0f / (et * (et * 0.5f / er + orc_log(er / et) + 1.0f / (1.0f + m->alpha)));
}

/* f_inv1 / f_inv2 (:28-63) */
double orc_pl_f_inv1(const orc_ic_model *m, float u, double gamma) {
  double er = 2.0f * gamma * m->emin;
  double A1 = pl_A1(m, er);
  if (u < A1 * er * 0.5f) return sqrt(2.0f * u * er / A1);
  if (u < 1.0f - A1 * er * ic_pow(m->es / er, -m->alpha) / (1.0f + m->alpha))
    return er * ic_pow(m->alpha * (1.0f / m->alpha + 0.5f - u / (A1 * er)), -1.0f / m->alpha);
  return er * ic_pow((1.0f - u) * (1.0f + m->alpha) / (A1 * m->es), -1.0f / (m->alpha + 1.0f));
}
double orc_pl_f_inv2(const orc_ic_model *m, float u, double gamma) {
  double er = 2.0f * gamma * m->emin;
  double et = er / (2.0f * er + 1.0f);
  double A2 = pl_A2(m, er, et);
  if (u < A2 * et * et * 0.5f / er) return sqrt(2.0f * u * er / A2);
  if (u < 1.0f - A2 * et / (1.0f + m->alpha)) return et * orc_exp(u / (A2 * et) - et * 0.5f / er);
  return er * ic_pow((1.0f - u) * (1.0f + m->alpha) / (A2 * et), -1.0f / (m->alpha + 1.0f));
}

/* draw_photon_ep (:79-121) */
double orc_pl_draw_ep(const orc_ic_model *m, double e1p, double gamma, float u) {
  double gemin2 = 2.0f * gamma * m->emin;
  double al = m->alpha;
  if (e1p < 0.5f && e1p / (1.0f - 2.0f * e1p) <= gemin2) {
    double e_lim = e1p / (1.0f - 2.0f * e1p);
    double a1 = (gemin2 * gemin2 * (al + 2.0f)) / (gamma * (e_lim * e_lim - e1p * e1p));
    return sqrt(u * (al + 2.0f) * gemin2 * gemin2 / (a1 * gamma) + e1p * e1p);
  }
  if (e1p > gemin2) {
    double a2 = (al * (al + 2.0f) * 0.5f / gamma) * ic_pow(e1p / gemin2, al);
    if (e1p < 0.5f) a2 /= (1.0f - ic_pow(1.0f - 2.0f * e1p, al));
    return gemin2 * ic_pow(ic_pow(gemin2 / e1p, al) - u * al * (al + 2.0f) / (2.0f * gamma * a2),
                           -1.0f / al);
  }
  double G = 0.0f;
  if (e1p < 0.5f) G = ic_pow((1.0f - 2.0f * e1p) * gemin2 / e1p, al);
  double U_0 = (gemin2 * gemin2 - e1p * e1p) * gamma / (gemin2 * gemin2 * (al + 2.0f));
  double a3 = 1.0f / (U_0 + (1.0f - G) * 2.0f * gamma / (al * (al + 2.0f)));
  if (u < U_0 * a3) return sqrt(u * (al + 2.0f) * gemin2 * gemin2 / (a3 * gamma) + e1p * e1p);
  return gemin2 * ic_pow(1.0f - (u - a3 * U_0) * al * (al + 2.0) / (2.0f * a3 * gamma), -1.0f / al);
}

/* emit_photon, draw_photon_energy, draw_photon_freepath (:149-182) and the
 * dummy (inverse_compton_dummy.cuh:19-35) over n particles; the draws of
 * particle k are the uniforms of Philox (seed, k, epoch, 12) and (..., 13) */
void orc_ic_model_emit(const orc_ic_model *m, const double *gamma, const double *p,
                       uint8_t *emit, double *eph, double *path, uint64_t n, uint64_t seed,
                       uint32_t epoch) {
  for (uint64_t k = 0; k < n; k++) {
    double a[2], b[2];
    orc_uniform2(seed, k, epoch, 12, a);
    orc_uniform2(seed, k, epoch, 13, b);
    float u[4] = {(float)a[0], (float)a[1], (float)b[0], (float)b[1]};
    double g = gamma[k];
    emit[k] = 0;
    eph[k] = 0.0;
    path[k] = 0.0;
    if (m->kind == 0) {
      if (g > m->cutoff) {
        emit[k] = 1;
        eph[k] = m->eph;
        path[k] = m->mfp;
      }
      continue;
    }
    float e_p = (float)(g * m->emin);
    float rate = (float)(e_p < 0.1f ? m->icrate : m->icrate * logf_(2.0f * e_p + 0.90517f) / e_p);
    if (!(u[0] < rate)) continue;
    emit[k] = 1;
    double e1p = g < m->es * 0.5f / m->emin ? orc_pl_f_inv1(m, u[1], g) : orc_pl_f_inv2(m, u[1], g);
    double ep = orc_pl_draw_ep(m, e1p, g, u[2]);
    double u1p = 1.0f - 1.0f / e1p + 1.0f / ep;
    double sg = (double)((0.0 < p[k]) - (p[k] < 0.0));
    double E = sg * (g + fabs(p[k]) * (-u1p)) * e1p;
    double r;
    if (E * m->emin < 2.0f)
      r = ic_pow(fabs(E) * m->emin / 2.0f, m->alpha);
    else
      r = 2.0f / (fabs(E) * m->emin);
    eph[k] = E;
    path[k] = -m->mfp * logf_(u[3]) / r;
  }
}
