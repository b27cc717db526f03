// kern_ic.hip -- inverse-Compton scattering tables and photon-energy
// sampling (the reference's inverse_compton class,
// src/cuda/radiation/rt_ic.cu:22-547, with the soft-photon spectra of
// src/radiation/spectra.h), and the triplet pair-production tables of
// triplet_pairs (src/cuda/radiation/rt_tpp.cu:21-294) on the same grids.
//
// The reference builds its tables with host loops (rt_ic.cu:292-421, the
// device kernel is commented out); here every table entry is one GPU thread
// summing its 800 (x 100) quadrature terms in the reference's order, and a
// second kernel takes the cumulative sums and normalises each row, one
// thread per row.  exp/log are the build's own (apt_math.hpp) so the CPU
// restatement reproduces every table bit for bit; pow(x, y) is
// exp(y ln x).  Sampling draws u from the Philox stream instead of curand.
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/aperture_amd.h"
#include "apt_common.hpp"
#include "apt_math.hpp"

namespace apt {
namespace {

// rt_ic.cu:15-20
constexpr double IC_MIN_GAMMA = 1.01f;  // a float literal in the reference
constexpr double IC_MAX_GAMMA = 1.0e15;
constexpr double IC_MIN_EP = 1.0e-10;
constexpr int IC_N_MU = 100, IC_N_E = 800;  // rt_ic.cu:283-284
constexpr double RE_SQUARE = 7.91402e-26;    // constant_defs.h:36
constexpr double CONST_PI = 3.14159265358979323846;

__host__ __device__ inline double ic_pow(double x, double y) { return apt_exp(y * apt_log(x)); }
__host__ __device__ inline double sqr(double x) { return x * x; }

// spectra.h:10-95
struct Spectrum {
  int kind;
  double p[5];
  __host__ __device__ double operator()(double e) const {
    switch (kind) {
      case APT_SPEC_POWER_LAW_HARD:  // delta, emin, emax
        return (e < p[2] && e > p[1]) ? ic_pow(e / p[2], p[0]) / e : 0.0;
      case APT_SPEC_POWER_LAW_SOFT:  // alpha, emin, emax
        return (e < p[2] && e > p[1]) ? ic_pow(e / p[1], -p[0]) / e : 0.0;
      case APT_SPEC_BROKEN_POWER_LAW:  // alpha, delta, epeak, emin, emax
        if (e < p[2] && e > p[3]) return ic_pow(e / p[2], p[1]) / e;
        if (e > p[2] && e < p[4]) return ic_pow(e / p[2], -p[0]) / e;
        return 0.0;
      case APT_SPEC_BLACK_BODY:  // kT
        return 1.75464e30 * e * e / (apt_exp(e / p[0]) - 1.0);
      default:  // APT_SPEC_MONO_ENERGETIC: e0, de
        return (e < p[0] + p[1] && e > p[0] - p[1]) ? 1.0 / (2.0 * p[1]) : 0.0;
    }
  }
};

// rt_ic.cu:22-60 (beta() returns 1/sqrt(1 - 1/gamma^2), as in the reference)
__host__ __device__ inline double ic_beta(double gamma) {
  return 1.0 / sqrt(1.0 - 1.0 / sqr(gamma));
}
__host__ __device__ inline double sigma_ic(double x) {
  if (x < 1.0e-3) return 1.0f - 2.0f * x + 26.0f * x * x / 5.0f;
  const double l = apt_log(1.0 + 2.0 * x);
  return 0.75 * ((1.0 + x) * (2.0 * x * (1.0 + x) / (1.0 + 2.0 * x) - l) / (x * x * x) +
                 0.5 * l / x - (1.0 + 3.0 * x) / sqr(1.0 + 2.0 * x));
}
__host__ __device__ inline double sigma_gg(double b) {
  return (1.0 - sqr(b)) *
         ((3.0 - b * b * b * b) * apt_log((1.0 + b) / (1.0 - b)) - 2.0 * b * (2.0 - b * b));
}
__host__ __device__ inline double x_ic(double gamma, double e, double mu) {
  return gamma * e * (1.0 - mu * ic_beta(gamma));
}
__host__ __device__ inline double sigma_lab(double q, double ge) {
  return 2.0 * q * apt_log(q) + (1.0 + 2.0 * q) * (1.0 - q) +
         0.5 * sqr(ge * q) * (1.0 - q) / (1.0 + q * ge);
}

struct Grid {
  double emin, de, dmu, n0;
};

// one IC rate per gamma (rt_ic.cu:288-333)
__global__ void k_ic_rate(Spectrum ne, Grid g, const double *gammas, int ng, double *rate) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= ng) return;
  const double gamma = gammas[n];
  double result = 0.0;
  for (int i_mu = 0; i_mu < IC_N_MU; i_mu++) {
    const double mu = i_mu * g.dmu - 1.0;
    for (int i_e = 0; i_e < IC_N_E; i_e++) {
      const double e = apt_exp(apt_log(g.emin) + i_e * g.de);
      const double x = x_ic(gamma, e, mu);
      const double sigma = sigma_ic(x);
      result += 0.5f * ne(e) * sigma * (1.0f - ic_beta(gamma) * mu) * e;
    }
  }
  // the reference's left-to-right product (rt_ic.cu:323-324)
  rate[n] = result * g.dmu * g.de * (RE_SQUARE * g.n0) * 8.0 * CONST_PI / 3.0;
}

// one gamma-gamma rate per photon energy (rt_ic.cu:337-368)
__global__ void k_gg_rate(Spectrum ne, Grid g, const double *gammas, int ng, double *rate) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= ng) return;
  const double eph = gammas[n];
  if (eph < 2.0) {
    rate[n] = 0.0;
    return;
  }
  double result = 0.0;
  for (int i_mu = 0; i_mu < IC_N_MU; i_mu++) {
    const double mu = i_mu * g.dmu - 1.0;
    for (int i_e = 0; i_e < IC_N_E; i_e++) {
      const double e = apt_exp(apt_log(g.emin) + i_e * g.de);
      const double s = eph * e * (1.0 - mu) * 0.5;
      if (s <= 1.0) continue;
      const double b = sqrt(1.0 - 1.0 / s);
      if (b == 1.0) continue;
      result += sigma_gg(b) * (1.0 - mu) * ne(e) * e;
    }
  }
  rate[n] = 0.25 * result * g.dmu * g.de * CONST_PI * (g.n0 * RE_SQUARE);
}

// lab-frame spectrum dN/de1 at (gamma n, e1 i) before the cumulative sum
// (rt_ic.cu:371-392 and, times e1, :398-412)
__global__ void k_ic_dnde(Spectrum ne, Grid g, const double *gammas, int ng,
                          const double *e1s, int nep, int times_e1, double *table) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ng * nep) return;
  const int n = t / nep, i = t % nep;
  const double gamma = gammas[n];
  const double e1 = e1s[i];
  double result = 0.0;
  for (int i_e = 0; i_e < IC_N_E; i_e++) {
    const double e = apt_exp(apt_log(g.emin) + i_e * g.de);
    const double nv = ne(e);
    const double ge = gamma * e * 4.0;
    const double q = e1 / (ge * (1.0 - e1));
    if (e1 < ge / (1.0 + ge) && e1 > e / gamma) result += nv * sigma_lab(q, ge) / gamma;
  }
  table[(size_t)n * nep + i] = times_e1 ? result * g.de * e1 : result * g.de;
}

// cumulative sum along e1 and normalisation by the last entry
__global__ void k_ic_cdf(double *table, int ng, int nep) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= ng) return;
  double *row = table + (size_t)n * nep;
  for (int i = 1; i < nep; i++) row[i] += row[i - 1];
  const double last = row[nep - 1];
  for (int i = 0; i < nep; i++) row[i] /= last;
}

// binary_search (rt_ic.cu:88-116) on one table row
__device__ int ic_search(float u, const double *a, int size, double &l, double &h) {
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
      l = h = v;
      return hi;
    }
  }
  if (v < u) {
    l = v;
    h = (lo < size ? a[lo] : v);
    return tmp;
  }
  h = v;
  l = (hi >= 0 ? a[hi] : v);
  return hi > 0 ? hi : 0;
}

struct Tables {
  const double *gammas, *ic_rate, *gg_rate, *dnde, *dnde_th;
  int ng, nep;
  double dg, dep, dlep;
};

// find_n_gamma (rt_ic.cu:74-79); the clamp only acts on a NaN gamma
__device__ int ic_find_n_gamma(const Tables &t, double gamma) {
  if (gamma < IC_MIN_GAMMA) return 0;
  if (gamma > IC_MAX_GAMMA) return t.ng - 1;
  const int gn = (int)((apt_log(gamma) - apt_log(IC_MIN_GAMMA)) / t.dg);
  return gn < 0 ? 0 : (gn > t.ng - 1 ? t.ng - 1 : gn);
}

// find_ic_rate (rt_ic.cu:156-162); at the last grid point (gamma = 1e15
// exactly) the upper neighbour is the point itself instead of one past the
// table
__device__ double ic_find_rate(const Tables &t, const double *rate, double g, int gn) {
  const int gu = gn + 1 < t.ng ? gn + 1 : gn;
  if (g > IC_MAX_GAMMA) return rate[gn];
  const double x = (apt_log(g) - apt_log(t.gammas[gn])) / t.dg;
  return rate[gn] * (1.0f - x) + rate[gu] * x;
}

// gen_photon_e (rt_ic.cu:127-152) for the draw u
__device__ double ic_gen_photon_e(const Tables &t, double g, float u) {
  const int gn = ic_find_n_gamma(t, g);
  double l = 0.0, h = 0.0, bb, result;
  int b = ic_search(u, t.dnde + (size_t)gn * t.nep, t.nep, l, h);
  if (b < 2) {
    b = ic_search(u, t.dnde_th + (size_t)gn * t.nep, t.nep, l, h);
    bb = (l == h ? b : (u - l) / (h - l) + b);
    result = apt_exp(t.dlep * bb + apt_log(IC_MIN_EP));
  } else {
    bb = (l == h ? b : (u - l) / (h - l) + b);
    result = t.dep * bb;
  }
  const double v = result * g;
  return fmax(fmin(v, g - 1.01), 0.0);
}

// accept == 0: eph = gen_photon_e(gamma) with u the first uniform of the
// Philox block (seed, k, epoch, 9).  accept == 1: gen_photon_energies
// (rt_ic.cu:198-212): the first uniform accepts a scattering when below
// find_ic_rate(gamma), the second draws its energy; eph = 0 otherwise.
__global__ void k_ic_sample(Tables t, const double *gamma, double *eph, uint64_t n,
                            uint64_t seed, uint32_t epoch, int accept) {
  const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  double u0, u1;
  uniform2(seed, k, epoch, 9, u0, u1);
  const double g = gamma[k];
  if (!accept) {
    eph[k] = ic_gen_photon_e(t, g, (float)u0);
  } else {
    const int gn = ic_find_n_gamma(t, g);
    eph[k] = (float)u0 < ic_find_rate(t, t.ic_rate, g, gn) ? ic_gen_photon_e(t, g, (float)u1)
                                                           : 0.0;
  }
}

// gen_rand_gammas (rt_ic.cu:214-225): log-uniform in [1.01, 1e10] with the
// first uniform of the Philox block (seed, k, epoch, 10)
__global__ void k_ic_rand_gammas(double *gammas, uint64_t n, uint64_t seed, uint32_t epoch) {
  const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  double u0, u1;
  uniform2(seed, k, epoch, 10, u0, u1);
  const float u = (float)u0;
  gammas[k] = apt_exp(apt_log(1.01) + u * (apt_log(1e10) - apt_log(1.01)));
}

// find_ic_rate / find_gg_rate (rt_ic.cu:156-171)
__global__ void k_ic_rates(Tables t, const double *gamma, double *ic, double *gg, uint64_t n) {
  const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const double g = gamma[k];
  const int gn = ic_find_n_gamma(t, g);
  ic[k] = ic_find_rate(t, t.ic_rate, g, gn);
  gg[k] = g < 2.0 ? 0.0 : ic_find_rate(t, t.gg_rate, g, gn);
}

// ---- triplet pair production (rt_tpp.cu:21-294) ----------------------

// total_f (rt_tpp.cu:21-43): the triplet cross section over alpha_f r_e^2
__host__ __device__ inline double tpp_total_f(double s) {
  if (s < 4.0) return 0.0;
  if (s < 4.6) {
    const double s4 = s - 4.0;
    return 0.001 * s4 * s4 *
           (5.6 + 20.4 * s4 - 10.9 * s4 * s4 - 3.6 * s4 * s4 * s4 + 7.4 * s4 * s4 * s4 * s4);
  }
  if (s < 6.0) return 0.582814 - 0.29842 * s + 0.04354 * s * s - 0.0012977 * s * s * s;
  if (s < 14.0)
    return (3.1247 - 1.3394 * s + 0.14612 * s * s) / (1 + 0.4648 * s + 0.016683 * s * s);
  const double l2s = apt_log(2.0 * s);
  return l2s * 28.0 / 9.0 - 218.0 / 27.0 +
         (-(4.0 / 3.0) * l2s * l2s * l2s + 3.863 * l2s * l2s - 11.0 * l2s + 27.9) / s;
}

// approx_Em (rt_tpp.cu:45-60): mean pair energy fraction
__host__ __device__ inline double tpp_approx_em(double s) {
  if (s <= 4.01) return 0.0;
  if (s <= 100.0)
    return apt_exp((apt_log(1.337) + apt_log(5.2495 / 1.337) * apt_log(s / 4.01) /
                                         apt_log(100.0 / 4.01))) /
           s;
  const double l2s = apt_log(2.0 * s);
  return 0.187 * l2s * l2s / s;
}

// E_plus_min / E_plus_max (rt_tpp.cu:62-74); NaN below s = 4
__host__ __device__ inline double tpp_e_plus_min(double gamma, double s) {
  return (gamma * (s - 1.0f) - sqrt((gamma * gamma - 1.0f) * s * (s - 4.0f))) / (1.0f + 2.0f * s);
}
__host__ __device__ inline double tpp_e_plus_max(double gamma, double s) {
  return (gamma * (s - 1.0f) + sqrt((gamma * gamma - 1.0f) * s * (s - 4.0f))) / (1.0f + 2.0f * s);
}

// rate and mean pair energy per gamma (rt_tpp.cu:200-229)
__global__ void k_tpp_rate(Spectrum ne, Grid g, const double *gammas, int ng, double *rate,
                           double *em) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= ng) return;
  const double gamma = gammas[n];
  double result = 0.0, emean = 0.0;
  for (int i_mu = 0; i_mu < IC_N_MU; i_mu++) {
    const double mu = i_mu * g.dmu - 1.0;
    for (int i_e = 0; i_e < IC_N_E; i_e++) {
      const double e = apt_exp(apt_log(g.emin) + i_e * g.de);
      const double b = ic_beta(gamma);
      const double s = gamma * e * (1.0 - b * mu);
      result += 0.5f * ne(e) * tpp_total_f(s) * (1.0 - b * mu) * e;
      emean += 0.5f * ne(e) * tpp_total_f(s) * tpp_approx_em(s) * (1.0 - b * mu) * e;
    }
  }
  rate[n] = result * g.dmu * g.de * g.n0 * RE_SQUARE / 137.0;
  em[n] = result > 0.0 ? emean / result : 0.0;
}

// pair-energy spectrum at (gamma n, E_p = gammas[i]) before the cumulative
// sum (rt_tpp.cu:233-262)
__global__ void k_tpp_dnde(Spectrum ne, Grid g, const double *gammas, int ng, const double *rate,
                           const double *em, double *table) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ng * ng) return;
  const int n = t / ng, i = t % ng;
  const double gamma = gammas[n];
  const double e_p = gammas[i];
  double *out = table + (size_t)n * ng + i;
  if (e_p > gamma) {
    *out = 0.0;
    return;
  }
  double result = 0.0;
  for (int i_e = 0; i_e < IC_N_E; i_e++) {
    const double e = apt_exp(apt_log(g.emin) + i_e * g.de);
    const double nv = ne(e);
    const double s = gamma * e;
    const double emin = tpp_e_plus_min(gamma, s);
    const double emax = tpp_e_plus_max(gamma, s);
    if (e_p < emin || e_p > emax) continue;
    const double norm = 4.0 * (ic_pow(emax, 0.25) - ic_pow(emin, 0.25));
    result += nv * rate[n] * em[n] * ic_pow(e_p, -1.75) * e / norm;
  }
  *out = result * g.de * g.dmu;
}

struct TppTables {
  const double *gammas, *rate, *em, *dnde;
  int ng;
  double dg;
};

__device__ int tpp_find_n_gamma(const TppTables &t, double gamma) {
  if (gamma < IC_MIN_GAMMA) return 0;
  if (gamma > IC_MAX_GAMMA) return t.ng - 1;
  const int gn = (int)((apt_log(gamma) - apt_log(IC_MIN_GAMMA)) / t.dg);
  return gn < 0 ? 0 : (gn > t.ng - 1 ? t.ng - 1 : gn);
}

// gen_tpp_Ep (rt_tpp.cu:132-143) with u the first uniform of the Philox
// block (seed, k, epoch, 11)
__global__ void k_tpp_sample(TppTables t, const double *gamma, double *ep, uint64_t n,
                             uint64_t seed, uint32_t epoch) {
  const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  double u0, u1;
  uniform2(seed, k, epoch, 11, u0, u1);
  const float u = (float)u0;
  const double g = gamma[k];
  const int gn = tpp_find_n_gamma(t, g);
  double l = 0.0, h = 0.0;
  const int b = ic_search(u, t.dnde + (size_t)gn * t.ng, t.ng, l, h);
  const double bb = (l == h ? b : (u - l) / (h - l) + b);
  const double result = apt_exp(apt_log(IC_MIN_GAMMA) + t.dg * bb);
  ep[k] = fmin(result, g - 1.01);
}

// find_tpp_rate / find_tpp_Em (rt_tpp.cu:114-130)
__global__ void k_tpp_rates(TppTables t, const double *gamma, double *rate, double *em,
                            uint64_t n) {
  const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const double g = gamma[k];
  const int gn = tpp_find_n_gamma(t, g);
  const int gu = gn + 1 < t.ng ? gn + 1 : gn;
  if (g > IC_MAX_GAMMA) {
    rate[k] = t.rate[gn];
    em[k] = t.em[gn];
  } else {
    const double x = (apt_log(g) - apt_log(t.gammas[gn])) / t.dg;
    rate[k] = t.rate[gn] * (1.0f - x) + t.rate[gu] * x;
    em[k] = t.em[gn] * (1.0f - x) + t.em[gu] * x;
  }
}

// ---- analytic emission models (inverse_compton_power_law.cuh,
// inverse_compton_dummy.cuh) ---------------------------------------------

struct IcModel {
  int kind;
  double alpha, es, emin, mfp, icrate, cutoff, eph;
};

// std::log of a float in the reference: the build's log, rounded to float
__host__ __device__ inline float logf_(float x) { return (float)apt_log((double)x); }

// InverseComptonPL1D::compute_A1 / compute_A2 (inverse_compton_power_law.cuh:11-26)
__host__ __device__ inline double pl_A1(const IcModel &m, double er) {
  return 1.0f / (er * (0.5f + 1.0f / m.alpha -
                       (1.0f / (m.alpha * (m.alpha + 1.0f))) * ic_pow(er / m.es, m.alpha)));
}
__host__ __device__ inline double pl_A2(const IcModel &m, double er, double et) {
  return 1.0f / (et * (et * 0.5f / er + apt_log(er / et) + 1.0f / (1.0f + m.alpha)));
}

// f_inv1 / f_inv2 (:28-63): inverse CDFs of the rest-frame photon energy
__host__ __device__ inline double pl_f_inv1(const IcModel &m, float u, double gamma) {
  const double er = 2.0f * gamma * m.emin;
  const double A1 = pl_A1(m, er);
  if (u < A1 * er * 0.5f) return sqrt(2.0f * u * er / A1);
  if (u < 1.0f - A1 * er * ic_pow(m.es / er, -m.alpha) / (1.0f + m.alpha))
    return er * ic_pow(m.alpha * (1.0f / m.alpha + 0.5f - u / (A1 * er)), -1.0f / m.alpha);
  return er * ic_pow((1.0f - u) * (1.0f + m.alpha) / (A1 * m.es), -1.0f / (m.alpha + 1.0f));
}
__host__ __device__ inline double pl_f_inv2(const IcModel &m, float u, double gamma) {
  const double er = 2.0f * gamma * m.emin;
  const double et = er / (2.0f * er + 1.0f);
  const double A2 = pl_A2(m, er, et);
  if (u < A2 * et * et * 0.5f / er) return sqrt(2.0f * u * er / A2);
  if (u < 1.0f - A2 * et / (1.0f + m.alpha)) return et * apt_exp(u / (A2 * et) - et * 0.5f / er);
  return er * ic_pow((1.0f - u) * (1.0f + m.alpha) / (A2 * et), -1.0f / (m.alpha + 1.0f));
}

// draw_photon_ep (:79-121): the original energy given the rest-frame one
__host__ __device__ inline double pl_draw_ep(const IcModel &m, double e1p, double gamma,
                                             float u) {
  const double gemin2 = 2.0f * gamma * m.emin;
  const double al = m.alpha;
  if (e1p < 0.5f && e1p / (1.0f - 2.0f * e1p) <= gemin2) {
    const double e_lim = e1p / (1.0f - 2.0f * e1p);
    const double a1 = (gemin2 * gemin2 * (al + 2.0f)) / (gamma * (e_lim * e_lim - e1p * e1p));
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
  const double U_0 = (gemin2 * gemin2 - e1p * e1p) * gamma / (gemin2 * gemin2 * (al + 2.0f));
  const double a3 = 1.0f / (U_0 + (1.0f - G) * 2.0f * gamma / (al * (al + 2.0f)));
  if (u < U_0 * a3) return sqrt(u * (al + 2.0f) * gemin2 * gemin2 / (a3 * gamma) + e1p * e1p);
  return gemin2 *
         ic_pow(1.0f - (u - a3 * U_0) * al * (al + 2.0) / (2.0f * a3 * gamma), -1.0f / al);
}

// emit_photon, draw_photon_energy, draw_photon_freepath (:149-182) for
// one particle; u[0..3] are the four draws in the reference's call order
__host__ __device__ inline void ic_model_one(const IcModel &m, double gamma, double p,
                                             const float u[4], uint8_t &emit, double &eph,
                                             double &path) {
  emit = 0;
  eph = 0.0;
  path = 0.0;
  if (m.kind == APT_ICM_DUMMY) {  // inverse_compton_dummy.cuh:19-35
    if (gamma > m.cutoff) {
      emit = 1;
      eph = m.eph;
      path = m.mfp;
    }
    return;
  }
  const float e_p = (float)(gamma * m.emin);
  const float rate = (float)(e_p < 0.1f ? m.icrate
                                        : m.icrate * logf_(2.0f * e_p + 0.90517f) / e_p);
  if (!(u[0] < rate)) return;
  emit = 1;
  const double e1p = gamma < m.es * 0.5f / m.emin ? pl_f_inv1(m, u[1], gamma)
                                                   : pl_f_inv2(m, u[1], gamma);
  const double ep = pl_draw_ep(m, e1p, gamma, u[2]);
  const double u1p = 1.0f - 1.0f / e1p + 1.0f / ep;
  const double sg = (double)((0.0 < p) - (p < 0.0));
  eph = sg * (gamma + fabs(p) * (-u1p)) * e1p;
  double r;
  if (eph * m.emin < 2.0f)
    r = ic_pow(fabs(eph) * m.emin / 2.0f, m.alpha);
  else
    r = 2.0f / (fabs(eph) * m.emin);
  path = -m.mfp * logf_(u[3]) / r;
}

// one thread per particle: the draws are the four uniforms of the Philox
// blocks (seed, k, epoch, 12) and (seed, k, epoch, 13)
__global__ void k_ic_model(IcModel m, const double *gamma, const double *p, uint8_t *emit,
                           double *eph, double *path, uint64_t n, uint64_t seed,
                           uint32_t epoch) {
  const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  double a0, a1, b0, b1;
  uniform2(seed, k, epoch, 12, a0, a1);
  uniform2(seed, k, epoch, 13, b0, b1);
  const float u[4] = {(float)a0, (float)a1, (float)b0, (float)b1};
  uint8_t e;
  double E, L;
  ic_model_one(m, gamma[k], p[k], u, e, E, L);
  emit[k] = e;
  eph[k] = E;
  path[k] = L;
}

}  // namespace
}  // namespace apt

using namespace apt;

struct apt_ic {
  hipStream_t stream = nullptr;
  int ng = 0, nep = 0;
  double dg = 0, dep = 0, dlep = 0;
  std::vector<double> gammas, ep, log_ep;
  double *d_gammas = nullptr, *d_ep = nullptr, *d_log_ep = nullptr;
  double *d_ic = nullptr, *d_gg = nullptr, *d_dnde = nullptr, *d_dnde_th = nullptr;
  bool ready = false;
  std::string err;
  apt_status fail(apt_status st, const char *msg) {
    err = msg;
    return st;
  }
};

struct apt_tpp {
  hipStream_t stream = nullptr;
  int ng = 0;
  double dg = 0;
  std::vector<double> gammas;
  double *d_gammas = nullptr, *d_rate = nullptr, *d_em = nullptr, *d_dnde = nullptr;
  bool ready = false;
  std::string err;
  apt_status fail(apt_status st, const char *msg) {
    err = msg;
    return st;
  }
};

#define ICCHK(x)                                               \
  do {                                                         \
    hipError_t e_ = (x);                                       \
    if (e_ != hipSuccess) return h->fail(APT_ERR_HIP, hipGetErrorString(e_)); \
  } while (0)

extern "C" {

// inverse_compton::inverse_compton (rt_ic.cu:227-282): the gamma, e1 and
// log e1 grids
apt_status apt_ic_create(int n_gamma, int n_ep, apt_ic **out) {
  // each cumulative table is n_gamma x n_ep doubles, one thread per entry
  if (!out || n_gamma < 3 || n_ep < 3 || (uint64_t)n_gamma * (uint64_t)n_ep > (1ull << 27))
    return APT_ERR_INVALID_ARG;
  apt_ic *h = new apt_ic();
  h->ng = n_gamma;
  h->nep = n_ep;
  h->dep = 1.0 / ((double)n_ep - 1.0);
  h->ep.resize(n_ep);
  h->log_ep.resize(n_ep);
  for (int n = 0; n < n_ep; n++) h->ep[n] = h->dep * n;
  h->dlep = (-apt_log(IC_MIN_EP)) / ((double)n_ep - 1.0);
  for (int n = 0; n < n_ep; n++) h->log_ep[n] = apt_exp(h->dlep * n + apt_log(IC_MIN_EP));
  h->dg = (apt_log(IC_MAX_GAMMA) - apt_log(IC_MIN_GAMMA)) / ((double)n_gamma - 1.0);
  h->gammas.resize(n_gamma);
  for (int n = 0; n < n_gamma; n++)
    h->gammas[n] = apt_exp(apt_log(IC_MIN_GAMMA) + h->dg * (double)n);
  hipError_t e = hipStreamCreate(&h->stream);
  const size_t tb = (size_t)n_gamma * n_ep * 8;
  if (e == hipSuccess) e = hipMalloc(&h->d_gammas, n_gamma * 8);
  if (e == hipSuccess) e = hipMalloc(&h->d_ep, n_ep * 8);
  if (e == hipSuccess) e = hipMalloc(&h->d_log_ep, n_ep * 8);
  if (e == hipSuccess) e = hipMalloc(&h->d_ic, n_gamma * 8);
  if (e == hipSuccess) e = hipMalloc(&h->d_gg, n_gamma * 8);
  if (e == hipSuccess) e = hipMalloc(&h->d_dnde, tb);
  if (e == hipSuccess) e = hipMalloc(&h->d_dnde_th, tb);
  if (e == hipSuccess)
    e = hipMemcpy(h->d_gammas, h->gammas.data(), n_gamma * 8, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(h->d_ep, h->ep.data(), n_ep * 8, hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemcpy(h->d_log_ep, h->log_ep.data(), n_ep * 8, hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    apt_ic_destroy(h);
    return APT_ERR_HIP;
  }
  *out = h;
  return APT_OK;
}

apt_status apt_ic_destroy(apt_ic *h) {
  if (!h) return APT_ERR_INVALID_ARG;
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  for (double *p : {h->d_gammas, h->d_ep, h->d_log_ep, h->d_ic, h->d_gg, h->d_dnde,
                    h->d_dnde_th})
    (void)hipFree(p);
  if (h->stream) (void)hipStreamDestroy(h->stream);
  delete h;
  return APT_OK;
}

const char *apt_ic_last_error(const apt_ic *h) { return h ? h->err.c_str() : "null handle"; }

// inverse_compton::init (rt_ic.cu:279-421)
apt_status apt_ic_init(apt_ic *h, const apt_ic_spectrum *spec, double emin, double emax,
                       double n0) {
  if (!h) return APT_ERR_INVALID_ARG;
  if (!spec || spec->kind < 0 || spec->kind > APT_SPEC_MONO_ENERGETIC || !(emin > 0.0) ||
      !(emax > emin))
    return h->fail(APT_ERR_INVALID_ARG, "ic_init: bad spectrum or energy range");
  Spectrum ne;
  ne.kind = spec->kind;
  for (int k = 0; k < 5; k++) ne.p[k] = spec->p[k];
  Grid g;
  g.emin = emin;
  g.n0 = n0;
  g.dmu = 2.0 / (IC_N_MU - 1.0);
  g.de = (apt_log(emax) - apt_log(emin)) / (IC_N_E - 1.0);
  const int ng = h->ng, nep = h->nep;
  k_ic_rate<<<(ng + 63) / 64, 64, 0, h->stream>>>(ne, g, h->d_gammas, ng, h->d_ic);
  k_gg_rate<<<(ng + 63) / 64, 64, 0, h->stream>>>(ne, g, h->d_gammas, ng, h->d_gg);
  const int cells = ng * nep;
  k_ic_dnde<<<(cells + 255) / 256, 256, 0, h->stream>>>(ne, g, h->d_gammas, ng, h->d_ep, nep,
                                                        0, h->d_dnde);
  k_ic_dnde<<<(cells + 255) / 256, 256, 0, h->stream>>>(ne, g, h->d_gammas, ng, h->d_log_ep,
                                                        nep, 1, h->d_dnde_th);
  k_ic_cdf<<<(ng + 63) / 64, 64, 0, h->stream>>>(h->d_dnde, ng, nep);
  k_ic_cdf<<<(ng + 63) / 64, 64, 0, h->stream>>>(h->d_dnde_th, ng, nep);
  ICCHK(hipGetLastError());
  ICCHK(hipStreamSynchronize(h->stream));
  h->ready = true;
  return APT_OK;
}

apt_status apt_ic_get_table(apt_ic *h, int which, double *dst, uint64_t count) {
  if (!h) return APT_ERR_INVALID_ARG;
  const uint64_t ng = h->ng, nep = h->nep;
  const double *src = nullptr;
  uint64_t n = 0;
  switch (which) {
    case APT_IC_GAMMAS: src = h->d_gammas; n = ng; break;
    case APT_IC_EP: src = h->d_ep; n = nep; break;
    case APT_IC_LOG_EP: src = h->d_log_ep; n = nep; break;
    case APT_IC_RATE: src = h->d_ic; n = ng; break;
    case APT_IC_GG_RATE: src = h->d_gg; n = ng; break;
    case APT_IC_DNDE: src = h->d_dnde; n = ng * nep; break;
    case APT_IC_DNDE_THOMSON: src = h->d_dnde_th; n = ng * nep; break;
    default: return h->fail(APT_ERR_INVALID_ARG, "ic_get_table: unknown table");
  }
  if (!dst || count != n) return h->fail(APT_ERR_INVALID_ARG, "ic_get_table: bad size");
  if (which >= APT_IC_RATE && !h->ready)
    return h->fail(APT_ERR_STATE, "ic_get_table: apt_ic_init has not run");
  ICCHK(hipStreamSynchronize(h->stream));
  ICCHK(hipMemcpy(dst, src, n * 8, hipMemcpyDeviceToHost));
  return APT_OK;
}

static Tables tables_of(const apt_ic *h) {
  Tables t;
  t.gammas = h->d_gammas;
  t.ic_rate = h->d_ic;
  t.gg_rate = h->d_gg;
  t.dnde = h->d_dnde;
  t.dnde_th = h->d_dnde_th;
  t.ng = h->ng;
  t.nep = h->nep;
  t.dg = h->dg;
  t.dep = h->dep;
  t.dlep = h->dlep;
  return t;
}

// one thread per sample: at most 2^31 - 1 blocks of 256
static constexpr uint64_t IC_MAX_SAMPLES = ((uint64_t)1 << 31) * 255;

static apt_status ic_sample_dev(apt_ic *h, const double *gamma, double *eph, uint64_t n,
                                uint64_t seed, uint32_t epoch, void *stream, int accept) {
  if (!h) return APT_ERR_INVALID_ARG;
  if (!h->ready) return h->fail(APT_ERR_STATE, "ic sampling: apt_ic_init has not run");
  if (n == 0) return APT_OK;
  if (!gamma || !eph) return h->fail(APT_ERR_INVALID_ARG, "ic sampling: null arrays");
  if (n > IC_MAX_SAMPLES) return h->fail(APT_ERR_INVALID_ARG, "ic sampling: too many samples");
  hipStream_t st = stream ? (hipStream_t)stream : h->stream;
  k_ic_sample<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(tables_of(h), gamma, eph, n, seed,
                                                           epoch, accept);
  ICCHK(hipGetLastError());
  return APT_OK;
}

// the device arrays variant: gamma and eph are device pointers, ordered on
// `stream` (NULL: the handle's stream)
apt_status apt_ic_gen_photon_e_dev(apt_ic *h, const double *gamma, double *eph, uint64_t n,
                                   uint64_t seed, uint32_t epoch, void *stream) {
  return ic_sample_dev(h, gamma, eph, n, seed, epoch, stream, 0);
}

// inverse_compton::generate_photon_energies (rt_ic.cu:515-522)
apt_status apt_ic_generate_photon_energies(apt_ic *h, double *e_ph, const double *gammas,
                                           uint64_t n, uint64_t seed, uint32_t epoch,
                                           void *stream) {
  return ic_sample_dev(h, gammas, e_ph, n, seed, epoch, stream, 1);
}

// inverse_compton::generate_random_gamma (rt_ic.cu:524-529)
apt_status apt_ic_generate_random_gamma(apt_ic *h, double *gammas, uint64_t n, uint64_t seed,
                                        uint32_t epoch, void *stream) {
  if (!h) return APT_ERR_INVALID_ARG;
  if (n == 0) return APT_OK;
  if (!gammas) return h->fail(APT_ERR_INVALID_ARG, "ic_generate_random_gamma: null array");
  if (n > IC_MAX_SAMPLES)
    return h->fail(APT_ERR_INVALID_ARG, "ic_generate_random_gamma: too many samples");
  hipStream_t st = stream ? (hipStream_t)stream : h->stream;
  k_ic_rand_gammas<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(gammas, n, seed, epoch);
  ICCHK(hipGetLastError());
  return APT_OK;
}

apt_status apt_ic_gen_photon_e(apt_ic *h, const double *gamma, double *eph, uint64_t n,
                               uint64_t seed, uint32_t epoch) {
  if (!h) return APT_ERR_INVALID_ARG;
  if (n == 0) return APT_OK;
  if (!gamma || !eph) return h->fail(APT_ERR_INVALID_ARG, "ic_gen_photon_e: null arrays");
  if (!h->ready) return h->fail(APT_ERR_STATE, "ic sampling: apt_ic_init has not run");
  if (n > IC_MAX_SAMPLES) return h->fail(APT_ERR_INVALID_ARG, "ic sampling: too many samples");
  double *dg = nullptr, *de = nullptr;
  hipError_t e = hipMalloc(&dg, n * 8);
  if (e == hipSuccess) e = hipMalloc(&de, n * 8);
  if (e == hipSuccess) e = hipMemcpy(dg, gamma, n * 8, hipMemcpyHostToDevice);
  apt_status st = e == hipSuccess ? apt_ic_gen_photon_e_dev(h, dg, de, n, seed, epoch, nullptr)
                                  : h->fail(APT_ERR_HIP, hipGetErrorString(e));
  if (st == APT_OK) {
    e = hipStreamSynchronize(h->stream);
    if (e == hipSuccess) e = hipMemcpy(eph, de, n * 8, hipMemcpyDeviceToHost);
    if (e != hipSuccess) st = h->fail(APT_ERR_HIP, hipGetErrorString(e));
  }
  (void)hipFree(dg);
  (void)hipFree(de);
  return st;
}

apt_status apt_ic_rates(apt_ic *h, const double *gamma, double *ic_rate, double *gg_rate,
                        uint64_t n) {
  if (!h) return APT_ERR_INVALID_ARG;
  if (!h->ready) return h->fail(APT_ERR_STATE, "ic_rates: apt_ic_init has not run");
  if (n == 0) return APT_OK;
  if (!gamma || !ic_rate || !gg_rate) return h->fail(APT_ERR_INVALID_ARG, "ic_rates: null arrays");
  if (n > IC_MAX_SAMPLES) return h->fail(APT_ERR_INVALID_ARG, "ic_rates: too many samples");
  double *d = nullptr;
  hipError_t e = hipMalloc(&d, n * 8 * 3);
  if (e == hipSuccess) e = hipMemcpy(d, gamma, n * 8, hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    k_ic_rates<<<(unsigned)((n + 255) / 256), 256, 0, h->stream>>>(tables_of(h), d, d + n,
                                                                  d + 2 * n, n);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  if (e == hipSuccess) e = hipMemcpy(ic_rate, d + n, n * 8, hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = hipMemcpy(gg_rate, d + 2 * n, n * 8, hipMemcpyDeviceToHost);
  (void)hipFree(d);
  return e == hipSuccess ? APT_OK : h->fail(APT_ERR_HIP, hipGetErrorString(e));
}

// triplet_pairs::triplet_pairs (rt_tpp.cu:147-184): the gamma grid, the
// same as the inverse-Compton one
apt_status apt_tpp_create(int n_gamma, apt_tpp **out) {
  // the pair-energy table is n_gamma x n_gamma doubles, one thread per entry
  if (!out || n_gamma < 3 || n_gamma > 11585) return APT_ERR_INVALID_ARG;
  apt_tpp *h = new apt_tpp();
  h->ng = n_gamma;
  h->dg = (apt_log(IC_MAX_GAMMA) - apt_log(IC_MIN_GAMMA)) / ((double)n_gamma - 1.0);
  h->gammas.resize(n_gamma);
  for (int n = 0; n < n_gamma; n++)
    h->gammas[n] = apt_exp(apt_log(IC_MIN_GAMMA) + h->dg * (double)n);
  hipError_t e = hipStreamCreate(&h->stream);
  if (e == hipSuccess) e = hipMalloc(&h->d_gammas, n_gamma * 8);
  if (e == hipSuccess) e = hipMalloc(&h->d_rate, n_gamma * 8);
  if (e == hipSuccess) e = hipMalloc(&h->d_em, n_gamma * 8);
  if (e == hipSuccess) e = hipMalloc(&h->d_dnde, (size_t)n_gamma * n_gamma * 8);
  if (e == hipSuccess)
    e = hipMemcpy(h->d_gammas, h->gammas.data(), n_gamma * 8, hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    apt_tpp_destroy(h);
    return APT_ERR_HIP;
  }
  *out = h;
  return APT_OK;
}

apt_status apt_tpp_destroy(apt_tpp *h) {
  if (!h) return APT_ERR_INVALID_ARG;
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  for (double *p : {h->d_gammas, h->d_rate, h->d_em, h->d_dnde}) (void)hipFree(p);
  if (h->stream) (void)hipStreamDestroy(h->stream);
  delete h;
  return APT_OK;
}

const char *apt_tpp_last_error(const apt_tpp *h) { return h ? h->err.c_str() : "null handle"; }

// triplet_pairs::init (rt_tpp.cu:188-275)
apt_status apt_tpp_init(apt_tpp *h, const apt_ic_spectrum *spec, double emin, double emax,
                        double n0) {
  if (!h) return APT_ERR_INVALID_ARG;
  if (!spec || spec->kind < 0 || spec->kind > APT_SPEC_MONO_ENERGETIC || !(emin > 0.0) ||
      !(emax > emin))
    return h->fail(APT_ERR_INVALID_ARG, "tpp_init: bad spectrum or energy range");
  Spectrum ne;
  ne.kind = spec->kind;
  for (int k = 0; k < 5; k++) ne.p[k] = spec->p[k];
  Grid g;
  g.emin = emin;
  g.n0 = n0;
  g.dmu = 2.0 / (IC_N_MU - 1.0);
  g.de = (apt_log(emax) - apt_log(emin)) / (IC_N_E - 1.0);
  const int ng = h->ng;
  k_tpp_rate<<<(ng + 63) / 64, 64, 0, h->stream>>>(ne, g, h->d_gammas, ng, h->d_rate, h->d_em);
  const int cells = ng * ng;
  k_tpp_dnde<<<(cells + 255) / 256, 256, 0, h->stream>>>(ne, g, h->d_gammas, ng, h->d_rate,
                                                         h->d_em, h->d_dnde);
  k_ic_cdf<<<(ng + 63) / 64, 64, 0, h->stream>>>(h->d_dnde, ng, ng);
  ICCHK(hipGetLastError());
  ICCHK(hipStreamSynchronize(h->stream));
  h->ready = true;
  return APT_OK;
}

apt_status apt_tpp_get_table(apt_tpp *h, int which, double *dst, uint64_t count) {
  if (!h) return APT_ERR_INVALID_ARG;
  const uint64_t ng = h->ng;
  const double *src = nullptr;
  uint64_t n = 0;
  switch (which) {
    case APT_TPP_GAMMAS: src = h->d_gammas; n = ng; break;
    case APT_TPP_RATE: src = h->d_rate; n = ng; break;
    case APT_TPP_EM: src = h->d_em; n = ng; break;
    case APT_TPP_DNDE: src = h->d_dnde; n = ng * ng; break;
    default: return h->fail(APT_ERR_INVALID_ARG, "tpp_get_table: unknown table");
  }
  if (!dst || count != n) return h->fail(APT_ERR_INVALID_ARG, "tpp_get_table: bad size");
  if (which != APT_TPP_GAMMAS && !h->ready)
    return h->fail(APT_ERR_STATE, "tpp_get_table: apt_tpp_init has not run");
  ICCHK(hipStreamSynchronize(h->stream));
  ICCHK(hipMemcpy(dst, src, n * 8, hipMemcpyDeviceToHost));
  return APT_OK;
}

static TppTables tpp_tables_of(const apt_tpp *h) {
  TppTables t;
  t.gammas = h->d_gammas;
  t.rate = h->d_rate;
  t.em = h->d_em;
  t.dnde = h->d_dnde;
  t.ng = h->ng;
  t.dg = h->dg;
  return t;
}

// host arrays in and out: gen_tpp_Ep and (find_tpp_rate, find_tpp_Em)
static apt_status tpp_run(apt_tpp *h, const double *gamma, double *o1, double *o2, uint64_t n,
                          uint64_t seed, uint32_t epoch, bool sample) {
  if (!h) return APT_ERR_INVALID_ARG;
  if (!h->ready) return h->fail(APT_ERR_STATE, "tpp: apt_tpp_init has not run");
  if (n == 0) return APT_OK;
  if (!gamma || !o1 || (!sample && !o2)) return h->fail(APT_ERR_INVALID_ARG, "tpp: null arrays");
  if (n > IC_MAX_SAMPLES) return h->fail(APT_ERR_INVALID_ARG, "tpp: too many samples");
  double *d = nullptr;
  hipError_t e = hipMalloc(&d, n * 8 * 3);
  if (e == hipSuccess) e = hipMemcpy(d, gamma, n * 8, hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    const unsigned blocks = (unsigned)((n + 255) / 256);
    if (sample)
      k_tpp_sample<<<blocks, 256, 0, h->stream>>>(tpp_tables_of(h), d, d + n, n, seed, epoch);
    else
      k_tpp_rates<<<blocks, 256, 0, h->stream>>>(tpp_tables_of(h), d, d + n, d + 2 * n, n);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  if (e == hipSuccess) e = hipMemcpy(o1, d + n, n * 8, hipMemcpyDeviceToHost);
  if (e == hipSuccess && !sample) e = hipMemcpy(o2, d + 2 * n, n * 8, hipMemcpyDeviceToHost);
  (void)hipFree(d);
  return e == hipSuccess ? APT_OK : h->fail(APT_ERR_HIP, hipGetErrorString(e));
}

apt_status apt_tpp_gen_ep(apt_tpp *h, const double *gamma, double *e_p, uint64_t n,
                          uint64_t seed, uint32_t epoch) {
  return tpp_run(h, gamma, e_p, nullptr, n, seed, epoch, true);
}

apt_status apt_tpp_rates(apt_tpp *h, const double *gamma, double *rate, double *em,
                         uint64_t n) {
  return tpp_run(h, gamma, rate, em, n, 0, 0, false);
}

// InverseComptonPL1D / InverseComptonDummy (inverse_compton_power_law.cuh,
// inverse_compton_dummy.cuh) over n particles, device arrays
apt_status apt_ic_model_emit(const apt_ic_model *model, const double *gamma, const double *p,
                             uint8_t *emit, double *eph, double *path, uint64_t n,
                             uint64_t seed, uint32_t epoch, void *stream) {
  if (!model || (model->kind != APT_ICM_DUMMY && model->kind != APT_ICM_POWER_LAW))
    return APT_ERR_INVALID_ARG;
  if (n == 0) return APT_OK;
  if (!gamma || !p || !emit || !eph || !path || n > IC_MAX_SAMPLES) return APT_ERR_INVALID_ARG;
  if (model->kind == APT_ICM_POWER_LAW &&
      !(model->alpha > 0.0 && model->e_s > 0.0 && model->e_min > 0.0))
    return APT_ERR_INVALID_ARG;
  IcModel m;
  m.kind = model->kind;
  m.alpha = model->alpha;
  m.es = model->e_s;
  m.emin = model->e_min;
  m.mfp = model->photon_path;
  m.icrate = model->ic_rate;
  m.cutoff = model->E_cutoff;
  m.eph = model->E_ph;
  k_ic_model<<<(unsigned)((n + 255) / 256), 256, 0, (hipStream_t)stream>>>(
      m, gamma, p, emit, eph, path, n, seed, epoch);
  return hipGetLastError() == hipSuccess ? APT_OK : APT_ERR_HIP;
}

}  // extern "C"
