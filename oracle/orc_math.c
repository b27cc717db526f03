/*
 * orc_math.c -- exp, sin/cos, atan2 and acos for the oracle's log-spherical
 * step.  TEST INFRASTRUCTURE ONLY.
 *
 * The log-spherical mover (ptc_updater_logsph.cu:163-230) goes through
 * exp, sin, cos, acos, log and atan2.  Library versions of these differ
 * between the host and the GPU in the last bit, so the FP64 form computes
 * them with the published algorithms below, written out the same way on
 * both sides (and built without contraction), which keeps the two
 * bit-identical.  Accuracy is pinned against numpy in
 * tests/test_oracle_math.py (a few ulp).
 *
 *   exp:   Cody & Waite reduction x = k ln2 + r, |r| <= ln2/2 (ln2 split
 *          hi/lo), Taylor series of e^r to r^13, scaled by 2^k.
 *   sin/cos: reduction by pi/2 with a three-part Cody & Waite constant,
 *          Taylor series to r^19 / r^18 on |r| <= pi/4, quadrant table.
 *   atan2: octant reduction to atan(z), z in [0,1]; two half-angle steps
 *          atan(z) = 2 atan(z / (1 + sqrt(1 + z^2))) bring |z| <= 0.199;
 *          alternating series to z^25.
 *   acos:  acos(t) = 2 atan2(sqrt(1 - t), sqrt(1 + t)), t clamped to [-1,1].
 */
#include <math.h>

#include "orc_internal.h"

double orc_exp(double x) {
  const double inv_ln2 = 1.4426950408889634;
  const double ln2_hi = 6.93147180369123816490e-01;
  const double ln2_lo = 1.90821492927058770002e-10;
  double k = rint(x * inv_ln2);
  double r = (x - k * ln2_hi) - k * ln2_lo;
  double p = 1.0 / 6227020800.0;
  p = p * r + 1.0 / 479001600.0;
  p = p * r + 1.0 / 39916800.0;
  p = p * r + 1.0 / 3628800.0;
  p = p * r + 1.0 / 362880.0;
  p = p * r + 1.0 / 40320.0;
  p = p * r + 1.0 / 5040.0;
  p = p * r + 1.0 / 720.0;
  p = p * r + 1.0 / 120.0;
  p = p * r + 1.0 / 24.0;
  p = p * r + 1.0 / 6.0;
  p = p * r + 0.5;
  p = p * r + 1.0;
  p = p * r + 1.0;
  return ldexp(p, (int)k);
}

static void sincos_reduced(double r, double *s, double *c) {
  double z = r * r;
  double ps = -1.0 / 121645100408832000.0;
  ps = ps * z + 1.0 / 355687428096000.0;
  ps = ps * z - 1.0 / 1307674368000.0;
  ps = ps * z + 1.0 / 6227020800.0;
  ps = ps * z - 1.0 / 39916800.0;
  ps = ps * z + 1.0 / 362880.0;
  ps = ps * z - 1.0 / 5040.0;
  ps = ps * z + 1.0 / 120.0;
  ps = ps * z - 1.0 / 6.0;
  *s = r + r * (z * ps);
  double pc = 1.0 / 6402373705728000.0;
  pc = pc * z - 1.0 / 20922789888000.0;
  pc = pc * z + 1.0 / 87178291200.0;
  pc = pc * z - 1.0 / 479001600.0;
  pc = pc * z + 1.0 / 3628800.0;
  pc = pc * z - 1.0 / 40320.0;
  pc = pc * z + 1.0 / 720.0;
  pc = pc * z - 1.0 / 24.0;
  pc = pc * z + 0.5;
  *c = 1.0 - z * pc;
}

void orc_sincos(double x, double *s, double *c) {
  const double two_over_pi = 6.36619772367581382433e-01;
  const double pio2_1 = 1.57079632673412561417e+00;
  const double pio2_2 = 6.07710050630396597660e-11;
  const double pio2_3 = 2.02226624871116645580e-21;
  double k = rint(x * two_over_pi);
  double r = ((x - k * pio2_1) - k * pio2_2) - k * pio2_3;
  double sr, cr;
  sincos_reduced(r, &sr, &cr);
  int q = ((int)k) & 3;
  switch (q) {
    case 0: *s = sr; *c = cr; break;
    case 1: *s = cr; *c = -sr; break;
    case 2: *s = -sr; *c = -cr; break;
    default: *s = -cr; *c = sr; break;
  }
}

double orc_sin(double x) {
  double s, c;
  orc_sincos(x, &s, &c);
  return s;
}

double orc_cos(double x) {
  double s, c;
  orc_sincos(x, &s, &c);
  return c;
}

static double atan01(double z) {
  double t = z / (1.0 + sqrt(1.0 + z * z));
  t = t / (1.0 + sqrt(1.0 + t * t));
  double s = t * t;
  double p = 1.0 / 25.0;
  p = -(p * s) + 1.0 / 23.0;
  p = -(p * s) + 1.0 / 21.0;
  p = -(p * s) + 1.0 / 19.0;
  p = -(p * s) + 1.0 / 17.0;
  p = -(p * s) + 1.0 / 15.0;
  p = -(p * s) + 1.0 / 13.0;
  p = -(p * s) + 1.0 / 11.0;
  p = -(p * s) + 1.0 / 9.0;
  p = -(p * s) + 1.0 / 7.0;
  p = -(p * s) + 1.0 / 5.0;
  p = -(p * s) + 1.0 / 3.0;
  p = -(p * s) + 1.0;
  return 4.0 * (t * p);
}

double orc_atan2(double y, double x) {
  const double pi = 3.14159265358979311600e+00;
  const double pio2 = 1.57079632679489655800e+00;
  double ax = fabs(x), ay = fabs(y);
  if (ax == 0.0 && ay == 0.0) return 0.0;
  int swap = ay > ax;
  double z = swap ? ax / ay : ay / ax;
  double a = atan01(z);
  if (swap) a = pio2 - a;
  if (x < 0.0) a = pi - a;
  return y < 0.0 ? -a : a;
}

double orc_acos(double t) {
  if (t > 1.0) t = 1.0;
  if (t < -1.0) t = -1.0;
  return 2.0 * orc_atan2(sqrt(1.0 - t), sqrt(1.0 + t));
}
