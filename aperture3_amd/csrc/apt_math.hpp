// apt_math.hpp -- FP64 exp, sin/cos, atan2 and acos for the log-spherical
// particle path, identical on the host and the device.
//
// The log-spherical mover (ptc_updater_logsph.cu:163-230) goes through exp,
// sin, cos, acos, log and atan2.  Library versions of these round
// differently on the GPU and the CPU, so the FP64 form uses the published
// algorithms below (contraction off, so every operation rounds as written):
//   exp:     Cody & Waite reduction x = k ln2 + r (ln2 split hi/lo), Taylor
//            series of e^r to r^13, scaled by 2^k.
//   sin/cos: reduction by pi/2 with a three-part Cody & Waite constant,
//            Taylor series to r^19 / r^18 on |r| <= pi/4, quadrant table.
//   atan2:   octant reduction to atan(z), z in [0,1]; two half-angle steps
//            atan(z) = 2 atan(z / (1 + sqrt(1 + z^2))) bring |z| <= 0.199;
//            alternating series to z^25.
//   acos:    2 atan2(sqrt(1 - t), sqrt(1 + t)), t clamped to [-1, 1].
// Accuracy is a few ulp (pinned against libm by the tests).
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>

namespace apt {

__host__ __device__ inline double apt_exp(double x) {
#pragma clang fp contract(off)
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

__host__ __device__ inline void apt_sincos(double x, double &s, double &c) {
#pragma clang fp contract(off)
  const double two_over_pi = 6.36619772367581382433e-01;
  const double pio2_1 = 1.57079632673412561417e+00;
  const double pio2_2 = 6.07710050630396597660e-11;
  const double pio2_3 = 2.02226624871116645580e-21;
  double k = rint(x * two_over_pi);
  double r = ((x - k * pio2_1) - k * pio2_2) - k * pio2_3;
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
  double sr = r + r * (z * ps);
  double pc = 1.0 / 6402373705728000.0;
  pc = pc * z - 1.0 / 20922789888000.0;
  pc = pc * z + 1.0 / 87178291200.0;
  pc = pc * z - 1.0 / 479001600.0;
  pc = pc * z + 1.0 / 3628800.0;
  pc = pc * z - 1.0 / 40320.0;
  pc = pc * z + 1.0 / 720.0;
  pc = pc * z - 1.0 / 24.0;
  pc = pc * z + 0.5;
  double cr = 1.0 - z * pc;
  switch (((int)k) & 3) {
    case 0: s = sr; c = cr; break;
    case 1: s = cr; c = -sr; break;
    case 2: s = -sr; c = -cr; break;
    default: s = -cr; c = sr; break;
  }
}

__host__ __device__ inline double apt_sin(double x) {
  double s, c;
  apt_sincos(x, s, c);
  return s;
}

__host__ __device__ inline double apt_atan01(double z) {
#pragma clang fp contract(off)
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

__host__ __device__ inline double apt_atan2(double y, double x) {
#pragma clang fp contract(off)
  const double pi = 3.14159265358979311600e+00;
  const double pio2 = 1.57079632679489655800e+00;
  double ax = fabs(x), ay = fabs(y);
  if (ax == 0.0 && ay == 0.0) return 0.0;
  bool swap = ay > ax;
  double z = swap ? ax / ay : ay / ax;
  double a = apt_atan01(z);
  if (swap) a = pio2 - a;
  if (x < 0.0) a = pi - a;
  return y < 0.0 ? -a : a;
}

__host__ __device__ inline double apt_acos(double t) {
#pragma clang fp contract(off)
  if (t > 1.0) t = 1.0;
  if (t < -1.0) t = -1.0;
  return 2.0 * apt_atan2(sqrt(1.0 - t), sqrt(1.0 + t));
}

}  // namespace apt
