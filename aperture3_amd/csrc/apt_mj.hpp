// apt_mj.hpp -- Maxwell-Juttner momentum sampler on the counter-based
// Philox stream, shared by the loader and the batched ensemble's wall
// injection.  Contraction off in every including file: the FP64 draw is
// bit-identical to the CPU restatement.
#pragma once
#include "apt_common.hpp"

namespace apt {

constexpr uint32_t MJ_MAX_TRIES = 4096u;

// Maxwell-Juttner momentum in the rest frame, then Zenitani's flipping
// method and a boost along x for a drift beta_x.  theta < 0.4: rejection
// from a Gamma(3/2, theta') proposal for the kinetic energy t = gamma-1
// (1/theta' = 1/theta - 5/4, using 1+x <= e^x to bound the MJ density
// sqrt(t(t+2))(t+1)e^{-t/theta}); Gamma(3/2) = Exp(1) + Z^2/2, Z^2/2 from
// Marsaglia's polar pair.  theta >= 0.4: Sobol's four-uniform algorithm.
// Direction: Marsaglia's (1972) point on the sphere.
__device__ inline bool sample_mj(double theta, double beta_x, uint64_t seed,
                                 uint64_t pid, double &ux, double &uy,
                                 double &uz) {
  double umag = 0.0;
  bool ok = false;
  if (theta > 0.0 && theta < 0.4) {
    double thp = 1.0 / (1.0 / theta - 1.25);
    for (uint32_t draw = 0; draw < MJ_MAX_TRIES; draw++) {
      double a0, a1, b0, b1;
      uniform2(seed, pid, 2 * draw, 1, a0, a1);
      uniform2(seed, pid, 2 * draw + 1, 1, b0, b1);
      double v1 = 2.0 * b0 - 1.0, v2 = 2.0 * b1 - 1.0;
      double s = v1 * v1 + v2 * v2;
      if (s >= 1.0 || s <= 0.0) continue;
      double t = thp * (-apt_log(a0) - v1 * v1 * apt_log(s) / s);
      double lhs = apt_log(a1);
      double rhs =
          0.5 * apt_log(1.0 + 0.5 * t) + apt_log(1.0 + t) - 1.25 * t;
      if (lhs < rhs) {
        umag = sqrt(t * (t + 2.0));
        ok = true;
        break;
      }
    }
  } else if (theta >= 0.4) {
    for (uint32_t draw = 0; draw < MJ_MAX_TRIES; draw++) {
      double a0, a1, b0, b1;
      uniform2(seed, pid, 2 * draw, 1, a0, a1);
      uniform2(seed, pid, 2 * draw + 1, 1, b0, b1);
      double x123 = a0 * a1 * b0;
      double uu = -theta * apt_log(x123);
      double eta = -theta * apt_log(x123 * b1);
      if (eta * eta - uu * uu > 1.0) {
        umag = uu;
        ok = true;
        break;
      }
    }
  } else {
    ok = true;
  }
  double d0 = 0.0, d1 = 0.0, d2 = 0.0;
  if (umag > 0.0) {
    for (uint32_t draw = 0; draw < MJ_MAX_TRIES; draw++) {
      double b0, b1;
      uniform2(seed, pid, draw, 2, b0, b1);
      double v1 = 2.0 * b0 - 1.0, v2 = 2.0 * b1 - 1.0;
      double s = v1 * v1 + v2 * v2;
      if (s >= 1.0) continue;
      double r = 2.0 * sqrt(1.0 - s);
      d0 = 1.0 - 2.0 * s;
      d1 = v1 * r;
      d2 = v2 * r;
      break;
    }
  }
  ux = umag * d0;
  uy = umag * d1;
  uz = umag * d2;
  if (beta_x != 0.0) {
    double g = sqrt(1.0 + ux * ux + uy * uy + uz * uz);
    double c0, c1;
    uniform2(seed, pid, 0, 3, c0, c1);
    if (-beta_x * ux / g > c0) ux = -ux;
    double G = 1.0 / sqrt(1.0 - beta_x * beta_x);
    ux = G * (ux + beta_x * g);
  }
  return ok;
}

}  // namespace apt
