// kern_cart_f64.hip -- plain FP64 Cartesian particle update (the drop-in
// form).  One fused pass per particle: push (ptc_push_cart_{1,2,3}d,
// ptc_updater.cu:21-218), then move + Esirkepov deposit
// (deposit_current_cart_{1,2,3}d, ptc_updater.cu:220-499), then the periodic
// wrap that the reference does through send_particles.  Expressions keep the
// reference's operation order and this file is built with -ffp-contract=off,
// so every particle's arithmetic is bit-identical to the CPU restatement.
// The current goes through two-limb int64 fixed point (order independent)
// with device-scope atomics on the wrapped (periodic) index; in 1D the lanes
// of a wave that add to the same element are summed first (one atomic pair
// per element: 4.4x faster at 200 particles per cell), in 2D and 3D each
// lane adds its own terms (the wave sums cost more than the atomics there).
#include "apt_common.hpp"
#include "apt_f64.hpp"
#include "apt_kernels.hpp"

namespace apt {

// Push (ptc_push_cart_1d, ptc_updater.cu:21-76) and move (the first part
// of deposit_current_cart_1d, :220-245) of one live particle: stores
// momenta, gamma, cell and offsets, and returns what the deposit needs.
__device__ __forceinline__ void push_move_1d_f64(const Mesh &m, const Species &sp, double dt,
                                                 const PtcF64 &p, const FieldsF64 &f,
                                                 unsigned long long *counters, uint64_t idx,
                                                 uint32_t c, int &s, bool &dep, double &weight,
                                                 double &old_x1, double &new_x1, int &dc1,
                                                 double &v2, double &v3) {
  uint32_t flag = p.flag[idx];
  s = ptc_species(flag);
  double x1 = p.x1[idx], x2 = p.x2[idx], x3 = p.x3[idx];
  double p1 = p.p1[idx], p2 = p.p2[idx], p3 = p.p3[idx];
  double gamma = p.E[idx];
  double q_over_m = dt * 0.5 * sp.q[s] / sp.m[s];
  if (p1 != p1 || p2 != p2 || p3 != p3)
    atomicAdd(&counters[CTR_NAN], 1ull);
  gamma = sqrt(1.0 + p1 * p1 + p2 * p2 + p3 * p3);
  if (!(flag & FLAG_IGNORE_EM)) {
    double E1 = gather1(f.E[0], x1, (int)c, ST_E1) * q_over_m;
    double E2 = gather1(f.E[1], x1, (int)c, ST_E2) * q_over_m;
    double E3 = gather1(f.E[2], x1, (int)c, ST_E3) * q_over_m;
    double B1 = gather1(f.B[0], x1, (int)c, ST_B1) * q_over_m;
    double B2 = gather1(f.B[1], x1, (int)c, ST_B2) * q_over_m;
    double B3 = gather1(f.B[2], x1, (int)c, ST_B3) * q_over_m;
    vay_push(p1, p2, p3, gamma, E1, E2, E3, B1, B2, B3);
    p.p1[idx] = p1;
    p.p2[idx] = p2;
    p.p3[idx] = p3;
    p.E[idx] = gamma;
  } else {
    // ignore_EM: stored momenta and gamma are left as they were
    p1 = p.p1[idx];
    p2 = p.p2[idx];
    p3 = p.p3[idx];
    gamma = p.E[idx];
  }
  double v1 = p1 / gamma;
  v2 = p2 / gamma;
  v3 = p3 / gamma;
  double w = p.w[idx];
  old_x1 = x1;
  new_x1 = old_x1 + (v1 * dt) / m.delta[0];
  dc1 = (int)floor(new_x1);
  if (dc1 > 1 || dc1 < -1) atomicAdd(&counters[CTR_MULTICELL], 1ull);
  new_x1 -= (double)dc1;
  p.cell[idx] = (uint32_t)m.wrap(0, (int)c + dc1);
  p.x1[idx] = new_x1;
  p.x2[idx] = x2 + v2 * dt;
  p.x3[idx] = x3 + v3 * dt;
  weight = sp.q[s] * w;
  dep = !(flag & FLAG_IGNORE_CURRENT);
}

template <int DIM>
__global__ void __launch_bounds__(256)
    k_push_deposit_f64(Mesh m, Species sp, double dt, int H, int rho,
                       PtcF64 p, FieldsF64 f, AccF64 a,
                       unsigned long long *__restrict__ counters,
                       uint64_t n) {
  uint64_t idx = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  // the 1D deposit goes through wave_acc_add, which every lane of the wave
  // enters: lanes without a particle take part with nothing to add
  bool live = idx < n;
  uint32_t c = live ? p.cell[idx] : MAX_CELL;
  live = live && c != MAX_CELL;
  int c1 = 0, c2 = 0, c3 = 0;
  if (live) {
    m.split(c, c1, c2, c3);
    if (!m.in_bulk(c1, c2, c3)) {
      p.cell[idx] = MAX_CELL;
      live = false;
    }
  }
  if (DIM != 1 && !live) return;
  if (DIM == 1) {
    bool dep = false;
    int s = 0, dc1 = 0;
    double weight = 0.0, old_x1 = 0.0, new_x1 = 0.0, v2 = 0.0, v3 = 0.0;
    if (live)
      push_move_1d_f64(m, sp, dt, p, f, counters, idx, c, s, dep, weight, old_x1, new_x1, dc1,
                       v2, v3);
    {
      // deposit_current_cart_1d (ptc_updater.cu:220-293): i over i0..i1
      const int i0 = (dc1 == -1 ? -2 : -1), i1 = (dc1 == 1 ? 1 : 0);
      double djx = 0.0;
      for (int i = -2; i <= 1; i++) {
        const bool in = dep && i >= i0 && i <= i1;
        double vj1 = 0.0, vj2 = 0.0, vj3 = 0.0, vr = 0.0;
        uint64_t o1 = 0, o = 0;
        if (in) {
          double sx0 = cic(-old_x1 + i + 1);
          double sx1 = cic(-new_x1 + (i + 1 - dc1));
          djx += sx1 - sx0;
          o1 = m.widx(i + (int)c + 1, 0, 0);
          o = m.widx(i + (int)c, 0, 0);
          vj1 = -weight * djx;
          double val1 = 0.5 * (sx0 + sx1);
          vj2 = weight * v2 * val1;
          vj3 = weight * v3 * val1;
          vr = weight * sx1;
        }
        wave_acc_add(a.J[0][0], a.J[0][1], o1, o1, vj1, H, in);
        wave_acc_add(a.J[1][0], a.J[1][1], o, o, vj2, H, in);
        wave_acc_add(a.J[2][0], a.J[2][1], o, o, vj3, H, in);
        if (rho)
          wave_acc_add(a.R[s][0], a.R[s][1], o, o | ((uint64_t)s << 56), vr, H, in);
      }
    }
    return;
  }
  const int d0 = m.dims[0], d1 = m.dims[1];
  uint32_t flag = p.flag[idx];
  int s = ptc_species(flag);
  double x1 = p.x1[idx], x2 = p.x2[idx], x3 = p.x3[idx];
  double p1 = p.p1[idx], p2 = p.p2[idx], p3 = p.p3[idx];
  double gamma = p.E[idx];
  double q_over_m = dt * 0.5 * sp.q[s] / sp.m[s];
  if (p1 != p1 || p2 != p2 || p3 != p3)
    atomicAdd(&counters[CTR_NAN], 1ull);
  gamma = sqrt(1.0 + p1 * p1 + p2 * p2 + p3 * p3);
  if (!(flag & FLAG_IGNORE_EM)) {
    double E1, E2, E3, B1, B2, B3;
    if (DIM == 2) {
      E1 = gather2(f.E[0], d0, x1, x2, c1, c2, ST_E1) * q_over_m;
      E2 = gather2(f.E[1], d0, x1, x2, c1, c2, ST_E2) * q_over_m;
      E3 = gather2(f.E[2], d0, x1, x2, c1, c2, ST_E3) * q_over_m;
      B1 = gather2(f.B[0], d0, x1, x2, c1, c2, ST_B1) * q_over_m;
      B2 = gather2(f.B[1], d0, x1, x2, c1, c2, ST_B2) * q_over_m;
      B3 = gather2(f.B[2], d0, x1, x2, c1, c2, ST_B3) * q_over_m;
    } else {
      E1 = gather3(f.E[0], d0, d1, x1, x2, x3, c1, c2, c3, ST_E1) * q_over_m;
      E2 = gather3(f.E[1], d0, d1, x1, x2, x3, c1, c2, c3, ST_E2) * q_over_m;
      E3 = gather3(f.E[2], d0, d1, x1, x2, x3, c1, c2, c3, ST_E3) * q_over_m;
      B1 = gather3(f.B[0], d0, d1, x1, x2, x3, c1, c2, c3, ST_B1) * q_over_m;
      B2 = gather3(f.B[1], d0, d1, x1, x2, x3, c1, c2, c3, ST_B2) * q_over_m;
      B3 = gather3(f.B[2], d0, d1, x1, x2, x3, c1, c2, c3, ST_B3) * q_over_m;
    }
    vay_push(p1, p2, p3, gamma, E1, E2, E3, B1, B2, B3);
    p.p1[idx] = p1;
    p.p2[idx] = p2;
    p.p3[idx] = p3;
    p.E[idx] = gamma;
  } else {
    // ignore_EM: stored momenta and gamma are left as they were
    p1 = p.p1[idx];
    p2 = p.p2[idx];
    p3 = p.p3[idx];
    gamma = p.E[idx];
  }

  // ---- move + deposit (reads the stored momenta / gamma) ----
  double v1 = p1 / gamma, v2 = p2 / gamma, v3 = p3 / gamma;
  double w = p.w[idx];
  const double old_x1 = x1, old_x2 = x2, old_x3 = x3;

  if (DIM == 2) {
    double new_x1 = old_x1 + (v1 * dt) * m.inv_delta[0];
    double new_x2 = old_x2 + (v2 * dt) * m.inv_delta[1];
    int dc1 = (int)floor(new_x1);
    int dc2 = (int)floor(new_x2);
    if (dc1 > 1 || dc1 < -1 || dc2 > 1 || dc2 < -1)
      atomicAdd(&counters[CTR_MULTICELL], 1ull);
    new_x1 -= (double)dc1;
    new_x2 -= (double)dc2;
    p.cell[idx] = (uint32_t)m.idx(m.wrap(0, c1 + dc1), m.wrap(1, c2 + dc2), 0);
    p.x1[idx] = new_x1;
    p.x2[idx] = new_x2;
    p.x3[idx] = old_x3 + v3 * dt;
    if (flag & FLAG_IGNORE_CURRENT) return;
    double weight = -sp.q[s] * w;
    int j0 = (dc2 == -1 ? -2 : -1), j1 = (dc2 == 1 ? 1 : 0);
    int i0 = (dc1 == -1 ? -2 : -1), i1 = (dc1 == 1 ? 1 : 0);
    double djy[3] = {0.0, 0.0, 0.0};
    for (int j = j0; j <= j1; j++) {
      double sy0 = cic(-old_x2 + j + 1);
      double sy1 = cic(-new_x2 + (j + 1 - dc2));
      double djx = 0.0;
      for (int i = i0; i <= i1; i++) {
        double sx0 = cic(-old_x1 + i + 1);
        double sx1 = cic(-new_x1 + (i + 1 - dc1));
        djx += movement2d(sy0, sy1, sx0, sx1);
        acc_add(a.J[0][0], a.J[0][1], m.widx(i + c1 + 1, j + c2, 0),
                weight * djx, H);
        djy[i - i0] += movement2d(sx0, sx1, sy0, sy1);
        acc_add(a.J[1][0], a.J[1][1], m.widx(i + c1, j + c2 + 1, 0),
                weight * djy[i - i0], H);
        double val2 = center2d(sx0, sx1, sy0, sy1);
        uint64_t o = m.widx(i + c1, j + c2, 0);
        acc_add(a.J[2][0], a.J[2][1], o, -weight * v3 * val2, H);
        if (rho) {
          double s1 = sx1 * sy1;
          acc_add(a.R[s][0], a.R[s][1], o, -weight * s1, H);
        }
      }
    }
  } else {
    double new_x1 = old_x1 + (v1 * dt) * m.inv_delta[0];
    double new_x2 = old_x2 + (v2 * dt) * m.inv_delta[1];
    double new_x3 = old_x3 + (v3 * dt) * m.inv_delta[2];
    int dc1 = (int)floor(new_x1);
    int dc2 = (int)floor(new_x2);
    int dc3 = (int)floor(new_x3);
    if (dc1 > 1 || dc1 < -1 || dc2 > 1 || dc2 < -1 || dc3 > 1 || dc3 < -1)
      atomicAdd(&counters[CTR_MULTICELL], 1ull);
    new_x1 -= (double)dc1;
    new_x2 -= (double)dc2;
    new_x3 -= (double)dc3;
    p.cell[idx] = (uint32_t)m.idx(m.wrap(0, c1 + dc1), m.wrap(1, c2 + dc2),
                                  m.wrap(2, c3 + dc3));
    p.x1[idx] = new_x1;
    p.x2[idx] = new_x2;
    p.x3[idx] = new_x3;
    if (flag & FLAG_IGNORE_CURRENT) return;
    double weight = -sp.q[s] * w;
    int k0 = (dc3 == -1 ? -2 : -1), k1 = (dc3 == 1 ? 1 : 0);
    int j0 = (dc2 == -1 ? -2 : -1), j1 = (dc2 == 1 ? 1 : 0);
    int i0 = (dc1 == -1 ? -2 : -1), i1 = (dc1 == 1 ? 1 : 0);
    double djz[3][3] = {{0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}};
    for (int k = k0; k <= k1; k++) {
      double sz0 = cic(-old_x3 + k + 1);
      double sz1 = cic(-new_x3 + (k + 1 - dc3));
      double djy[3] = {0.0, 0.0, 0.0};
      for (int j = j0; j <= j1; j++) {
        double sy0 = cic(-old_x2 + j + 1);
        double sy1 = cic(-new_x2 + (j + 1 - dc2));
        double djx = 0.0;
        for (int i = i0; i <= i1; i++) {
          double sx0 = cic(-old_x1 + i + 1);
          double sx1 = cic(-new_x1 + (i + 1 - dc1));
          djx += movement3d(sy0, sy1, sz0, sz1, sx0, sx1);
          acc_add(a.J[0][0], a.J[0][1], m.widx(i + c1 + 1, j + c2, k + c3),
                  weight * djx, H);
          djy[i - i0] += movement3d(sx0, sx1, sz0, sz1, sy0, sy1);
          acc_add(a.J[1][0], a.J[1][1], m.widx(i + c1, j + c2 + 1, k + c3),
                  weight * djy[i - i0], H);
          djz[j - j0][i - i0] += movement3d(sx0, sx1, sy0, sy1, sz0, sz1);
          acc_add(a.J[2][0], a.J[2][1], m.widx(i + c1, j + c2, k + c3 + 1),
                  weight * djz[j - j0][i - i0], H);
          if (rho) {
            double s1 = sx1 * sy1 * sz1;
            acc_add(a.R[s][0], a.R[s][1], m.widx(i + c1, j + c2, k + c3),
                    -weight * s1, H);
          }
        }
      }
    }
  }
}

// J = hi 2^-H + lo 2^-(H+40), accumulators cleared for the next step.
__global__ void k_finalize_f64(long long *__restrict__ hi,
                               long long *__restrict__ lo,
                               double *__restrict__ out, uint64_t n, int H) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  out[i] = fx_value(hi[i], lo[i], H);
  hi[i] = 0;
  lo[i] = 0;
}

// ---------------------------------------------------------------------
// host launchers
// ---------------------------------------------------------------------
hipError_t launch_push_deposit_f64(const Mesh &m, const Species &sp,
                                   double dt, int H, int rho,
                                   const PtcF64 &p, const FieldsF64 &f,
                                   const AccF64 &a,
                                   unsigned long long *counters, uint64_t n,
                                   hipStream_t st) {
  if (n == 0) return hipSuccess;
  dim3 block(256), grid((unsigned)((n + 255) / 256));
  if (m.dim == 1)
    k_push_deposit_f64<1><<<grid, block, 0, st>>>(m, sp, dt, H, rho, p, f, a,
                                                  counters, n);
  else if (m.dim == 2)
    k_push_deposit_f64<2><<<grid, block, 0, st>>>(m, sp, dt, H, rho, p, f, a,
                                                  counters, n);
  else
    k_push_deposit_f64<3><<<grid, block, 0, st>>>(m, sp, dt, H, rho, p, f, a,
                                                  counters, n);
  return hipGetLastError();
}

hipError_t launch_finalize_f64(long long *hi, long long *lo, double *out,
                               uint64_t n, int H, hipStream_t st) {
  dim3 block(256), grid((unsigned)((n + 255) / 256));
  k_finalize_f64<<<grid, block, 0, st>>>(hi, lo, out, n, H);
  return hipGetLastError();
}

}  // namespace apt
