// kern_cart_mixed.hip -- mixed-precision Cartesian particle update: the
// same fused push + move + Esirkepov deposit as the FP64 form
// (ptc_updater.cu:21-499), computed in FP32 on the 36-byte particle layout
// (int32 cell, FP32 in-cell offsets, FP32 u = gamma v, FP32 weight, flag).
// gamma is recomputed from u instead of stored.  Every current
// contribution is rounded to 2^-H and summed in int64, so the deposit is
// deterministic and charge conserving to the round-off of the FP32 shapes.
#include "apt_common.hpp"
#include "apt_f32.hpp"
#include "apt_kernels.hpp"
#include "apt_tiles.hpp"

namespace apt {

#pragma clang fp contract(fast)

// shape arrays over the four node offsets -2..1 around the old cell:
// s0 = old position, s1 = new position (moved by dc in {-1,0,1})
__device__ __forceinline__ void shapes(float xo, float xn, int dc, float s0[4],
                                       float s1[4]) {
#pragma unroll
  for (int t = 0; t < 4; t++) {
    s0[t] = 0.0f;
    s1[t] = 0.0f;
  }
  s0[1] = 1.0f - xo;
  s0[2] = xo;
  s1[1 + dc] = 1.0f - xn;
  s1[2 + dc] = xn;
}

template <int DIM>
__global__ void __launch_bounds__(256)
    k_push_deposit_mixed(Mesh m, Species sp, float dt, float scale, PtcF32 p,
                         FieldsF32 f, AccMixed a,
                         unsigned long long *__restrict__ counters,
                         uint64_t first, uint64_t n) {
  long long *__restrict__ J0 = a.J[0];
  long long *__restrict__ J1 = a.J[1];
  long long *__restrict__ J2 = a.J[2];
  uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const uint64_t idx = first + t;
  uint32_t c = p.cell[idx];
  if (c == MAX_CELL) return;
  int c1, c2, c3;
  m.split(c, c1, c2, c3);
  if (!m.in_bulk(c1, c2, c3)) {
    p.cell[idx] = MAX_CELL;
    return;
  }
  uint32_t flag = p.flag[idx];
  int s = ptc_species(flag);
  float x1 = p.x1[idx], x2 = p.x2[idx], x3 = p.x3[idx];
  float u1 = p.u1[idx], u2 = p.u2[idx], u3 = p.u3[idx];
  float w = p.w[idx];
  if (u1 != u1 || u2 != u2 || u3 != u3) atomicAdd(&counters[CTR_NAN], 1ull);
  float gamma = sqrtf(1.0f + u1 * u1 + u2 * u2 + u3 * u3);
  if (!(flag & FLAG_IGNORE_EM)) {
    float qm = (float)(0.5 * sp.q[s] / sp.m[s]) * dt;
    float E1 = gatherf<DIM>(f.E[0], m, c1, c2, c3, x1, x2, x3, ST_E1) * qm;
    float E2 = gatherf<DIM>(f.E[1], m, c1, c2, c3, x1, x2, x3, ST_E2) * qm;
    float E3 = gatherf<DIM>(f.E[2], m, c1, c2, c3, x1, x2, x3, ST_E3) * qm;
    float B1 = gatherf<DIM>(f.B[0], m, c1, c2, c3, x1, x2, x3, ST_B1) * qm;
    float B2 = gatherf<DIM>(f.B[1], m, c1, c2, c3, x1, x2, x3, ST_B2) * qm;
    float B3 = gatherf<DIM>(f.B[2], m, c1, c2, c3, x1, x2, x3, ST_B3) * qm;
    vay_pushf(u1, u2, u3, gamma, E1, E2, E3, B1, B2, B3);
    gamma = sqrtf(1.0f + u1 * u1 + u2 * u2 + u3 * u3);
    p.u1[idx] = u1;
    p.u2[idx] = u2;
    p.u3[idx] = u3;
  }
  float ig = 1.0f / gamma;
  float v1 = u1 * ig, v2 = u2 * ig, v3 = u3 * ig;
  float qw = (float)sp.q[s] * w;

  // ---- move ----
  float n1 = x1 + v1 * dt * (float)m.inv_delta[0];
  int dc1 = (int)floorf(n1);
  n1 -= (float)dc1;
  float n2 = x2, n3 = x3;
  int dc2 = 0, dc3 = 0;
  if (DIM >= 2) {
    n2 = x2 + v2 * dt * (float)m.inv_delta[1];
    dc2 = (int)floorf(n2);
    n2 -= (float)dc2;
  } else {
    n2 = x2 + v2 * dt;
  }
  if (DIM == 3) {
    n3 = x3 + v3 * dt * (float)m.inv_delta[2];
    dc3 = (int)floorf(n3);
    n3 -= (float)dc3;
  } else {
    n3 = x3 + v3 * dt;
  }
  // an offset that rounds to 1.0f belongs to the next cell
  if (n1 >= 1.0f) { n1 -= 1.0f; dc1 += 1; }
  if (DIM >= 2 && n2 >= 1.0f) { n2 -= 1.0f; dc2 += 1; }
  if (DIM == 3 && n3 >= 1.0f) { n3 -= 1.0f; dc3 += 1; }
  if (dc1 > 1 || dc1 < -1 || dc2 > 1 || dc2 < -1 || dc3 > 1 || dc3 < -1) {
    atomicAdd(&counters[CTR_MULTICELL], 1ull);
    dc1 = max(-1, min(1, dc1));
    dc2 = max(-1, min(1, dc2));
    dc3 = max(-1, min(1, dc3));
  }
  p.cell[idx] = (uint32_t)m.idx(m.wrap(0, c1 + dc1),
                                DIM >= 2 ? m.wrap(1, c2 + dc2) : c2,
                                DIM == 3 ? m.wrap(2, c3 + dc3) : c3);
  p.x1[idx] = n1;
  p.x2[idx] = n2;
  p.x3[idx] = n3;
  if (flag & FLAG_IGNORE_CURRENT) return;

  // ---- Esirkepov deposit (ptc_updater.cu:264-291, 343-384, 443-497) ----
  float sx0[4], sx1[4];
  shapes(x1, n1, dc1, sx0, sx1);
  if (DIM == 1) {
    float acc = 0.0f;
#pragma unroll
    for (int t = 0; t < 4; t++) {
      acc += sx1[t] - sx0[t];
      int i = c1 + t - 2;
      if (t < 3) acc_fx(J0, m.widx(i + 1, c2, c3), -qw * acc, scale);
      float val = 0.5f * (sx0[t] + sx1[t]);
      if (val != 0.0f) {
        uint64_t o = m.widx(i, c2, c3);
        acc_fx(J1, o, qw * v2 * val, scale);
        acc_fx(J2, o, qw * v3 * val, scale);
        if (a.rho && sx1[t] != 0.0f) acc_fx(a.R[s], o, qw * sx1[t], scale);
      }
    }
    return;
  }
  float sy0[4], sy1[4];
  shapes(x2, n2, dc2, sy0, sy1);
  if (DIM == 2) {
    const float wt = -qw;  // reference sign convention (ptc_updater.cu:345)
    float djy[4] = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int tj = 0; tj < 4; tj++) {
      float ay = 0.5f * (sy0[tj] + sy1[tj]);
      float dsy = sy1[tj] - sy0[tj];
      float djx = 0.0f;
      int j = c2 + tj - 2;
#pragma unroll
      for (int ti = 0; ti < 4; ti++) {
        int i = c1 + ti - 2;
        djx += (sx1[ti] - sx0[ti]) * ay;
        if (ti < 3 && djx != 0.0f) acc_fx(J0, m.widx(i + 1, j, c3), wt * djx, scale);
        djy[ti] += dsy * 0.5f * (sx0[ti] + sx1[ti]);
        if (tj < 3 && djy[ti] != 0.0f)
          acc_fx(J1, m.widx(i, j + 1, c3), wt * djy[ti], scale);
        float cen = (2.0f * sx1[ti] * sy1[tj] + sx0[ti] * sy1[tj] +
                     sx1[ti] * sy0[tj] + 2.0f * sx0[ti] * sy0[tj]) *
                    (1.0f / 6.0f);
        if (cen != 0.0f) acc_fx(J2, m.widx(i, j, c3), -wt * v3 * cen, scale);
        if (a.rho && sx1[ti] * sy1[tj] != 0.0f)
          acc_fx(a.R[s], m.widx(i, j, c3), qw * (sx1[ti] * sy1[tj]), scale);
      }
    }
    return;
  }
  float sz0[4], sz1[4];
  shapes(x3, n3, dc3, sz0, sz1);
  {
    const float wt = -qw;
    float djz[4][4];
#pragma unroll
    for (int a = 0; a < 4; a++)
#pragma unroll
      for (int b = 0; b < 4; b++) djz[a][b] = 0.0f;
#pragma unroll
    for (int tk = 0; tk < 4; tk++) {
      int k = c3 + tk - 2;
      float djy[4] = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
      for (int tj = 0; tj < 4; tj++) {
        int j = c2 + tj - 2;
        float djx = 0.0f;
        float cyz = (2.0f * sy1[tj] * sz1[tk] + sy0[tj] * sz1[tk] +
                     sy1[tj] * sz0[tk] + 2.0f * sy0[tj] * sz0[tk]) *
                    (1.0f / 6.0f);
#pragma unroll
        for (int ti = 0; ti < 4; ti++) {
          int i = c1 + ti - 2;
          djx += (sx1[ti] - sx0[ti]) * cyz;
          if (ti < 3 && djx != 0.0f)
            acc_fx(J0, m.widx(i + 1, j, k), wt * djx, scale);
          float cxz = (2.0f * sx1[ti] * sz1[tk] + sx0[ti] * sz1[tk] +
                       sx1[ti] * sz0[tk] + 2.0f * sx0[ti] * sz0[tk]) *
                      (1.0f / 6.0f);
          djy[ti] += (sy1[tj] - sy0[tj]) * cxz;
          if (tj < 3 && djy[ti] != 0.0f)
            acc_fx(J1, m.widx(i, j + 1, k), wt * djy[ti], scale);
          float cxy = (2.0f * sx1[ti] * sy1[tj] + sx0[ti] * sy1[tj] +
                       sx1[ti] * sy0[tj] + 2.0f * sx0[ti] * sy0[tj]) *
                      (1.0f / 6.0f);
          djz[tj][ti] += (sz1[tk] - sz0[tk]) * cxy;
          if (tk < 3 && djz[tj][ti] != 0.0f)
            acc_fx(J2, m.widx(i, j, k + 1), wt * djz[tj][ti], scale);
          float s3 = sx1[ti] * sy1[tj] * sz1[tk];
          if (a.rho && s3 != 0.0f) acc_fx(a.R[s], m.widx(i, j, k), qw * s3, scale);
        }
      }
    }
  }
}

__global__ void k_finalize_mixed(long long *__restrict__ acc,
                                 float *__restrict__ out, uint64_t n,
                                 double inv_scale) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  out[i] = (float)((double)acc[i] * inv_scale);
  acc[i] = 0;
}

hipError_t launch_push_deposit_mixed_range(const Mesh &m, const Species &sp,
                                           float dt, int S, const PtcF32 &p,
                                           const FieldsF32 &f,
                                           const AccMixed &a,
                                           unsigned long long *counters,
                                           uint64_t first, uint64_t n,
                                           hipStream_t st) {
  if (n == 0) return hipSuccess;
  float scale = ldexpf(1.0f, S);
  dim3 block(256), grid((unsigned)((n + 255) / 256));
  if (m.dim == 1)
    k_push_deposit_mixed<1><<<grid, block, 0, st>>>(m, sp, dt, scale, p, f, a,
                                                    counters, first, n);
  else if (m.dim == 2)
    k_push_deposit_mixed<2><<<grid, block, 0, st>>>(m, sp, dt, scale, p, f, a,
                                                    counters, first, n);
  else
    k_push_deposit_mixed<3><<<grid, block, 0, st>>>(m, sp, dt, scale, p, f, a,
                                                    counters, first, n);
  return hipGetLastError();
}

hipError_t launch_push_deposit_mixed(const Mesh &m, const Species &sp,
                                     float dt, int S, const PtcF32 &p,
                                     const FieldsF32 &f, const AccMixed &a,
                                     const TileInfo &tiles,
                                     unsigned long long *counters,
                                     uint64_t n, hipStream_t st) {
  if (tiles.tile_off == nullptr || tiles.ntiles == 0)
    return launch_push_deposit_mixed_range(m, sp, dt, S, p, f, a, counters, 0,
                                           n, st);
  hipError_t e = launch_push_deposit_tiled(m, sp, dt, S, p, f, a, tiles,
                                           counters, st);
  if (e != hipSuccess) return e;
  // slots past the tiled range (appended since the last sort)
  if (n > tiles.tiled_end)
    e = launch_push_deposit_mixed_range(m, sp, dt, S, p, f, a, counters,
                                        tiles.tiled_end, n - tiles.tiled_end,
                                        st);
  return e;
}

hipError_t launch_finalize_mixed(long long *acc, float *out, uint64_t n,
                                 int S, hipStream_t st) {
  dim3 block(256), grid((unsigned)((n + 255) / 256));
  k_finalize_mixed<<<grid, block, 0, st>>>(acc, out, n, ldexp(1.0, -S));
  return hipGetLastError();
}

}  // namespace apt
