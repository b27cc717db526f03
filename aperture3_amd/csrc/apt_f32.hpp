// apt_f32.hpp -- FP32 device building blocks of the mixed form: CIC taps
// and gathers (interpolation.cuh:28-209 restricted to the non-zero taps),
// the Vay pusher in FP32 (vay_push.cuh:8-30) and the int64 fixed-point
// current add.  Including files decide contraction (the mixed kernels
// allow it).
#pragma once
#include "apt_common.hpp"

namespace apt {

__device__ __forceinline__ float cicf(float d) {
  return fmaxf(1.0f - fabsf(d), 0.0f);
}

// CIC weights of the two non-zero taps of a component with stagger bit st
// (interp_cell, interpolation.cuh:79-87): upper node (st = 1) -> taps c-1,
// c; cell centre (st = 0) -> taps c-1, c when x < 1/2 else c, c+1.
__device__ __forceinline__ void taps(float x, int st, int &o, float &w0,
                                     float &w1) {
  if (st) {
    o = -1;
    w0 = 1.0f - x;
    w1 = x;
  } else if (x < 0.5f) {
    o = -1;
    w0 = 0.5f - x;
    w1 = 0.5f + x;
  } else {
    o = 0;
    w0 = 1.5f - x;
    w1 = x - 0.5f;
  }
}

template <int DIM>
__device__ __forceinline__ float gatherf(const float *__restrict__ f,
                                         const Mesh &m, int c1, int c2,
                                         int c3, float x1, float x2, float x3,
                                         int st) {
  int o1, o2, o3;
  float a0, a1, b0, b1, g0, g1;
  taps(x1, st & 1, o1, a0, a1);
  if (DIM == 1) {
    const float *p = f + (c1 + o1);
    return p[0] * a0 + p[1] * a1;
  }
  taps(x2, (st >> 1) & 1, o2, b0, b1);
  const int d0 = m.dims[0];
  if (DIM == 2) {
    const float *p = f + (int64_t)(c2 + o2) * d0 + (c1 + o1);
    return (p[0] * a0 + p[1] * a1) * b0 + (p[d0] * a0 + p[d0 + 1] * a1) * b1;
  }
  taps(x3, (st >> 2) & 1, o3, g0, g1);
  const int64_t d01 = (int64_t)d0 * m.dims[1];
  const float *p = f + (int64_t)(c3 + o3) * d01 + (int64_t)(c2 + o2) * d0 +
                   (c1 + o1);
  float lo = (p[0] * a0 + p[1] * a1) * b0 + (p[d0] * a0 + p[d0 + 1] * a1) * b1;
  const float *q = p + d01;
  float hi = (q[0] * a0 + q[1] * a1) * b0 + (q[d0] * a0 + q[d0 + 1] * a1) * b1;
  return lo * g0 + hi * g1;
}

__device__ __forceinline__ void vay_pushf(float &p1, float &p2, float &p3,
                                          float gamma, float E1, float E2,
                                          float E3, float B1, float B2,
                                          float B3) {
  float ig = 1.0f / gamma;
  float up1 = p1 + 2.0f * E1 + (p2 * B3 - p3 * B2) * ig;
  float up2 = p2 + 2.0f * E2 + (p3 * B1 - p1 * B3) * ig;
  float up3 = p3 + 2.0f * E3 + (p1 * B2 - p2 * B1) * ig;
  float tt = B1 * B1 + B2 * B2 + B3 * B3;
  float ut = up1 * B1 + up2 * B2 + up3 * B3;
  float sigma = 1.0f + up1 * up1 + up2 * up2 + up3 * up3 - tt;
  float inv_gamma2 =
      2.0f / (sigma + sqrtf(sigma * sigma + 4.0f * (tt + ut * ut)));
  float s = 1.0f / (1.0f + inv_gamma2 * tt);
  float ign = sqrtf(inv_gamma2);
  p1 = (up1 + B1 * ut * inv_gamma2 + (up2 * B3 - up3 * B2) * ign) * s;
  p2 = (up2 + B2 * ut * inv_gamma2 + (up3 * B1 - up1 * B3) * ign) * s;
  p3 = (up3 + B3 * ut * inv_gamma2 + (up1 * B2 - up2 * B1) * ign) * s;
}

__device__ __forceinline__ void acc_fx(long long *__restrict__ acc,
                                       uint64_t o, float c, float scale) {
  long long v = (long long)__float2int_rn(c * scale);
  if (v != 0)
    __hip_atomic_fetch_add((unsigned long long *)(acc + o),
                           (unsigned long long)v, __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
}

// hardware FP32 square root and reciprocal (<= 1 ulp); arguments here are
// >= 1 or well inside the normal range
__device__ __forceinline__ float hw_sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ __forceinline__ float hw_rcp(float x) { return __builtin_amdgcn_rcpf(x); }

// Vay push (vay_push.cuh:8-30) in FP32 with the hardware reciprocal and
// square root; returns 1/gamma of the new momentum.  Contraction follows
// the including file or block.
__device__ __forceinline__ float vay_push_f(float &p1, float &p2, float &p3,
                                           float gamma, float E1, float E2,
                                           float E3, float B1, float B2,
                                           float B3) {
#pragma clang fp contract(fast)
  float ig = hw_rcp(gamma);
  float up1 = p1 + 2.0f * E1 + (p2 * B3 - p3 * B2) * ig;
  float up2 = p2 + 2.0f * E2 + (p3 * B1 - p1 * B3) * ig;
  float up3 = p3 + 2.0f * E3 + (p1 * B2 - p2 * B1) * ig;
  float tt = B1 * B1 + B2 * B2 + B3 * B3;
  float ut = up1 * B1 + up2 * B2 + up3 * B3;
  float sigma = 1.0f + up1 * up1 + up2 * up2 + up3 * up3 - tt;
  float inv_gamma2 =
      2.0f * hw_rcp(sigma + hw_sqrt(sigma * sigma + 4.0f * (tt + ut * ut)));
  float s = hw_rcp(1.0f + inv_gamma2 * tt);
  float ign = hw_sqrt(inv_gamma2);
  p1 = (up1 + B1 * ut * inv_gamma2 + (up2 * B3 - up3 * B2) * ign) * s;
  p2 = (up2 + B2 * ut * inv_gamma2 + (up3 * B1 - up1 * B3) * ign) * s;
  p3 = (up3 + B3 * ut * inv_gamma2 + (up1 * B2 - up2 * B1) * ign) * s;
  return ign;
}

}  // namespace apt
