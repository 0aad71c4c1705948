// FP32-compensated bond geometry and sums shared by the untiled and tiled
// FP32C kernels (the form is described in pd_step32.hip).
#pragma once
#include "pd_wave.h"

__device__ __forceinline__ void kahan(float& s, float& c, float v) {
  const float yk = v - c;
  const float t = s + yk;
  c = (t - s) - yk;
  s = t;
}

struct G32 {
  float Y1, Y2, Y3, inv_x, inv_y, xij, ext, s;
};

__device__ __forceinline__ void geom32(const float4& xp, const double4& yp, const float4& xj, const double4& yj,
                                       G32& g) {
  const float X1 = xj.x - xp.x, X2 = xj.y - xp.y, X3 = xj.z - xp.z;
  const float U1 = (float)((yj.x - yp.x) - (double)X1);
  const float U2 = (float)((yj.y - yp.y) - (double)X2);
  const float U3 = (float)((yj.z - yp.z) - (double)X3);
  g.Y1 = X1 + U1; g.Y2 = X2 + U2; g.Y3 = X3 + U3;
  const float xx2 = X1 * X1 + X2 * X2 + X3 * X3;
  const float xu = 2.0f * (X1 * U1 + X2 * U2 + X3 * U3) + (U1 * U1 + U2 * U2 + U3 * U3);
  const float yy2 = xx2 + xu;
  g.inv_x = __builtin_amdgcn_rsqf(xx2);
  const float xij = xx2 * g.inv_x;
  g.xij = xij;
  g.inv_y = __builtin_amdgcn_rsqf(yy2);
  const float yij = yy2 * g.inv_y;
  g.ext = xu * __builtin_amdgcn_rcpf(yij + xij);
  g.s = g.ext * g.inv_x;
}

__device__ __forceinline__ double comp_total(float s, float c, int C) {
  return group_sum((double)s - (double)c, C);
}


// sticky "has yielded" flag of an elastoplastic particle in the sign of its
// FP32 record's td_norm2 (aux32.z), as pd_has_yielded reads the FP64 record
__device__ __forceinline__ bool pd_has_yielded32(const float4& a) { return __builtin_signbit(a.z); }
