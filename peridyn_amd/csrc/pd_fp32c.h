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

// Compensated sum of FP32 terms taken in groups: the terms of one group (a
// lane's kGroupSteps = 8 consecutive bond steps, the bonds of two packed words
// in the tiled kernels) are accumulated in plain FP32, and each group's total
// enters a Kahan sum.  The error stays a constant multiple of eps * sum |term|
// (at most 8 roundings per group, 2 for the compensated sum), independent of
// the number of groups, at an eighth of the per-term compensation cost.  Tiled
// and untiled kernels group the same bond steps, so their sums stay
// bit-identical: a group ends after an odd packed word or the chunk's last.
// A group's running total starts at -0 (the additive identity, so its first
// term enters exactly) and takes each term a * b as a fused multiply-add.
// Bond steps per group: two packed words of the tiled kernels.
constexpr int kGroupSteps = 8;
__device__ __forceinline__ bool group_end_word(int t4, int nt4) { return (t4 & 1) || t4 == nt4 - 1; }
__device__ __forceinline__ int group_count(int width) { return (width + kGroupSteps - 1) / kGroupSteps; }

struct GroupKahan {
  float g = -0.0f, s = 0.0f, c = 0.0f;
  __device__ __forceinline__ void add(float a, float b) { g = __builtin_fmaf(a, b, g); }
  __device__ __forceinline__ void flush() {
    kahan(s, c, g);
    g = -0.0f;
  }
};

struct G32 {
  float Y1, Y2, Y3, inv_x, inv_y, xij, ext, s;
};

// Displacement of a particle from its FP32-rounded reference position, in FP64:
// w = y - (corner + x32), with x32 the FP32 offset from the block's corner (exact
// for lattice positions).  Formed once per particle (per halo record in the
// tiled kernels); a bond's displacement difference is then U = float(w_j - w_i).
__device__ __forceinline__ double3 disp32(const double3& corner, const float4& x32, const double4& y) {
  return make_double3(y.x - (corner.x + (double)x32.x), y.y - (corner.y + (double)x32.y),
                      y.z - (corner.z + (double)x32.z));
}
__device__ __forceinline__ double3 corner_of(const BlockDev& B) {
  return make_double3(B.corner[0], B.corner[1], B.corner[2]);
}
__device__ __forceinline__ double3 disp32(const BlockDev& B, const float4& x32, const double4& y) {
  return disp32(corner_of(B), x32, y);
}

// Bond geometry from the FP32 reference offsets and the FP64 displacements:
//   X = x32_j - x32_i,  U = float(w_j - w_i),  Y = X + U,
//   |Y|^2 = |X|^2 + (2 X.U + U.U),  e = (2 X.U + U.U) / (|Y| + |X|),  s = e / |X|
// (no |Y| - |X| cancellation); dot products as fused multiply-adds.
__device__ __forceinline__ void geom32(const float4& xp, const double3& wp, const float4& xj, const double3& wj,
                                       G32& g) {
  const float X1 = xj.x - xp.x, X2 = xj.y - xp.y, X3 = xj.z - xp.z;
  const float U1 = (float)(wj.x - wp.x);
  const float U2 = (float)(wj.y - wp.y);
  const float U3 = (float)(wj.z - wp.z);
  g.Y1 = X1 + U1; g.Y2 = X2 + U2; g.Y3 = X3 + U3;
  const float xx2 = __builtin_fmaf(X1, X1, __builtin_fmaf(X2, X2, X3 * X3));
  const float xdu = __builtin_fmaf(X1, U1, __builtin_fmaf(X2, U2, X3 * U3));
  const float udu = __builtin_fmaf(U1, U1, __builtin_fmaf(U2, U2, U3 * U3));
  const float xu = __builtin_fmaf(2.0f, xdu, udu);
  const float yy2 = xx2 + xu;
  g.inv_x = __builtin_amdgcn_rsqf(xx2);
  const float xij = xx2 * g.inv_x;
  g.xij = xij;
  g.inv_y = __builtin_amdgcn_rsqf(yy2);
  const float yij = yy2 * g.inv_y;
  g.ext = xu * __builtin_amdgcn_rcpf(yij + xij);
  g.s = g.ext * g.inv_x;
}

// ordinary state-based t (ordinary_state_based.jl:176-187) in FP32 from the
// owner's (theta, 1/m) record ap, its product apxy = theta/m, and the
// neighbour's (theta_j, 1/m_j): A = 3K - 5G, B15 = 15G;
// t = A (theta_i/m_i + theta_j/m_j) + 15G s (1/m_i + 1/m_j)
__device__ __forceinline__ float osb_t32(const float4& ap, float apxy, float thj, float rmj, float A, float B15,
                                         const G32& g) {
  const float sm = ap.y + rmj;
  const float twm = __builtin_fmaf(thj, rmj, apxy);
  return __builtin_fmaf(A, twm, (B15 * g.s) * sm);
}

// Per-neighbour factor of the force term xx = (t / |Y|) vol_j / density: for
// bond-based blocks t = c11 s, so the factor c11 vol_j / density is taken with
// it; formed once per halo record in the tiled kernels, per bond in the others
__device__ __forceinline__ float vd32(float vol, float inv_den) { return vol * inv_den; }
__device__ __forceinline__ float cvd32(float c11, float vol, float inv_den) { return c11 * (vol * inv_den); }

__device__ __forceinline__ double comp_total(float s, float c, int C) {
  return group_sum((double)s - (double)c, C);
}


// sticky "has yielded" flag of an elastoplastic particle in the sign of its
// FP32 record's td_norm2 (aux32.z), as pd_has_yielded reads the FP64 record
__device__ __forceinline__ bool pd_has_yielded32(const float4& a) { return __builtin_signbit(a.z); }
