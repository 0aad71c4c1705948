// Device helpers shared by the kernels.  The library is compiled with
// -ffp-contract=off: every a*b+c below rounds twice, as in the reference's Julia
// code (no @fastmath / @muladd there).  The reference-order (exact) form uses
// IEEE FP64 division and square root only; the fast forms use pd_rsqrt/pd_rcp.
#pragma once
#include "pd_internal.h"

// get_magnitude (src/misc/const_macros.jl:116): sqrt((a1*a1 + a2*a2) + a3*a3)
__device__ __forceinline__ double pd_mag(double a1, double a2, double a3) {
  return __builtin_sqrt(a1 * a1 + a2 * a2 + a3 * a3);
}

// Julia `^(x::Float64, n::Integer)` power by squaring in double-double
// (base/math.jl pow_body; muladd taken as fma).  n == 1 returns x, n == 2 x*x.
__device__ __forceinline__ double pd_pow_int(double x, int64_t n) {
  if (n == 0) return 1.0;
  if (n == 3) return x * x * x;
  double y = 1.0, xnlo = 0.0, ynlo = 0.0;
  if (n < 0) {
    double rx = 1.0 / x;
    if (n == -2) return rx * rx;
    if (__builtin_isfinite(x)) xnlo = -__builtin_fma(x, rx, -1.0) * rx;
    x = rx;
    n = -n;
  }
  while (n > 1) {
    if (n & 1) {
      double err = __builtin_fma(y, xnlo, x * ynlo);
      double p = x * y;
      double plo = __builtin_fma(x, y, -p);
      y = p;
      ynlo = plo + err;
    }
    double err = x * 2.0 * xnlo;
    double p = x * x;
    double plo = __builtin_fma(x, x, -p);
    x = p;
    xnlo = plo + err;
    n >>= 1;
  }
  double err = __builtin_fma(y, xnlo, x * ynlo);
  return (__builtin_isfinite(x) && __builtin_isfinite(err)) ? __builtin_fma(x, y, err) : x * y;
}

// Direction of a ToFroBC during step s = time_step + 1 (1-based), i.e. after the
// check! calls of the completed steps, at time_step values t = 0 .. time_step-1
// (ToFroBC.jl:107-111: flip when t > applyafter and t % freq == 0).
__device__ __forceinline__ double pd_bc_dir(const BcDev& bc, int64_t time_step) {
  const int64_t hi = time_step - 1;                  // last t checked
  int64_t lo = bc.applyafter + 1;
  if (lo < 0) lo = 0;
  if (hi < lo) return 1.0;
  int64_t flips;
  if (__builtin_isinf(bc.freq)) {
    flips = (lo == 0) ? 1 : 0;                       // t % Inf == t
  } else {
    const int64_t fr = (int64_t)bc.freq;
    // multiples of fr in [lo, hi]
    const int64_t upto_hi = hi / fr + 1;             // multiples in [0, hi]
    const int64_t upto_lo = (lo > 0) ? ((lo - 1) / fr + 1) : 0;
    flips = upto_hi - upto_lo;
  }
  return (flips & 1) ? -1.0 : 1.0;
}

__device__ __forceinline__ int pd_wave_id() {
  return __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
}
__device__ __forceinline__ int pd_lane() { return threadIdx.x & 63; }

// FP64 form: a squared length with fused multiply-adds, fma(c, c, fma(b, b,
// a * a)), one rounding per term (the reference-order form keeps the
// reference's separate products and sums).  The tiled and untiled FP64
// kernels use the same fused expressions, so they stay bit-identical.
__device__ __forceinline__ double pd_norm2(double a, double b, double c) {
  return __builtin_fma(c, c, __builtin_fma(b, b, a * a));
}

// Wavefront-level LDS ordering: every lane's earlier LDS writes are visible to
// every lane of the wavefront after this point.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// number of set bits of m below this lane
__device__ __forceinline__ int lanes_below(uint64_t m) {
  return (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

// XCD-aware block id: consecutive logical blocks (spatial neighbours sharing
// neighbour data) land on the same XCD / L2.  Hardware dispatches workgroup b
// to XCD b % 8.
// XCD x receives q + (x < r) workgroups (q = nblocks / 8, r = nblocks % 8), so
// the map b -> start(x) + b / 8 is a bijection onto [0, nblocks).
__device__ __forceinline__ int pd_xcd_block(int nblocks) {
  const int b = blockIdx.x;
  const int q = nblocks / 8, r = nblocks % 8;
  const int xcd = b % 8, idx = b / 8;
  const int start = xcd * q + (xcd < r ? xcd : r);
  return start + idx;
}

// ---- global-memory accessors ------------------------------------------------
// The array pointers live in the device-side EnvDev table, so the compiler
// cannot prove they address global memory and would emit generic (flat)
// loads; these casts make the hot loops use global_load/global_store.  The
// byte offset is formed in 32 bits and added to the (wave-uniform) base, so the
// loads take the scalar-base form (one VGPR of offset, no 64-bit address
// arithmetic per lane).  Every array of an env therefore spans < 4 GiB, which
// pd_env_commit checks (PD_MAX_ARRAY_BYTES).
typedef double pd_d4v __attribute__((ext_vector_type(4)));
typedef float pd_f4v __attribute__((ext_vector_type(4)));
#define PD_GLOBAL __attribute__((address_space(1)))

template <typename T>
__device__ __forceinline__ const PD_GLOBAL T* gaddr(const T* p, uint32_t i) {
  return (const PD_GLOBAL T*)((const PD_GLOBAL char*)p + (uint32_t)(i * (uint32_t)sizeof(T)));
}
template <typename T>
__device__ __forceinline__ PD_GLOBAL T* gaddr_w(T* p, uint32_t i) {
  return (PD_GLOBAL T*)((PD_GLOBAL char*)p + (uint32_t)(i * (uint32_t)sizeof(T)));
}
template <typename T>
__device__ __forceinline__ T gld(const T* p, uint32_t i) {
  return *gaddr(p, i);
}
template <typename T>
__device__ __forceinline__ void gst(T* p, uint32_t i, T v) {
  *gaddr_w(p, i) = v;
}
__device__ __forceinline__ double4 gld4(const double4* p, uint32_t i) {
  const pd_d4v v = *(const PD_GLOBAL pd_d4v*)gaddr(p, i);
  return make_double4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ float4 gld4(const float4* p, uint32_t i) {
  const pd_f4v v = *(const PD_GLOBAL pd_f4v*)gaddr(p, i);
  return make_float4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ void gst4(double4* p, uint32_t i, double4 v) {
  pd_d4v w;
  w.x = v.x; w.y = v.y; w.z = v.z; w.w = v.w;
  *(PD_GLOBAL pd_d4v*)gaddr_w(p, i) = w;
}

// ---- fast-form FP64 reciprocals (not used by the reference-order form) -------
// Hardware estimates (relative error <= 5.3e-8 measured on gfx950) refined by
// one Newton-Raphson step: <= 4.3e-15 relative (2.2e-15 for the reciprocal),
// far inside the 1e-10 agreement of the FP64 form, at a fraction of the cost
// of the correctly-rounded division and square-root sequences.
__device__ __forceinline__ double pd_rsqrt(double a) {
  const double r = __builtin_amdgcn_rsq(a);
  const double e = __builtin_fma(-a * r, r, 1.0);
  return __builtin_fma(0.5 * r, e, r);
}
__device__ __forceinline__ double pd_rcp(double a) {
  const double r = __builtin_amdgcn_rcp(a);
  const double e = __builtin_fma(-a, r, 1.0);
  return __builtin_fma(r, e, r);
}

// Fast forms, elastoplastic blocks: the td pass keeps a sticky "has yielded"
// flag in the sign bit of aux.z (|aux.z| = td_norm2 >= 0).  A particle that has
// never yielded (C > 1 at every step so far) has never changed its plastic
// extensions, which are therefore exactly 0, so their loads and stores can be
// skipped without changing any result.
__device__ __forceinline__ bool pd_has_yielded(const double4& aux) { return __builtin_signbit(aux.z); }

// Slot of bond k of a particle whose bond 0 sits at `base` (layout in pd_internal.h).
__host__ __device__ __forceinline__ uint32_t pd_slot(uint32_t base, int k, int lgG) {
  const int G = 1 << lgG;
  const int C = 64 >> lgG;
  return base + 64u * (uint32_t)(k >> lgG) + (uint32_t)((k & (G - 1)) * C);
}
