// Per-step kernels of the explicit-dynamics hot path (gfx950, wave64).
//
// One wavefront owns one chunk of C = 64/G particles; lane L = g*C + q serves
// particle first+q and its bonds k = g, g+G, ... (bond layout in
// pd_internal.h): at every iteration the wavefront reads 64 consecutive column
// and state slots.  Each particle sums its own force (a gather: no atomics,
// deterministic); with G > 1 the G partial sums are combined by an xor-butterfly
// over lanes q, q+C, ..., which gives every lane the same total.
//
// Forms (template PREC):
//   PD_PREC_EXACT  (G = 1) reproduces the single-threaded reference's
//                  accumulation order for every particle p
//                  (general_material.jl:330-368): reverse terms from neighbours
//                  j < p in ascending j, then p's own row in descending j (the
//                  family order), then reverse terms from j > p in ascending j,
//                  each rounded as ((xx*Y)*vol)/rho; theta and td_norm2 are summed
//                  in the family's descending order.  Bit-identical to the oracle.
//   PD_PREC_FP64   FP64 with one evaluation per bond: 1/|X| and 1/|Y| are formed
//                  once, both halves of the pair force are combined, and the
//                  neighbour's (theta, 1/m, td_norm2, C) come from one 32-byte
//                  gather of `aux`.  Agrees with the exact form to ~1e-15 per step.
#include <algorithm>

#include "pd_wave.h"

namespace {

struct Geom {
  double X1, X2, X3, Y1, Y2, Y3, xij, yij, ext, s;
};

// X = x_j - x_i, Y = y_j - y_i (get_ij, const_macros.jl:150-178), stretch as in
// general_material.jl:336-341.
__device__ __forceinline__ void geom(const double4& xp, const double4& yp, const double4& xj,
                                     const double4& yj, Geom& g) {
  g.X1 = xj.x - xp.x; g.X2 = xj.y - xp.y; g.X3 = xj.z - xp.z;
  g.Y1 = yj.x - yp.x; g.Y2 = yj.y - yp.y; g.Y3 = yj.z - yp.z;
  g.yij = pd_mag(g.Y1, g.Y2, g.Y3);
  g.xij = pd_mag(g.X1, g.X2, g.X3);
  g.ext = g.yij - g.xij;
  g.s = g.ext / g.xij;
}

// elasto_plastic.jl:228-255 for owner o and neighbour q (a = w/m_o, b = w/m_q).
__device__ __forceinline__ double ep_t(double C, double K, double G, double xij, double ext,
                                       double a, double b, double th_o, double th_q,
                                       double edp_old, double* edp_new) {
  const double tw_m = th_o * a + th_q * b;
  const double theta_ij = tw_m / (a + b);
  const double ed = ext - (theta_ij * xij) / 3.0;
  const double ede = ed - edp_old;
  const double td_trial = ((15.0 * G) * (a + b)) * ede;
  double td;
  if (C > 1.0) {
    td = td_trial;
    *edp_new = edp_old;
  } else {
    td = C * td_trial;
    *edp_new = edp_old + ede * (1.0 - C);
  }
  const double tn = ((3.0 * K) * xij) * tw_m;
  return tn + td;
}

// many small device-to-device copies in one launch: one workgroup per
// descriptor (the run set's per-env chunk and tile lists)
__global__ __launch_bounds__(256) void k_copy_words(const PdCopy* __restrict__ d, int n) {
  const PdCopy c = d[blockIdx.x];
  for (uint32_t i = threadIdx.x; i < c.words; i += 256) c.dst[i] = c.src[i];
}

}  // namespace

// ---------------------------------------------------------------------------
// Dilatation theta (src/peri.jl:155-172) for OSB and EP blocks.
// ---------------------------------------------------------------------------
template <int PREC>
__global__ __launch_bounds__(256) void k_theta(const Chunk* __restrict__ chunks, int nchunks,
                                               EnvDev* __restrict__ envs) {
  WaveCtx w;
  if (!wave_chunk(chunks, nchunks, envs, w)) return;
  const EnvDev& E = envs[w.ch.env];
  const BlockDev& B = E.blocks[w.ch.block];
  if (B.kind == PD_BB || !B.deformed) return;
  if (E.precision == PD_PREC_FP32C) return;  // done by k_theta32
  const double4* __restrict__ x = E.x;
  const double4* __restrict__ y = E.y[E.cur];
  const int32_t* __restrict__ col = E.col;
  const uint8_t* __restrict__ st = E.state;
  const int p = w.active ? w.p : w.ch.first;
  const double4 xp = x[p], yp = y[p];
  const int d = w.active ? E.deg[p] : 0;
  const uint32_t base = w.ch.slot0 + w.lane;
  if (PREC == PD_PREC_EXACT) {
    if (!w.active) return;
    const double c_p = 3.0 / E.m[p];
    double th = 0.0;
    for (int k = d - 1; k >= 0; --k) {
      const uint32_t slot = base + 64u * (uint32_t)k;
      if (st[slot]) {
        const int j = col[slot];
        Geom gm;
        geom(xp, yp, x[j], y[j], gm);
        const double wgt = 1.0 / gm.xij;
        th += ((((c_p * wgt) * gm.xij) * gm.ext) * x[j].w) * 1.0;
      }
    }
    E.theta[p] = th;
    E.aux[p].x = th;
  } else {
    double acc = 0.0;
    constexpr int U = 1;  // bonds per batch, streamed data one batch ahead (material_fast)
    const int ni = (d > w.g) ? ((d - w.g + w.G - 1) >> w.lgG) : 0;
    int jn[U];
    uint8_t sn[U];
    auto load_stream = [&](int i0) {
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const bool ok = (i0 + u) < ni;
        const uint32_t slot = base + 64u * (uint32_t)(i0 + u);
        jn[u] = ok ? gld(col, slot) : p;
        sn[u] = ok ? gld(st, slot) : (uint8_t)0;
      }
    };
    load_stream(0);
    for (int i0 = 0; i0 < ni; i0 += U) {
      int jv[U];
      uint8_t sv[U];
#pragma unroll
      for (int u = 0; u < U; ++u) { jv[u] = jn[u]; sv[u] = sn[u]; }
      double4 xj[U], yj[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        xj[u] = gld4(x, jv[u]);
        yj[u] = gld4(y, jv[u]);
      }
      if (i0 + U < ni) load_stream(i0 + U);
#pragma unroll
      for (int u = 0; u < U; ++u) {
        if (!sv[u]) continue;
        const double X1 = xj[u].x - xp.x, X2 = xj[u].y - xp.y, X3 = xj[u].z - xp.z;
        const double Y1 = yj[u].x - yp.x, Y2 = yj[u].y - yp.y, Y3 = yj[u].z - yp.z;
        const double xx2 = pd_norm2(X1, X2, X3);
        const double yy2 = pd_norm2(Y1, Y2, Y3);
        const double ext = yy2 * pd_rsqrt(yy2) - xx2 * pd_rsqrt(xx2);
        acc = __builtin_fma(ext, xj[u].w, acc);
      }
    }
    acc = group_sum(acc, w.C);
    if (w.active && w.g == 0) {
      const double th = (3.0 * E.aux[p].y) * acc;
      E.theta[p] = th;
      E.aux[p].x = th;
    }
  }
}

// ---------------------------------------------------------------------------
// td_norm2 (elasto_plastic.jl:168-172,282-295 through loop_over_neighs!,
// src/peri.jl:411-445); breaks bonds with s >= s0.  A bond broken here keeps
// its plastic strain in both edp buffers.
// ---------------------------------------------------------------------------
template <int PREC>
__global__ __launch_bounds__(256) void k_td(const Chunk* __restrict__ chunks, int nchunks,
                                            EnvDev* __restrict__ envs) {
  WaveCtx w;
  if (!wave_chunk(chunks, nchunks, envs, w)) return;
  const EnvDev& E = envs[w.ch.env];
  const BlockDev& B = E.blocks[w.ch.block];
  if (B.kind != PD_EP || !B.deformed) return;
  const double4* __restrict__ x = E.x;
  const double4* __restrict__ y = E.y[E.cur];
  const int32_t* __restrict__ col = E.col;
  uint8_t* __restrict__ st = E.state;
  const double* __restrict__ edp = E.edp[E.edp_cur];
  double* __restrict__ edp_o = E.edp[E.edp_cur ^ 1];
  const int p = w.active ? w.p : w.ch.first;
  const double4 xp = x[p], yp = y[p];
  const int d = w.active ? E.deg[p] : 0;
  const int nt = B.ntype;
  const int tp = E.ptype[p];
  const uint32_t base = w.ch.slot0 + w.lane;
  if (PREC == PD_PREC_EXACT) {
    if (!w.active) return;
    const double m_p = E.m[p], th_p = E.theta[p];
    double acc = 0.0;
    for (int k = d - 1; k >= 0; --k) {
      const uint32_t slot = base + 64u * (uint32_t)k;
      if (!st[slot]) continue;
      const int j = col[slot];
      const double4 xj = x[j];
      Geom gm;
      geom(xp, yp, xj, y[j], gm);
      const int tj = (nt > 1) ? E.ptype[j] : 0;
      if (gm.s < B.s0[tp + nt * tj]) {
        const double wij = 1.0 / gm.xij;
        const double wji = 1.0 / pd_mag(-gm.X1, -gm.X2, -gm.X3);
        const double Gij = B.G[tp + nt * tj];
        const double a = wij / m_p, b = wji / E.m[j];
        const double w_m = a + b;
        const double tw_m = th_p * a + E.theta[j] * b;
        const double theta_ij = tw_m / w_m;
        const double ed = gm.ext - (theta_ij * gm.xij) / 3.0;
        const double td = ((15.0 * Gij) * (ed - edp[slot])) * w_m;
        acc += (td * td) * xj.w;
      } else {
        st[slot] = 0;
        edp_o[slot] = edp[slot];
      }
    }
    E.td[p] = acc;
    E.aux[p].z = acc;
    E.aux[p].w = __builtin_sqrt((2.0 * B.psi[tp + nt * tp]) / acc);
  } else {
    const double4 ap = E.aux[p];  // (theta, 1/m, +-td_norm2 of the last step, C)
    // p's plastic extensions are all exactly 0 until it first yields
    const bool yielded = pd_has_yielded(ap);
    double acc = 0.0;
    constexpr int U = 1;  // bonds per batch, streamed data one batch ahead (material_fast)
    const int ni = (d > w.g) ? ((d - w.g + w.G - 1) >> w.lgG) : 0;
    int jn[U];
    uint8_t sn[U];
    double en[U];
    auto load_stream = [&](int i0) {
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const bool ok = (i0 + u) < ni;
        const uint32_t slot = base + 64u * (uint32_t)(i0 + u);
        jn[u] = ok ? gld(col, slot) : p;
        sn[u] = ok ? gld(st, slot) : (uint8_t)0;
        en[u] = (ok && yielded) ? gld(edp, slot) : 0.0;
      }
    };
    load_stream(0);
    for (int i0 = 0; i0 < ni; i0 += U) {
      int jv[U];
      uint8_t sv[U];
      double eo[U];
#pragma unroll
      for (int u = 0; u < U; ++u) { jv[u] = jn[u]; sv[u] = sn[u]; eo[u] = en[u]; }
      double4 xj[U], yj[U], aj[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        xj[u] = gld4(x, jv[u]);
        yj[u] = gld4(y, jv[u]);
        aj[u] = gld4(E.aux, jv[u]);
      }
      if (i0 + U < ni) load_stream(i0 + U);
#pragma unroll
      for (int u = 0; u < U; ++u) {
        if (!sv[u]) continue;
        const uint32_t slot = base + 64u * (uint32_t)(i0 + u);
        const double X1 = xj[u].x - xp.x, X2 = xj[u].y - xp.y, X3 = xj[u].z - xp.z;
        const double Y1 = yj[u].x - yp.x, Y2 = yj[u].y - yp.y, Y3 = yj[u].z - yp.z;
        const double xx2 = pd_norm2(X1, X2, X3);
        const double yy2 = pd_norm2(Y1, Y2, Y3);
        const double inv_x = pd_rsqrt(xx2);
        const double xij = xx2 * inv_x;
        const double ext = yy2 * pd_rsqrt(yy2) - xij;
        const int tj = (nt > 1) ? gld(E.ptype, jv[u]) : 0;
        if (ext * inv_x < B.s0[tp + nt * tj]) {
          const double sm = ap.y + aj[u].y;
          const double twm = ap.x * ap.y + aj[u].x * aj[u].y;
          const double ed = ext - (twm * pd_rcp(sm)) * xij * (1.0 / 3.0);
          const double td = (15.0 * B.G[tp + nt * tj]) * (ed - eo[u]) * (inv_x * sm);
          acc = __builtin_fma(td * td, xj[u].w, acc);
        } else {
          gst(st, slot, (uint8_t)0);
          if (yielded) gst(edp_o, slot, eo[u]);  // else 0 in both buffers already
        }
      }
    }
    acc = group_sum(acc, w.C);
    if (w.active && w.g == 0) {
      const double C = __builtin_sqrt((2.0 * B.psi[tp + nt * tp]) / acc);
      // sticky yield flag in the sign of aux.z (multi-type blocks yield per type
      // pair: flagged from the start)
      const bool flag = yielded || !(C > 1.0) || nt > 1;
      E.td[p] = acc;
      E.aux[p].z = flag ? -acc : acc;
      E.aux[p].w = C;
    }
  }
}

// ---------------------------------------------------------------------------
// Fused force density + velocity-Verlet:
//   f_p = material force (force_density_T!, general_material.jl:296-369, skipped
//         while the block's deformed latch is off) + contact terms
//         (short_range_contact! RM12, contacts.jl:210-247) in model order;
//   KICK : v += dt/2 f; velocity BCs           (velocity_verlet.jl:74-79)
//   DRIFT: v += dt/2 f; y' = y + dt v; position BCs for the next step
//          (velocity_verlet.jl:63-69), written to the other y buffer.
// ---------------------------------------------------------------------------
template <int KIND>
__device__ __forceinline__ void material_exact(const EnvDev& E, const BlockDev& B, const WaveCtx& w,
                                               const double4& xp, const double4& yp, double& a1,
                                               double& a2, double& a3) {
  const double4* __restrict__ x = E.x;
  const double4* __restrict__ y = E.y[E.cur];
  const int32_t* __restrict__ col = E.col;
  uint8_t* __restrict__ st = E.state;
  const int p = w.p;
  const int nt = B.ntype;
  const int tp = E.ptype[p];
  const double den_p = B.density[tp];
  const int d = E.deg[p];
  const uint32_t s0 = w.ch.slot0 + w.lane;
  double m_p = 0.0, th_p = 0.0, td_p = 0.0, Cp1 = 0.0;
  const double* __restrict__ edp_in = nullptr;
  double* __restrict__ edp_out = nullptr;
  if (KIND != PD_BB) { m_p = E.m[p]; th_p = E.theta[p]; }
  if (KIND == PD_EP) {
    td_p = E.td[p];
    edp_in = E.edp[E.edp_cur];
    edp_out = E.edp[E.edp_cur ^ 1];
    Cp1 = __builtin_sqrt((2.0 * B.psi[tp + nt * tp]) / td_p);
  }
  const double c11 = B.c11;
  // t for owner p (general_material.jl:347-350 plug-in)
  auto t_own = [&](int j, int tj, uint32_t slot, const Geom& g) -> double {
    if (KIND == PD_BB) return g.s * c11;
    const double wij = 1.0 / g.xij;
    const double wji = 1.0 / pd_mag(-g.X1, -g.X2, -g.X3);
    const int ij = tp + nt * tj;
    if (KIND == PD_OSB) {
      const double a = wij / m_p, b = wji / E.m[j];
      return (3.0 * B.K[ij] - 5.0 * B.G[ij]) * (th_p * g.xij * a + E.theta[j] * g.xij * b) +
             15.0 * B.G[ij] * g.ext * (a + b);
    }
    const double C = (nt > 1) ? __builtin_sqrt((2.0 * B.psi[ij]) / td_p) : Cp1;
    double en;
    const double t = ep_t(C, B.K[ij], B.G[ij], g.xij, g.ext, wij / m_p, wji / E.m[j], th_p, E.theta[j],
                          edp_in[slot], &en);
    edp_out[slot] = en;
    return t;
  };
  // t_jp seen from neighbour j as owner
  auto t_rev = [&](int j, int tj, uint32_t slot, const Geom& g) -> double {
    if (KIND == PD_BB) return g.s * c11;
    const int ji = tj + nt * tp;
    const double wij = 1.0 / g.xij;
    const double wji = 1.0 / pd_mag(-g.X1, -g.X2, -g.X3);
    if (KIND == PD_OSB) {
      const double a = wij / E.m[j], b = wji / m_p;
      return (3.0 * B.K[ji] - 5.0 * B.G[ji]) * (E.theta[j] * g.xij * a + th_p * g.xij * b) +
             15.0 * B.G[ji] * g.ext * (a + b);
    }
    const double C = __builtin_sqrt((2.0 * B.psi[ji]) / E.td[j]);
    double en;
    return ep_t(C, B.K[ji], B.G[ji], g.xij, g.ext, wij / E.m[j], wji / m_p, E.theta[j], th_p,
                edp_in[E.rev[slot]], &en);
  };
  // phase A: reverse terms from neighbours j < p, ascending j
  int kmid = 0;
  for (; kmid < d; ++kmid) {
    const uint32_t slot = s0 + 64u * (uint32_t)kmid;
    const int j = col[slot];
    if (j > p) break;
    if (!st[slot]) continue;
    const int tj = (nt > 1) ? E.ptype[j] : 0;
    Geom g;
    const double4 xj = x[j];
    geom(xp, yp, xj, y[j], g);
    if (g.s < B.s0[tj + nt * tp]) {
      const double xx = 0.5 * t_rev(j, tj, slot, g) / g.yij;
      a1 += xx * g.Y1 * xj.w / den_p;
      a2 += xx * g.Y2 * xj.w / den_p;
      a3 += xx * g.Y3 * xj.w / den_p;
    }
  }
  // phase B: own row in family order (descending j); breaks bonds
  for (int k = d - 1; k >= 0; --k) {
    const uint32_t slot = s0 + 64u * (uint32_t)k;
    if (!st[slot]) continue;
    const int j = col[slot];
    const int tj = (nt > 1) ? E.ptype[j] : 0;
    Geom g;
    const double4 xj = x[j];
    geom(xp, yp, xj, y[j], g);
    if (g.s < B.s0[tp + nt * tj]) {
      const double xx = 0.5 * t_own(j, tj, slot, g) / g.yij;
      a1 += xx * g.Y1 * xj.w / den_p;
      a2 += xx * g.Y2 * xj.w / den_p;
      a3 += xx * g.Y3 * xj.w / den_p;
    } else {
      st[slot] = 0;
      if (KIND == PD_EP) edp_out[slot] = edp_in[slot];
    }
  }
  // phase C: reverse terms from neighbours j > p, ascending j
  for (int k = kmid; k < d; ++k) {
    const uint32_t slot = s0 + 64u * (uint32_t)k;
    if (!st[slot]) continue;
    const int j = col[slot];
    const int tj = (nt > 1) ? E.ptype[j] : 0;
    Geom g;
    const double4 xj = x[j];
    geom(xp, yp, xj, y[j], g);
    if (g.s < B.s0[tj + nt * tp]) {
      const double xx = 0.5 * t_rev(j, tj, slot, g) / g.yij;
      a1 += xx * g.Y1 * xj.w / den_p;
      a2 += xx * g.Y2 * xj.w / den_p;
      a3 += xx * g.Y3 * xj.w / den_p;
    }
  }
}

// FP64 fast form: partial sums of this lane's bonds (k = g, g+G, ...)
template <int KIND>
__device__ __forceinline__ void material_fast(const EnvDev& E, const BlockDev& B, const WaveCtx& w, int p,
                                              const double4& xp, const double4& yp, double& a1, double& a2,
                                              double& a3) {
  const double4* __restrict__ x = E.x;
  const double4* __restrict__ y = E.y[E.cur];
  const int32_t* __restrict__ col = E.col;
  uint8_t* __restrict__ st = E.state;
  const int nt = B.ntype;
  const int tp = E.ptype[p];
  const double inv_den = 1.0 / B.density[tp];
  const int d = w.active ? E.deg[p] : 0;
  const uint32_t base = w.ch.slot0 + w.lane;
  double4 ap = make_double4(0.0, 0.0, 0.0, 0.0);
  const double* __restrict__ edp_in = nullptr;
  double* __restrict__ edp_out = nullptr;
  if (KIND != PD_BB) ap = E.aux[p];
  bool yielded_p = false;  // p's plastic extensions are exactly 0 until it yields
  if (KIND == PD_EP) {
    edp_in = E.edp[E.edp_cur];
    edp_out = E.edp[E.edp_cur ^ 1];
    yielded_p = pd_has_yielded(ap);
  }
  const double c11 = B.c11;
  // Bonds in batches of U: all column/state (and reverse-slot) loads of the
  // batch first, then all neighbour gathers, then the arithmetic.  Loading the
  // next batch's streamed data ahead (as the theta / td passes do) measured
  // slower here: its registers cost these kernels a wave per SIMD.
  constexpr int U = (KIND == PD_EP) ? 1 : 2;
  const int ni = (d > w.g) ? ((d - w.g + w.G - 1) >> w.lgG) : 0;
  const double s0_pp = B.s0[tp + nt * tp];
  for (int i0 = 0; i0 < ni; i0 += U) {
    int jv[U];
    uint8_t sv[U];
    uint32_t rv[U];
    double eo[U], er[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const bool ok = (i0 + u) < ni;
      const uint32_t slot = base + 64u * (uint32_t)(i0 + u);
      jv[u] = ok ? gld(col, slot) : p;
      sv[u] = ok ? gld(st, slot) : (uint8_t)0;
      if (KIND == PD_EP) eo[u] = (ok && yielded_p) ? gld(edp_in, slot) : 0.0;
    }
    double4 xj[U], yj[U], aj[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      xj[u] = gld4(x, jv[u]);
      yj[u] = gld4(y, jv[u]);
      if (KIND != PD_BB) aj[u] = gld4(E.aux, jv[u]);
    }
    if (KIND == PD_EP) {
      // the reverse bond's plastic extension, 0 unless j has yielded
#pragma unroll
      for (int u = 0; u < U; ++u) {
        er[u] = 0.0;
        if (sv[u] && pd_has_yielded(aj[u])) {
          rv[u] = gld(E.rev, base + 64u * (uint32_t)(i0 + u));
          er[u] = gld(edp_in, rv[u]);
        }
      }
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      if (!sv[u]) continue;
      const uint32_t slot = base + 64u * (uint32_t)(i0 + u);
      const int j = jv[u];
      const double X1 = xj[u].x - xp.x, X2 = xj[u].y - xp.y, X3 = xj[u].z - xp.z;
      const double Y1 = yj[u].x - yp.x, Y2 = yj[u].y - yp.y, Y3 = yj[u].z - yp.z;
      const double xx2 = pd_norm2(X1, X2, X3);
      const double yy2 = pd_norm2(Y1, Y2, Y3);
      const double inv_x = pd_rsqrt(xx2);
      const double inv_y = pd_rsqrt(yy2);
      const double xij = xx2 * inv_x;
      const double yij = yy2 * inv_y;
      const double ext = yij - xij;
      const int tj = (nt > 1) ? gld(E.ptype, j) : tp;
      const int ij = tp + nt * tj;
      if (ext * inv_x < ((nt > 1) ? B.s0[ij] : s0_pp)) {
        double t;
        if (KIND == PD_BB) {
          t = ext * inv_x * c11;
        } else {
          const double sm = ap.y + aj[u].y;
          const double twm = ap.x * ap.y + aj[u].x * aj[u].y;
          if (KIND == PD_OSB) {
            t = (3.0 * B.K[ij] - 5.0 * B.G[ij]) * twm + 15.0 * B.G[ij] * ext * (inv_x * sm);
          } else {
            const double ed = ext - (twm * pd_rcp(sm)) * xij * (1.0 / 3.0);
            const double g15 = 15.0 * B.G[ij] * (inv_x * sm);
            double Cp = ap.w, Cj = aj[u].w;
            if (nt > 1) {
              Cp = __builtin_sqrt((2.0 * B.psi[ij]) / __builtin_fabs(ap.z));
              Cj = __builtin_sqrt((2.0 * B.psi[ij]) / __builtin_fabs(aj[u].z));
            }
            const double ede = ed - eo[u];
            double td_own = g15 * ede;
            double e_new = eo[u];
            if (!(Cp > 1.0)) {
              td_own *= Cp;
              e_new = eo[u] + ede * (1.0 - Cp);
            }
            if (yielded_p) gst(edp_out, slot, e_new);  // else e_new = 0 = stored
            double td_rev = g15 * (ed - er[u]);
            if (!(Cj > 1.0)) td_rev *= Cj;
            t = 3.0 * B.K[ij] * twm + 0.5 * (td_own + td_rev);
          }
        }
        const double xx = t * inv_y * (xj[u].w * inv_den);
        a1 = __builtin_fma(xx, Y1, a1);
        a2 = __builtin_fma(xx, Y2, a2);
        a3 = __builtin_fma(xx, Y3, a3);
      } else {
        gst(st, slot, (uint8_t)0);
        if (KIND == PD_EP && yielded_p) gst(edp_out, slot, eo[u]);
      }
    }
  }
}

template <int KIND, int PREC, bool GPOW>
__device__ __forceinline__ void force_body(const Chunk* __restrict__ chunks, int nchunks, EnvDev* __restrict__ envs,
                                           int flags) {
  WaveCtx w;
  if (!wave_chunk(chunks, nchunks, envs, w)) return;
  EnvDev& E = envs[w.ch.env];
  const BlockDev& B = E.blocks[w.ch.block];
  const int p = w.active ? w.p : w.ch.first;
  const double4* __restrict__ x = E.x;
  const double4* __restrict__ y = E.y[E.cur];
  const double4 xp = x[p], yp = y[p];
  double a1 = 0.0, a2 = 0.0, a3 = 0.0;
  if (PREC == PD_PREC_EXACT) {
    if (!w.active) return;
    if (B.deformed) material_exact<KIND>(E, B, w, xp, yp, a1, a2, a3);
  } else {
    if (B.deformed) {
      material_fast<KIND>(E, B, w, p, xp, yp, a1, a2, a3);
      a1 = group_sum(a1, w.C);
      a2 = group_sum(a2, w.C);
      a3 = group_sum(a3, w.C);
    }
    if (!w.active || w.g != 0) return;
  }
  add_contacts<GPOW>(E, p, yp, y, a1, a2, a3);
  finish_particle(E, p, xp, yp, a1, a2, a3, flags);
}

template <int KIND, int PREC, bool GPOW>
__global__ __launch_bounds__(256) void k_force(const Chunk* __restrict__ chunks, int nchunks,
                                               EnvDev* __restrict__ envs, int flags) {
  force_body<KIND, PREC, GPOW>(chunks, nchunks, envs, flags);
}

// FP64 form with integer contact exponents: held to at most 96 VGPRs (5 waves
// per SIMD), which the elastoplastic bond loop reaches without spilling (it
// would take 98 unconstrained, 4 waves); the other kinds use fewer anyway.
template <int KIND>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(5, 8))) void k_force_lean(
    const Chunk* __restrict__ chunks, int nchunks, EnvDev* __restrict__ envs, int flags) {
  force_body<KIND, PD_PREC_FP64, false>(chunks, nchunks, envs, flags);
}

// short_range_contact! of one contact model at the resident positions (RM12,
// contacts.jl:189-247; RM11, contacts.jl:357-380): its terms added one by one to
// acc, the caller's f in internal order, in the reference's single-threaded
// order.  One thread per particle.
template <bool GPOW>
__global__ __launch_bounds__(256) void k_contact_model(const EnvDev* __restrict__ envs, int slot, int cidx,
                                                       double4* __restrict__ acc) {
  const EnvDev& E = envs[slot];
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= E.N) return;
  const double4* __restrict__ y = E.y[E.cur];
  const double4 a = acc[p];
  double a1 = a.x, a2 = a.y, a3 = a.z;
  add_contact_model<GPOW>(E, E.contacts[cidx], p, y[p], y, a1, a2, a3);
  acc[p] = make_double4(a1, a2, a3, 0.0);
}

// Start of a run call: kick + drift + position BCs of step time_step + 1 from
// the stored force (velocity_verlet.jl:63-69).
__global__ __launch_bounds__(256) void k_prologue(const Chunk* __restrict__ chunks, int nchunks,
                                                  EnvDev* __restrict__ envs) {
  WaveCtx w;
  if (!wave_chunk(chunks, nchunks, envs, w)) return;
  EnvDev& E = envs[w.ch.env];
  if (!w.active || w.g != 0) return;
  const int p = w.p;
  const double dt = E.dt, c = dt / 2.0;
  const double4 yp = E.y[E.cur][p];
  const double4 f = E.f[p];
  double4 v = E.v[p];
  v.x = v.x + c * f.x; v.y = v.y + c * f.y; v.z = v.z + c * f.z;
  double4 yn = make_double4(yp.x + dt * v.x, yp.y + dt * v.y, yp.z + dt * v.z, 0.0);
  const int bci = E.bcid[p];
  if (bci >= 0) {
    const BcDev& bc = E.bcs[bci];
    if (bc.kind == PD_BC_FIX) {
      yn = yp;
    } else {
      // this drift belongs to step time_step + 1 (the fused kernels drift the
      // following step and use time_step + 1 there)
      const double ddir = dt * pd_bc_dir(bc, E.time_step);
      yn.x = yp.x + ddir * bc.rate[0]; yn.y = yp.y + ddir * bc.rate[1]; yn.z = yp.z + ddir * bc.rate[2];
    }
  }
  E.y[E.cur ^ 1][p] = yn;
  E.v[p] = v;
}

// Per-env bookkeeping between kernels of consecutive steps.
__global__ void k_book(EnvDev* __restrict__ envs, const int32_t* __restrict__ slots, int n, int flags) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  EnvDev& E = envs[slots[i]];
  if (flags & PD_B_FLIP_Y) E.cur ^= 1;
  if ((flags & PD_B_FLIP_EDP) && E.has_ep) E.edp_cur ^= 1;
  if (flags & PD_B_STEP) E.time_step += 1;
}

// Deformed latch (general_material.jl:243-255): isapprox(y .- y[:,1] .+ x[:,1], x)
// with rtol = sqrt(eps(Float64)).  Two kernels: k_latch_part sums |a - x|^2,
// |a|^2 and |x|^2 over the slices of each (env, block) entry, one workgroup per
// slice (the launch's units list every entry's slices); k_latch_fin adds the
// slices of an entry in a fixed order and sets the latch.  Latched blocks skip
// both.
__device__ __forceinline__ void block_sum3(double& a, double& b, double& c) {
  for (int o = 32; o > 0; o >>= 1) {
    a += __shfl_down(a, o);
    b += __shfl_down(b, o);
    c += __shfl_down(c, o);
  }
  __shared__ double red[3][4];
  const int wv = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) { red[0][wv] = a; red[1][wv] = b; red[2][wv] = c; }
  __syncthreads();
  if (threadIdx.x == 0) {
    a = (red[0][0] + red[0][1]) + (red[0][2] + red[0][3]);
    b = (red[1][0] + red[1][1]) + (red[1][2] + red[1][3]);
    c = (red[2][0] + red[2][1]) + (red[2][2] + red[2][3]);
  }
}

// A block's slices follow from its own size only (kLatchSlice particles each,
// two per thread, at most kLatchMaxSlices), so its sums, and the step it
// latches on, do not depend on the other entries of the launch: a specimen
// latches on the same step alone and batched.  Small slices keep every
// workgroup's dependent loads short: a block that never deforms (a rigid
// indenter moved by its BCs) is tested at every step.
constexpr int kLatchSlice = 256 * 2;
constexpr int kLatchMaxSlices = 16384;
__host__ __device__ __forceinline__ int latch_slices(int64_t n) {
  const int64_t s = (n + kLatchSlice - 1) / kLatchSlice;
  return (int)(s < 1 ? 1 : (s > kLatchMaxSlices ? kLatchMaxSlices : s));
}

__global__ __launch_bounds__(256) void k_latch_part(EnvDev* __restrict__ envs, const int2* __restrict__ items,
                                                    const int2* __restrict__ units, double4* __restrict__ part) {
  const int2 un = units[blockIdx.x];  // (entry, slice)
  const int2 it = items[un.x];
  const EnvDev& E = envs[it.x];
  const BlockDev& B = E.blocks[it.y];
  if (B.deformed) return;  // uniform per workgroup
  const int ns = latch_slices(B.n);
  const double4* __restrict__ y = E.y[E.cur];
  const double4* __restrict__ x = E.x;
  const double4 y1 = y[B.first], x1 = x[B.first];
  const int per = (B.n + ns - 1) / ns;
  const int lo = un.y * per;
  const int hi = min(B.n, lo + per);
  double sd = 0.0, sa = 0.0, sx = 0.0;
  for (int i = lo + (int)threadIdx.x; i < hi; i += 256) {
    const double4 yi = gld4(y, B.first + i), xi = gld4(x, B.first + i);
    const double a[3] = {(yi.x - y1.x) + x1.x, (yi.y - y1.y) + x1.y, (yi.z - y1.z) + x1.z};
    const double xx[3] = {xi.x, xi.y, xi.z};
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const double df = a[d] - xx[d];
      sd += df * df;
      sa += a[d] * a[d];
      sx += xx[d] * xx[d];
    }
  }
  block_sum3(sd, sa, sx);
  if (threadIdx.x == 0) part[blockIdx.x] = make_double4(sd, sa, sx, 0.0);
}

__global__ __launch_bounds__(64) void k_latch_fin(EnvDev* __restrict__ envs, const int2* __restrict__ items,
                                                  const int32_t* __restrict__ ustart, const double4* __restrict__ part) {
  const int2 it = items[blockIdx.x];
  EnvDev& E = envs[it.x];
  BlockDev& B = E.blocks[it.y];
  if (B.deformed) return;
  const int ns = latch_slices(B.n);
  const double4* __restrict__ mine = part + ustart[blockIdx.x];
  double D = 0.0, A = 0.0, X = 0.0;
  for (int q = threadIdx.x; q < ns; q += 64) {
    const double4 v = mine[q];
    D += v.x; A += v.y; X += v.z;
  }
  for (int o = 32; o > 0; o >>= 1) {
    D += __shfl_down(D, o);
    A += __shfl_down(A, o);
    X += __shfl_down(X, o);
  }
  if (threadIdx.x != 0) return;
  const double d = __builtin_sqrt(D);
  bool approx;
  if (__builtin_isfinite(d)) {
    const double rtol = 1.4901161193847656e-08;  // sqrt(eps(Float64))
    approx = d <= rtol * fmax(__builtin_sqrt(A), __builtin_sqrt(X));
  } else {
    approx = false;
  }
  if (!approx) B.deformed = 1;
}

// ---- observables (printdata, src/environment.jl:308-320) ---------------------
// Partial sums per chunk, then per env in chunk order (deterministic):
//   MODE 0: p = mass .* v (update_misc!, solvers.jl:289-291), mass = volume *
//           density[type] (environment.jl:173-180);
//   MODE 1: f, and damage = 1 - intact / intact0 (cal_damage, peri.jl:587-594).
__device__ __forceinline__ double wave_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

template <int MODE>
__global__ __launch_bounds__(256) void k_obs_part(const Chunk* __restrict__ chunks, int nchunks,
                                                  const EnvDev* __restrict__ envs, double4* __restrict__ part) {
  WaveCtx w;
  if (!wave_chunk(chunks, nchunks, envs, w)) return;
  const EnvDev& E = envs[w.ch.env];
  const BlockDev& B = E.blocks[w.ch.block];
  const int p = w.active ? w.p : w.ch.first;
  double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
  if (MODE == 0) {
    if (w.active && w.g == 0) {
      const double4 v = E.v[p];
      const double mass = E.x[p].w * B.density[E.ptype[p]];
      s0 = mass * v.x; s1 = mass * v.y; s2 = mass * v.z;
    }
  } else {
    const int d = w.active ? E.deg[p] : 0;
    const uint32_t base = w.ch.slot0 + w.lane;
    int cnt = 0;
    const int ni = (d > w.g) ? ((d - w.g + w.G - 1) >> w.lgG) : 0;
    for (int i = 0; i < ni; ++i) cnt += E.state[base + 64u * (uint32_t)i];
    for (int off = w.C; off < 64; off <<= 1) cnt += __shfl_xor(cnt, off);
    if (w.active && w.g == 0) {
      const double4 f = E.f[p];
      s0 = f.x; s1 = f.y; s2 = f.z;
      s3 = 1.0 - (double)cnt / (double)d;
    }
  }
  s0 = wave_sum(s0); s1 = wave_sum(s1); s2 = wave_sum(s2); s3 = wave_sum(s3);
  if (w.lane == 0) part[w.cid] = make_double4(s0, s1, s2, s3);
}

// One wavefront per env: the env's chunk partials [first, first + count) in
// order; MODE 0 writes row[1..3], MODE 1 row[4..7] of the env's 8-column row.
template <int MODE>
__global__ __launch_bounds__(64) void k_obs_fin(const double4* __restrict__ part, const int2* __restrict__ ranges,
                                                double* __restrict__ rows) {
  const int2 r = ranges[blockIdx.x];
  double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
  for (int i = threadIdx.x; i < r.y; i += 64) {
    const double4 v = part[r.x + i];
    s0 += v.x; s1 += v.y; s2 += v.z; s3 += v.w;
  }
  s0 = wave_sum(s0); s1 = wave_sum(s1); s2 = wave_sum(s2); s3 = wave_sum(s3);
  if (threadIdx.x != 0) return;
  double* row = rows + 8 * (size_t)blockIdx.x;
  if (MODE == 0) {
    row[1] = s0; row[2] = s1; row[3] = s2;
  } else {
    row[4] = s0; row[5] = s1; row[6] = s2; row[7] = s3;
  }
}

// One wavefront, one env: its chunk partials [0, nchunks) in order, as
// k_obs_fin sums them; MODE 1 also writes the time column (time_step * dt)
// and divides the damage sum by the particle count, the two columns the host
// finishes for pd_env_observables.
template <int MODE>
__global__ __launch_bounds__(64) void k_obs_fin_env(const double4* __restrict__ part, int nchunks,
                                                    const EnvDev* __restrict__ envs, int slot,
                                                    double* __restrict__ row) {
  double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
  for (int i = threadIdx.x; i < nchunks; i += 64) {
    const double4 v = part[i];
    s0 += v.x; s1 += v.y; s2 += v.z; s3 += v.w;
  }
  s0 = wave_sum(s0); s1 = wave_sum(s1); s2 = wave_sum(s2); s3 = wave_sum(s3);
  if (threadIdx.x != 0) return;
  if (MODE == 0) {
    row[1] = s0; row[2] = s1; row[3] = s2;
  } else {
    const EnvDev& E = envs[slot];
    row[0] = (double)E.time_step * E.dt;
    row[4] = s0; row[5] = s1; row[6] = s2;
    row[7] = s3 / (double)E.N;
  }
}

// ---- quasi-static minimiser (QSDrag minimize!, src/solvers/minimize.jl:105-231) --
// apply_bc!(:position) and (:velocity) of every BC once, as minimize! does
// before iterating (the last BC listed for a particle wins, as in list order).
__global__ __launch_bounds__(256) void k_qs_bc(const Chunk* __restrict__ chunks, int nchunks,
                                               EnvDev* __restrict__ envs) {
  WaveCtx w;
  if (!wave_chunk(chunks, nchunks, envs, w)) return;
  EnvDev& E = envs[w.ch.env];
  if (!w.active || w.g != 0) return;
  const int p = w.p;
  const int bci = E.bcid[p];
  if (bci < 0) return;
  const BcDev& bc = E.bcs[bci];
  double4 v = E.v[p];
  if (bc.kind == PD_BC_FIX) {
    v.x = 0.0; v.y = 0.0; v.z = 0.0;           // y = last = y
  } else {
    const double dir = pd_bc_dir(bc, E.time_step);
    const double ddir = E.dt * dir;             // last + dt * direction * rate
    double4 y = E.y[E.cur][p];
    y.x = y.x + ddir * bc.rate[0]; y.y = y.y + ddir * bc.rate[1]; y.z = y.z + ddir * bc.rate[2];
    E.y[E.cur][p] = y;
    v.x = dir * bc.rate[0]; v.y = dir * bc.rate[1]; v.z = dir * bc.rate[2];
  }
  E.v[p] = v;
}

// Julia's min / max on Float64 (NaN if either argument is NaN); a is finite.
__device__ __forceinline__ double jl_clip(double x, double a) {
  if (__builtin_isnan(x)) return x;
  const double lo = x < a ? x : a;              // min.(x, a)
  return lo > -a ? lo : -a;                     // max.(., -a)
}
__device__ __forceinline__ unsigned long long abs_bits(double x) {
  return (unsigned long long)__double_as_longlong(__builtin_fabs(x));  // orders like |x|, NaN above all
}

// One minimize! iteration on the free particles (no BC) after update_acc!:
//   force = f - (lambda * opt) * (1 + k |opt|),  k = lambda / 10
//   dy = clip((0.5 * force) * dt^2 + opt * dt, ps/10);  opt += dt * force;  y += dy
// (the reference's clipv result is not stored back into opt: apply_! rebinds
// a local).  maxes[0] / [1] take max |f| and max |dy| as non-negative bits.
__global__ __launch_bounds__(256) void k_qs_update(const Chunk* __restrict__ chunks, int nchunks,
                                                   EnvDev* __restrict__ envs, double4* __restrict__ opt,
                                                   double dt, double dt2, double lambda, double kk, double a,
                                                   unsigned long long* __restrict__ maxes) {
  WaveCtx w;
  if (!wave_chunk(chunks, nchunks, envs, w)) return;
  EnvDev& E = envs[w.ch.env];
  unsigned long long mf = 0, mx = 0;
  if (w.active && w.g == 0 && !(E.qheld ? E.qheld[w.p] != 0 : E.bcid[w.p] >= 0)) {
    const int p = w.p;
    const double4 f = E.f[p];
    double4 o = opt[p];
    double4 y = E.y[E.cur][p];
    const double fc[3] = {f.x, f.y, f.z};
    double oc[3] = {o.x, o.y, o.z};
    double yc[3] = {y.x, y.y, y.z};
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const double force = fc[c] - (lambda * oc[c]) * (1.0 + kk * __builtin_fabs(oc[c]));
      double dy = (0.5 * force) * dt2 + oc[c] * dt;
      oc[c] = oc[c] + dt * force;
      dy = jl_clip(dy, a);
      yc[c] = yc[c] + dy;
      mf = max(mf, abs_bits(fc[c]));
      mx = max(mx, abs_bits(dy));
    }
    opt[p] = make_double4(oc[0], oc[1], oc[2], 0.0);
    E.y[E.cur][p] = make_double4(yc[0], yc[1], yc[2], y.w);
  }
  for (int off = 32; off > 0; off >>= 1) {
    mf = max(mf, (unsigned long long)__shfl_xor((long long)mf, off));
    mx = max(mx, (unsigned long long)__shfl_xor((long long)mx, off));
  }
  if (w.lane == 0) {
    atomicMax(&maxes[0], mf);
    atomicMax(&maxes[1], mx);
  }
}

// 1/m into aux.y after the weighted volume is built
__global__ void k_aux_init(const double* __restrict__ m, double4* __restrict__ aux, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  aux[i] = make_double4(0.0, 1.0 / m[i], 0.0, 0.0);
}

// ---- launch wrappers --------------------------------------------------------
static inline int grid_for(int nchunks) {
  return (nchunks + PD_WAVES_PER_BLOCK - 1) / PD_WAVES_PER_BLOCK;
}

void pd_launch_theta(hipStream_t s, int prec, const Chunk* chunks, int n, EnvDev* envs) {
  if (n <= 0) return;
  if (prec == PD_PREC_FP32C) {
    pd_launch_theta32(s, chunks, n, envs);
    return;
  }
  if (prec == PD_PREC_EXACT)
    hipLaunchKernelGGL(k_theta<PD_PREC_EXACT>, dim3(grid_for(n)), dim3(256), 0, s, chunks, n, envs);
  else
    hipLaunchKernelGGL(k_theta<PD_PREC_FP64>, dim3(grid_for(n)), dim3(256), 0, s, chunks, n, envs);
}

void pd_launch_td(hipStream_t s, int prec, const Chunk* chunks, int n, EnvDev* envs) {
  if (n <= 0) return;
  if (prec == PD_PREC_FP32C) {
    pd_launch_td32(s, chunks, n, envs);
    return;
  }
  if (prec == PD_PREC_EXACT)
    hipLaunchKernelGGL(k_td<PD_PREC_EXACT>, dim3(grid_for(n)), dim3(256), 0, s, chunks, n, envs);
  else
    hipLaunchKernelGGL(k_td<PD_PREC_FP64>, dim3(grid_for(n)), dim3(256), 0, s, chunks, n, envs);
}

template <int KIND>
static void launch_force_kind(hipStream_t s, int prec, const Chunk* chunks, int n, EnvDev* envs, int flags,
                              bool gpow) {
  const dim3 g(grid_for(n)), b(256);
  if (prec == PD_PREC_EXACT)
    hipLaunchKernelGGL((k_force<KIND, PD_PREC_EXACT, true>), g, b, 0, s, chunks, n, envs, flags);
  else if (gpow)
    hipLaunchKernelGGL((k_force<KIND, PD_PREC_FP64, true>), g, b, 0, s, chunks, n, envs, flags);
  else
    hipLaunchKernelGGL((k_force_lean<KIND>), g, b, 0, s, chunks, n, envs, flags);
}

void pd_launch_force(hipStream_t s, int kind, int prec, const Chunk* chunks, int n, EnvDev* envs, int flags,
                     bool gpow) {
  if (n <= 0) return;
  if (prec == PD_PREC_FP32C) {
    pd_launch_force32(s, kind, chunks, n, envs, flags, gpow);
    return;
  }
  if (kind == PD_BB) launch_force_kind<PD_BB>(s, prec, chunks, n, envs, flags, gpow);
  else if (kind == PD_OSB) launch_force_kind<PD_OSB>(s, prec, chunks, n, envs, flags, gpow);
  else launch_force_kind<PD_EP>(s, prec, chunks, n, envs, flags, gpow);
}

// ---- contact finish of tiled envs ------------------------------------------
// The listed envs (slots[0 .. n)) number their particles with contact terms:
// counts zeroed, then every particle flagged and the flagged ones appended to
// cpart (any order: each particle is finished on its own).  starts[k] is the
// first workgroup of env k (a prefix over ceil(N / 256)).
__global__ void k_cpart_zero(EnvDev* __restrict__ envs, const int32_t* __restrict__ slots, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) *envs[slots[i]].ccount = 0;
}

__global__ __launch_bounds__(256) void k_cpart_mark(EnvDev* __restrict__ envs, const int32_t* __restrict__ slots,
                                                    const int32_t* __restrict__ starts, int n) {
  int lo = 0, hi = n;  // starts[lo] <= blockIdx.x < starts[lo + 1]
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (starts[mid] <= (int)blockIdx.x) lo = mid;
    else hi = mid;
  }
  EnvDev& E = envs[slots[lo]];
  const int p = ((int)blockIdx.x - starts[lo]) * 256 + (int)threadIdx.x;
  const bool in = p < E.N;
  const bool has = in && pd_has_contact_terms(E, p);
  if (in) E.cflag[p] = has ? 1 : 0;
  // one counter atomic per wavefront; its flagged lanes take consecutive entries
  const uint64_t m = __builtin_amdgcn_ballot_w64(has);
  if (!m) return;
  const int leader = __builtin_ctzll(m);
  int base = 0;
  if (pd_lane() == leader) base = atomicAdd(E.ccount, __builtin_popcountll(m));
  base = __shfl(base, leader);
  if (!has) return;
  const int at = base + lanes_below(m);
  E.cpart[at] = p;
  if (E.crec) {
    const ContactDev& C = E.contacts[0];
    const int i1 = C.inv1[p], i2 = C.inv2[p];
    int4* r = (int4*)(E.crec + 8 * (size_t)at);
    r[0] = make_int4(p, i1, i1 >= 0 ? C.own_cnt[i1] : 0, i2);
    r[1] = make_int4(i2 >= 0 ? C.rev_cnt[i2] : 0, i2 >= 0 ? C.rev_off[i2] : 0, 0, 0);
  }
}

// The contact terms (model order, list order) and the Verlet update of the
// particles the tiled force kernels left unfinished: f holds their bond sums.
// The grid's wavefronts stride over the concatenation of the listed envs'
// cpart lists (a wavefront-wide prefix of their counts maps an entry to its
// env), in one of two shapes, both summing every particle's terms in list
// order, so both give add_contacts' own sums:
//  * few particles (fewer than lane_min in all): one wavefront per particle;
//    the lanes form the terms of 64 list entries at once, then every lane adds
//    them in list order (readlane).  Every particle's gathers are in flight
//    together, which is what a short list needs;
//  * many (deep contact: tens of thousands of particles, tens of entries
//    each): one lane per particle summing its own lists, 64 particles per
//    wavefront, about a ninth of the instructions per list entry.

__device__ __forceinline__ double readlane_d(double v, int l) {
  const long long b = __builtin_bit_cast(long long, v);
  const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffll), l);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
  return __builtin_bit_cast(double, ((long long)hi << 32) | (long long)(unsigned)lo);
}

// Adds (sign +1) or subtracts (-1) the terms t(k) = term(y_q(k), vol_q(k)) for
// k = 0 .. n-1 in order; before entry `at` (0 <= at <= n) it runs ins() once.
template <class Q, class T, class I>
__device__ __forceinline__ void wave_list_terms(const double4* __restrict__ x, const double4* __restrict__ y, int n,
                                                int at, Q&& q, T&& term, I&& ins, bool sub, double& a1,
                                                double& a2, double& a3) {
  const int lane = pd_lane();
  for (int k0 = 0; k0 < n; k0 += 64) {
    const int k = k0 + lane;
    double t1 = 0.0, t2 = 0.0, t3 = 0.0;
    if (k < n) {
      const int qq = q(k);
      const double4 yq = y[qq];
      term(yq, x[qq].w, t1, t2, t3);
    }
    const int m = n - k0 < 64 ? n - k0 : 64;
    for (int u = 0; u < m; ++u) {
      if (k0 + u == at) ins();
      const double u1 = readlane_d(t1, u), u2 = readlane_d(t2, u), u3 = readlane_d(t3, u);
      if (sub) {
        a1 -= u1; a2 -= u2; a3 -= u3;
      } else {
        a1 += u1; a2 += u2; a3 += u3;
      }
    }
  }
  if (at >= n) ins();
}

// The same for the entry at of an env with one pair model, from its record
// (k_cpart_mark): the own list, then the transposed list, as one run of
// n1 + n2 entries whose gathers are in flight together (entries below n1
// added, the others subtracted, in that order).
__device__ __forceinline__ void contact_finish_rec(EnvDev& E, int at, int flags) {
  const int4* r = (const int4*)(E.crec + 8 * (size_t)at);
  const int4 r0 = r[0], r1 = r[1];
  const int p = r0.x, i1 = r0.y, n1 = r0.z, n2 = r1.x, off2 = r1.y;
  const double4* __restrict__ x = E.x;
  const double4* __restrict__ y = E.y[E.cur];
  const double4 yp = y[p];
  const double4 f = E.f[p];
  const double4 xp = x[p];
  const double4 vp = E.v[p];
  const int bci = E.bcid[p];
  const double inv_den_p = E.inv_den[p];
  const ContactDev& C = E.contacts[0];
  const int32_t* __restrict__ own = C.own + (size_t)(i1 >= 0 ? i1 : 0) * C.max_neighs;
  double a1 = f.x, a2 = f.y, a3 = f.z;
  const int nn = n1 + n2;
  for (int k0 = 0; k0 < nn; k0 += 64) {
    const int k = k0 + pd_lane();
    double t1 = 0.0, t2 = 0.0, t3 = 0.0;
    if (k < nn) {
      const bool mine = k < n1;
      const int q = mine ? C.idx2[own[k]] : C.idx1[C.rev[off2 + (k - n1)]];
      const double4 yq = y[q];
      const double a = x[q].w * inv_den_p;
      double F1, F2, F3;
      if (mine) contact_fn<false>(C, yp.x - yq.x, yp.y - yq.y, yp.z - yq.z, F1, F2, F3);
      else contact_fn<false>(C, yq.x - yp.x, yq.y - yp.y, yq.z - yp.z, F1, F2, F3);
      t1 = F1 * a; t2 = F2 * a; t3 = F3 * a;
    }
    const int m = nn - k0 < 64 ? nn - k0 : 64;
    for (int u = 0; u < m; ++u) {
      const double u1 = readlane_d(t1, u), u2 = readlane_d(t2, u), u3 = readlane_d(t3, u);
      if (k0 + u < n1) {
        a1 += u1; a2 += u2; a3 += u3;
      } else {
        a1 -= u1; a2 -= u2; a3 -= u3;
      }
    }
  }
  if (pd_lane() == 0) finish_particle_v(E, p, xp, yp, a1, a2, a3, flags, vp, bci);
}

// Contact terms and Verlet update of particle p of E (every lane of the
// wavefront calls it with the same p).
__device__ __forceinline__ void contact_finish_particle(EnvDev& E, int p, int flags) {
  const double4* __restrict__ x = E.x;
  const double4* __restrict__ y = E.y[E.cur];
  auto none = [] {};
  const double4 yp = y[p];
  const double4 f = E.f[p];
  double a1 = f.x, a2 = f.y, a3 = f.z;
  for (int cm = 0; cm < E.ncontacts; ++cm) {
    const ContactDev& C = E.contacts[cm];
    const int i1 = C.inv1[p];
    if (C.self_contact) {
      // add_self_contact: reverse entries of members i' < i1, the own term, the rest
      if (i1 < 0) continue;
      const double vol_p = x[p].w;
      const double den_p = (vol_p * E.blocks[E.pblock[p]].density[E.ptype[p]]) / vol_p;
      const int off = C.rev_off[i1], n = C.rev_cnt[i1];
      const double b = vol_p / den_p;
      int split = 0;  // entries of members < i1 (the list is in member order)
      for (int k0 = 0; k0 < n; k0 += 64) {
        const int k = k0 + pd_lane();
        const bool lt = k < n && C.rev[off + k] < i1;
        split += __builtin_popcountll(__builtin_amdgcn_ballot_w64(lt));
      }
      auto own = [&] {
        if (!C.own_cnt[i1]) return;
        const int q = C.idx1[C.own[(size_t)i1 * C.max_neighs]];
        const double4 yq = y[q];
        double F1, F2, F3;
        contact_fn<false>(C, yp.x - yq.x, yp.y - yq.y, yp.z - yq.z, F1, F2, F3);
        const double a = x[q].w / den_p;
        a1 += F1 * a; a2 += F2 * a; a3 += F3 * a;
      };
      wave_list_terms(
          x, y, n, split, [&](int k) { return C.idx1[C.rev[off + k]]; },
          [&](const double4& yq, double, double& t1, double& t2, double& t3) {
            double F1, F2, F3;
            contact_fn<false>(C, yq.x - yp.x, yq.y - yp.y, yq.z - yp.z, F1, F2, F3);
            t1 = F1 * b; t2 = F2 * b; t3 = F3 * b;
          },
          own, true, a1, a2, a3);
      continue;
    }
    // add_contact_model: own list, then the transposed list
    const int i2 = C.inv2[p];
    if (i1 >= 0) {
      const double inv_den_p = E.inv_den[p];
      wave_list_terms(
          x, y, C.own_cnt[i1], 1 << 30, [&](int k) { return C.idx2[C.own[(size_t)i1 * C.max_neighs + k]]; },
          [&](const double4& yq, double vq, double& t1, double& t2, double& t3) {
            double F1, F2, F3;
            contact_fn<false>(C, yp.x - yq.x, yp.y - yq.y, yp.z - yq.z, F1, F2, F3);
            const double a = vq * inv_den_p;
            t1 = F1 * a; t2 = F2 * a; t3 = F3 * a;
          },
          none, false, a1, a2, a3);
    }
    if (i2 >= 0) {
      const double inv_den_p = E.inv_den[p];
      const int off = C.rev_off[i2];
      wave_list_terms(
          x, y, C.rev_cnt[i2], 1 << 30, [&](int k) { return C.idx1[C.rev[off + k]]; },
          [&](const double4& yq, double vq, double& t1, double& t2, double& t3) {
            double F1, F2, F3;
            contact_fn<false>(C, yq.x - yp.x, yq.y - yp.y, yq.z - yp.z, F1, F2, F3);
            const double bq = vq * inv_den_p;
            t1 = F1 * bq; t2 = F2 * bq; t3 = F3 * bq;
          },
          none, true, a1, a2, a3);
    }
  }
  if (pd_lane() == 0) finish_particle(E, p, x[p], yp, a1, a2, a3, flags);  // tiled envs: integer exponents
}

// One pair-model list of particle p on its own lane: the terms of entries
// k = 0 .. n-1 (partner q(k)), added (sub: subtracted) in list order, with the
// gathers of 4 entries in flight at once.  term(yq, vol_q) -> (t1, t2, t3).
template <class Q, class T>
__device__ __forceinline__ void lane_list_terms(const double4* __restrict__ x, const double4* __restrict__ y, int n,
                                                int pself, Q&& q, T&& term, bool sub, double& a1, double& a2,
                                                double& a3) {
  for (int k = 0; k < n; k += 4) {
    int qs[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) qs[u] = (k + u < n) ? q(k + u) : pself;
    double4 yq[4];
    double vq[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      yq[u] = y[qs[u]];
      vq[u] = x[qs[u]].w;
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      if (k + u >= n) break;
      double t1, t2, t3;
      term(yq[u], vq[u], t1, t2, t3);
      if (sub) {
        a1 -= t1; a2 -= t2; a3 -= t3;
      } else {
        a1 += t1; a2 += t2; a3 += t3;
      }
    }
  }
}

// add_contacts (integer exponents) for particle p on its own lane, the pair
// models' lists through lane_list_terms: the same terms in the same order.
__device__ __forceinline__ void lane_contact_terms(const EnvDev& E, int p, const double4& yp,
                                                   const double4* __restrict__ y, double& a1, double& a2,
                                                   double& a3) {
  const double4* __restrict__ x = E.x;
  for (int cm = 0; cm < E.ncontacts; ++cm) {
    const ContactDev& C = E.contacts[cm];
    if (C.self_contact) {
      add_self_contact<false>(E, C, p, yp, y, a1, a2, a3);
      continue;
    }
    const int i1 = C.inv1[p], i2 = C.inv2[p];
    if (i1 >= 0) {
      const double inv_den_p = E.inv_den[p];
      const int32_t* __restrict__ own = C.own + (size_t)i1 * C.max_neighs;
      lane_list_terms(
          x, y, C.own_cnt[i1], p, [&](int k) { return C.idx2[own[k]]; },
          [&](const double4& yq, double vq, double& t1, double& t2, double& t3) {
            double F1, F2, F3;
            contact_fn<false>(C, yp.x - yq.x, yp.y - yq.y, yp.z - yq.z, F1, F2, F3);
            const double a = vq * inv_den_p;
            t1 = F1 * a; t2 = F2 * a; t3 = F3 * a;
          },
          false, a1, a2, a3);
    }
    if (i2 >= 0) {
      const double inv_den_p = E.inv_den[p];
      const int32_t* __restrict__ rev = C.rev + C.rev_off[i2];
      lane_list_terms(
          x, y, C.rev_cnt[i2], p, [&](int k) { return C.idx1[rev[k]]; },
          [&](const double4& yq, double vq, double& t1, double& t2, double& t3) {
            double F1, F2, F3;
            contact_fn<false>(C, yq.x - yp.x, yq.y - yp.y, yq.z - yp.z, F1, F2, F3);
            const double bq = vq * inv_den_p;
            t1 = F1 * bq; t2 = F2 * bq; t3 = F3 * bq;
          },
          true, a1, a2, a3);
    }
  }
}

// lane_contact_terms for an env with one pair model, from the record of its
// cpart entry at (k_cpart_mark); returns p.
__device__ __forceinline__ int lane_contact_rec(const EnvDev& E, int at, const double4* __restrict__ y,
                                                double4& yp, double& a1, double& a2, double& a3) {
  const int4* r = (const int4*)(E.crec + 8 * (size_t)at);
  const int4 r0 = r[0], r1 = r[1];
  const int p = r0.x, i1 = r0.y, n1 = r0.z, n2 = r1.x, off2 = r1.y;
  const double4* __restrict__ x = E.x;
  yp = y[p];
  const double4 f = E.f[p];
  const double inv_den_p = E.inv_den[p];
  const ContactDev& C = E.contacts[0];
  a1 = f.x; a2 = f.y; a3 = f.z;
  const int32_t* __restrict__ own = C.own + (size_t)(i1 >= 0 ? i1 : 0) * C.max_neighs;
  const int32_t* __restrict__ rev = C.rev + off2;
  const double4 ypp = yp;
  lane_list_terms(
      x, y, n1, p, [&](int k) { return C.idx2[own[k]]; },
      [&](const double4& yq, double vq, double& t1, double& t2, double& t3) {
        double F1, F2, F3;
        contact_fn<false>(C, ypp.x - yq.x, ypp.y - yq.y, ypp.z - yq.z, F1, F2, F3);
        const double a = vq * inv_den_p;
        t1 = F1 * a; t2 = F2 * a; t3 = F3 * a;
      },
      false, a1, a2, a3);
  lane_list_terms(
      x, y, n2, p, [&](int k) { return C.idx1[rev[k]]; },
      [&](const double4& yq, double vq, double& t1, double& t2, double& t3) {
        double F1, F2, F3;
        contact_fn<false>(C, yq.x - ypp.x, yq.y - ypp.y, yq.z - ypp.z, F1, F2, F3);
        const double bq = vq * inv_den_p;
        t1 = F1 * bq; t2 = F2 * bq; t3 = F3 * bq;
      },
      true, a1, a2, a3);
  return p;
}

// Total entries of the cpart lists of envs slots[0 .. n) (wave-uniform).
__device__ __forceinline__ int contact_total(const EnvDev* __restrict__ envs, const int32_t* __restrict__ slots,
                                             int n) {
  int total = 0;
  for (int k0 = 0; k0 < n; k0 += 64) {
    const int k = k0 + pd_lane();
    int cnt = k < n ? *envs[slots[k]].ccount : 0;
    for (int off = 32; off > 0; off >>= 1) cnt += __shfl_xor(cnt, off);
    total += cnt;
  }
  return total;
}

// Entry t of the concatenated cpart lists of envs slots[0 .. n), one per lane
// (t may differ between lanes; every lane must call it): its env index (into
// slots) and its position in that env's list; -1 past the last entry.
__device__ __forceinline__ int contact_entry_env(const EnvDev* __restrict__ envs, const int32_t* __restrict__ slots,
                                                 int n, int t, int& pos) {
  const int lane = pd_lane();
  int found = -1;
  pos = 0;
  for (int k0 = 0; k0 < n; k0 += 64) {
    const int k = k0 + lane;
    const int cnt = k < n ? *envs[slots[k]].ccount : 0;
    int incl = cnt;  // inclusive prefix over this group of envs
    for (int off = 1; off < 64; off <<= 1) {
      const int v = __shfl_up(incl, off);
      if (lane >= off) incl += v;
    }
    const int total = __shfl(incl, 63);
    // first lane of the group whose prefix exceeds t (64 if none)
    int lo = 0, hi = 64;
    for (int step = 0; step < 7; ++step) {
      const int mid = (lo + hi) >> 1;
      const int v = __shfl(incl, mid < 64 ? mid : 63);
      if (lo < hi) {
        if (v > t) hi = mid;
        else lo = mid + 1;
      }
    }
    const int excl = __shfl(incl - cnt, lo < 64 ? lo : 63);
    if (found < 0 && t >= 0 && t < total) {
      found = k0 + lo;
      pos = t - excl;
    }
    t -= total;
  }
  return found;
}

// Up to kCfinEnvs envs, every workgroup first puts the prefix of their counts
// in LDS, so that an entry finds its env without global loads.
constexpr int kCfinEnvs = 512;

__global__ __launch_bounds__(256) void k_contact_finish(EnvDev* __restrict__ envs, const int32_t* __restrict__ slots,
                                                        int n, int lane_min, int flags) {
  const int nwaves = (int)gridDim.x * PD_WAVES_PER_BLOCK;
  const int w0 = (int)blockIdx.x * PD_WAVES_PER_BLOCK + pd_wave_id();
  __shared__ int32_t pre[kCfinEnvs + 1];  // pre[k] = entries of envs 0 .. k-1
  const bool lds = n <= kCfinEnvs;
  if (lds) {
    if (threadIdx.x < 64) {
      const int lane = pd_lane();
      int run = 0;
      for (int k0 = 0; k0 < n; k0 += 64) {
        const int k = k0 + lane;
        const int cnt = k < n ? *envs[slots[k]].ccount : 0;
        int incl = cnt;
        for (int off = 1; off < 64; off <<= 1) {
          const int v = __shfl_up(incl, off);
          if (lane >= off) incl += v;
        }
        if (k < n) pre[k + 1] = run + incl;
        run += __shfl(incl, 63);
      }
      if (lane == 0) pre[0] = 0;
    }
    __syncthreads();
  }
  // env of entry t (0 <= t < total, lds only): the k with pre[k] <= t < pre[k + 1]
  auto find = [&](int t, int& pos) {
    int lo = 0, hi = n - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (pre[mid] <= t) lo = mid;
      else hi = mid - 1;
    }
    pos = t - pre[lo];
    return lo;
  };
  const int total = lds ? pre[n] : contact_total(envs, slots, n);
  if (total >= lane_min) {
    for (int t0 = w0 * 64; t0 < total; t0 += nwaves * 64) {
      const int t = t0 + pd_lane();
      int pos, k;
      if (lds) {
        if (t >= total) continue;
        k = find(t, pos);
      } else {  // every lane searches (shuffles); past the end k = -1
        k = contact_entry_env(envs, slots, n, t, pos);
        if (k < 0) continue;
      }
      EnvDev& E = envs[slots[k]];
      const double4* __restrict__ y = E.y[E.cur];
      double4 yp;
      double a1, a2, a3;
      int p;
      if (E.crec) {
        p = lane_contact_rec(E, pos, y, yp, a1, a2, a3);
      } else {
        p = E.cpart[pos];
        yp = y[p];
        const double4 f = E.f[p];
        a1 = f.x; a2 = f.y; a3 = f.z;
        lane_contact_terms(E, p, yp, y, a1, a2, a3);
      }
      finish_particle(E, p, E.x[p], yp, a1, a2, a3, flags);  // tiled envs: integer exponents
    }
    return;
  }
  for (int t = w0; t < total; t += nwaves) {
    int pos;
    const int k = lds ? find(t, pos) : contact_entry_env(envs, slots, n, t, pos);
    EnvDev& E = envs[slots[k]];
    if (E.crec) contact_finish_rec(E, pos, flags);
    else contact_finish_particle(E, E.cpart[pos], flags);
  }
}

void pd_launch_cpart_build(hipStream_t s, EnvDev* envs, const int32_t* slots, const int32_t* starts, int n,
                           int nblocks) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_cpart_zero, dim3((n + 255) / 256), dim3(256), 0, s, envs, slots, n);
  if (nblocks > 0) hipLaunchKernelGGL(k_cpart_mark, dim3(nblocks), dim3(256), 0, s, envs, slots, starts, n);
}

void pd_launch_contact_finish(hipStream_t s, EnvDev* envs, const int32_t* slots, int n, int nblocks, int lane_min,
                              int flags) {
  if (n <= 0 || nblocks <= 0) return;
  hipLaunchKernelGGL(k_contact_finish, dim3(nblocks), dim3(256), 0, s, envs, slots, n, lane_min, flags);
}

void pd_launch_contact_model(hipStream_t s, const EnvDev* envs, int slot, int cidx, bool gpow, double4* acc, int n) {
  if (n <= 0) return;
  const dim3 g((n + 255) / 256), b(256);
  if (gpow) hipLaunchKernelGGL(k_contact_model<true>, g, b, 0, s, envs, slot, cidx, acc);
  else hipLaunchKernelGGL(k_contact_model<false>, g, b, 0, s, envs, slot, cidx, acc);
}

void pd_launch_prologue(hipStream_t s, const Chunk* chunks, int n, EnvDev* envs) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_prologue, dim3(grid_for(n)), dim3(256), 0, s, chunks, n, envs);
}

void pd_launch_book(hipStream_t s, EnvDev* envs, const int32_t* slots, int n, int flags) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_book, dim3((n + 255) / 256), dim3(256), 0, s, envs, slots, n, flags);
}

int pd_latch_slices(int64_t n) { return latch_slices(n); }

// deformed latch of each (env slot, block) item, for one readback
__global__ void k_latch_flags(const EnvDev* __restrict__ envs, const int2* __restrict__ items, int n,
                              int32_t* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = envs[items[i].x].blocks[items[i].y].deformed;
}

void pd_launch_copy_words(hipStream_t s, const PdCopy* d, int n) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_copy_words, dim3(n), dim3(256), 0, s, d, n);
}

void pd_launch_latch_flags(hipStream_t s, EnvDev* envs, const int2* items, int n, int32_t* out) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_latch_flags, dim3((n + 255) / 256), dim3(256), 0, s, envs, items, n, out);
}

void pd_launch_latch(hipStream_t s, EnvDev* envs, const int2* items, int n, const int2* units, int nunits,
                     const int32_t* ustart, double4* part) {
  if (n <= 0 || nunits <= 0) return;
  hipLaunchKernelGGL(k_latch_part, dim3(nunits), dim3(256), 0, s, envs, items, units, part);
  hipLaunchKernelGGL(k_latch_fin, dim3(n), dim3(64), 0, s, envs, items, ustart, (const double4*)part);
}

void pd_launch_observables(hipStream_t s, int mode, const Chunk* chunks, int n, EnvDev* envs, double4* part,
                           const int2* ranges, int nenv, double* rows) {
  if (n <= 0 || nenv <= 0) return;
  if (mode == 0) {
    hipLaunchKernelGGL(k_obs_part<0>, dim3(grid_for(n)), dim3(256), 0, s, chunks, n, envs, part);
    hipLaunchKernelGGL(k_obs_fin<0>, dim3(nenv), dim3(64), 0, s, (const double4*)part, ranges, rows);
  } else {
    hipLaunchKernelGGL(k_obs_part<1>, dim3(grid_for(n)), dim3(256), 0, s, chunks, n, envs, part);
    hipLaunchKernelGGL(k_obs_fin<1>, dim3(nenv), dim3(64), 0, s, (const double4*)part, ranges, rows);
  }
}

void pd_launch_observables_env(hipStream_t s, const Chunk* chunks, int n, EnvDev* envs, int slot, double4* part,
                               double* row) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_obs_part<0>, dim3(grid_for(n)), dim3(256), 0, s, chunks, n, envs, part);
  hipLaunchKernelGGL(k_obs_fin_env<0>, dim3(1), dim3(64), 0, s, (const double4*)part, n, envs, slot, row);
  hipLaunchKernelGGL(k_obs_part<1>, dim3(grid_for(n)), dim3(256), 0, s, chunks, n, envs, part);
  hipLaunchKernelGGL(k_obs_fin_env<1>, dim3(1), dim3(64), 0, s, (const double4*)part, n, envs, slot, row);
}

void pd_launch_qs_bc(hipStream_t s, const Chunk* chunks, int n, EnvDev* envs) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_qs_bc, dim3(grid_for(n)), dim3(256), 0, s, chunks, n, envs);
}

void pd_launch_qs_update(hipStream_t s, const Chunk* chunks, int n, EnvDev* envs, double4* opt, double dt,
                         double lambda, double a, unsigned long long* maxes) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_qs_update, dim3(grid_for(n)), dim3(256), 0, s, chunks, n, envs, opt, dt, dt * dt, lambda,
                     lambda / 10.0, a, maxes);
}

void pd_launch_aux_init(hipStream_t s, const double* m, double4* aux, int n) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_aux_init, dim3((n + 255) / 256), dim3(256), 0, s, m, aux, n);
}
