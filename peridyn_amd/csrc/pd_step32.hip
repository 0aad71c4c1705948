// FP32-compensated form of the bond passes (PD_PREC_FP32C) for bond-based,
// ordinary state-based and elastoplastic blocks.
//
// Per bond:
//   X  = x32_j - x32_i      reference bond vector in FP32; x32 holds positions
//                           relative to the block's minimum corner, so lattice
//                           offsets (and any FP32-representable ones) are exact
//   U  = float((y_j - y_i) - X)   displacement difference formed in FP64 from the
//                           resident FP64 positions, then rounded: its relative
//                           error is that of FP32 on U itself, whatever the rigid
//                           motion of the body
//   |Y|^2 = |X|^2 + (2 X.U + U.U),  e = (2 X.U + U.U) / (|Y| + |X|),  s = e / |X|
//                           (no |Y| - |X| cancellation), all FP32
// Force and theta sums are Kahan-compensated FP32 per lane over groups of 8
// bond steps (GroupKahan), td_norm2 sums per bond; the G lanes of a particle
// are combined in FP64.  Positions, velocities, the Verlet
// update and the weighted volume stay FP64.  Elastoplastic blocks
// (elasto_plastic.jl:228-255,282-295) keep their plastic state in FP64: the
// plastic extensions edp (accumulated over the run from small increments that
// FP32 would absorb), the deviatoric extension ed = e - theta_ij x / 3 formed
// from the FP32 extension, the return map ede = ed - edp with its update
// edp += ede (1 - C), and the particle's own C; the neighbour's theta, 1/m,
// td_norm2 (its sign the yield flag) and C come from its FP32 record aux32.
#include "pd_fp32c.h"


// Dilatation theta (src/peri.jl:155-172) of OSB and EP blocks.
template <int DUMMY>
__global__ __launch_bounds__(256) void k_theta32(const Chunk* __restrict__ chunks, int nchunks,
                                                 EnvDev* __restrict__ envs) {
  WaveCtx w;
  if (!wave_chunk(chunks, nchunks, envs, w)) return;
  const EnvDev& E = envs[w.ch.env];
  const BlockDev& B = E.blocks[w.ch.block];
  if (B.kind == PD_BB || !B.deformed) return;
  const float4* __restrict__ x = E.x32;
  const double4* __restrict__ y = E.y[E.cur];
  const int32_t* __restrict__ col = E.col;
  const uint8_t* __restrict__ st = E.state;
  const int p = w.active ? w.p : w.ch.first;
  const float4 xp = x[p];
  const double3 wp = disp32(B, xp, y[p]);
  const int d = w.active ? E.deg[p] : 0;
  const uint32_t base = w.ch.slot0 + w.lane;
  GroupKahan acc;  // groups of kGroupSteps bond steps, as the tiled kernel's
  constexpr int U = 2;  // bonds per batch; column/state loaded one batch ahead
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
    float4 xj[U];
    double4 yj[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      xj[u] = gld4(x, jv[u]);
      yj[u] = gld4(y, jv[u]);
    }
    if (i0 + U < ni) load_stream(i0 + U);
#pragma unroll
    for (int u = 0; u < U; ++u) {
      if (sv[u]) {
        G32 g;
        geom32(xp, wp, xj[u], disp32(B, xj[u], yj[u]), g);
        acc.add(g.ext, xj[u].w);
      }
      if (i0 + u < ni && (((i0 + u) % kGroupSteps) == kGroupSteps - 1 || i0 + u == ni - 1)) acc.flush();
    }
  }
  // the tiled kernel runs every packed word of the chunk, empty ones too
  for (int q = group_count(ni); q < group_count(w.ch.width); ++q) acc.flush();
  const double tot = comp_total(acc.s, acc.c, w.C);
  if (w.active && w.g == 0) {
    const float th = (float)((3.0 * (double)E.aux32[p].y) * tot);
    E.aux32[p].x = th;
    E.theta[p] = (double)th;
  }
}

// td_norm2 of EP blocks (elasto_plastic.jl:168-172,282-295 through
// loop_over_neighs!, src/peri.jl:411-445); breaks bonds with s >= s0 (a bond
// broken here keeps its plastic extension in both edp buffers).  Writes the
// particle's td_norm2 (sign: the sticky yield flag) and C = sqrt(2 psi /
// td_norm2) to its FP64 record aux and its FP32 record aux32.
template <int DUMMY>
__global__ __launch_bounds__(256) void k_td32(const Chunk* __restrict__ chunks, int nchunks,
                                              EnvDev* __restrict__ envs) {
  WaveCtx w;
  if (!wave_chunk(chunks, nchunks, envs, w)) return;
  EnvDev& E = envs[w.ch.env];
  const BlockDev& B = E.blocks[w.ch.block];
  if (B.kind != PD_EP || !B.deformed) return;
  const float4* __restrict__ x = E.x32;
  const double4* __restrict__ y = E.y[E.cur];
  const int32_t* __restrict__ col = E.col;
  uint8_t* __restrict__ st = E.state;
  const double* __restrict__ edp = E.edp[E.edp_cur];
  double* __restrict__ edp_o = E.edp[E.edp_cur ^ 1];
  const int p = w.active ? w.p : w.ch.first;
  const float4 xp = x[p];
  const double3 wp = disp32(B, xp, y[p]);
  const float4 ap = E.aux32[p];  // (theta, 1/m, +-td_norm2, C)
  const bool yielded = pd_has_yielded(E.aux[p]);  // p's plastic extensions are 0 until it yields
  const int d = w.active ? E.deg[p] : 0;
  const int nt = B.ntype;
  const int tp = E.ptype[p];
  const uint32_t base = w.ch.slot0 + w.lane;
  float s = 0.0f, c = 0.0f;
  constexpr int U = 2;  // bonds per batch; column/state/edp loaded one batch ahead
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
    float4 xj[U], aj[U];
    double4 yj[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      xj[u] = gld4(x, jv[u]);
      yj[u] = gld4(y, jv[u]);
      aj[u] = gld4(E.aux32, jv[u]);
    }
    if (i0 + U < ni) load_stream(i0 + U);
#pragma unroll
    for (int u = 0; u < U; ++u) {
      if (!sv[u]) continue;
      const uint32_t slot = base + 64u * (uint32_t)(i0 + u);
      G32 g;
      geom32(xp, wp, xj[u], disp32(B, xj[u], yj[u]), g);
      const int tj = (nt > 1) ? E.ptype[jv[u]] : 0;
      const int ij = tp + nt * tj;
      if ((double)g.s < B.s0[ij]) {
        const double sm = (double)ap.y + (double)aj[u].y;
        const double twm = (double)ap.x * (double)ap.y + (double)aj[u].x * (double)aj[u].y;
        const double ed = (double)g.ext - (twm * pd_rcp(sm)) * (double)g.xij * (1.0 / 3.0);
        const double td = (15.0 * B.G[ij]) * (ed - eo[u]) * ((double)g.inv_x * sm);
        kahan(s, c, (float)(td * td) * xj[u].w);
      } else {
        gst(st, slot, (uint8_t)0);
        if (yielded) gst(edp_o, slot, eo[u]);  // else 0 in both buffers already
      }
    }
  }
  const double acc = comp_total(s, c, w.C);
  if (w.active && w.g == 0) {
    const double C = __builtin_sqrt((2.0 * B.psi[tp + nt * tp]) / acc);
    // multi-type blocks yield per type pair: flagged from the start
    const bool flag = yielded || !(C > 1.0) || nt > 1;
    const double tz = flag ? -acc : acc;
    E.td[p] = acc;
    E.aux[p].z = tz;
    E.aux[p].w = C;
    E.aux32[p].z = flag ? -(float)acc : (float)acc;
    E.aux32[p].w = (float)C;
  }
}

// Fused force density (BB / OSB / EP) + FP64 velocity-Verlet (finish_particle).
template <int KIND, bool GPOW>
__global__ __launch_bounds__(256) void k_force32(const Chunk* __restrict__ chunks, int nchunks,
                                                 EnvDev* __restrict__ envs, int flags) {
  WaveCtx w;
  if (!wave_chunk(chunks, nchunks, envs, w)) return;
  EnvDev& E = envs[w.ch.env];
  const BlockDev& B = E.blocks[w.ch.block];
  const int p = w.active ? w.p : w.ch.first;
  const double4* __restrict__ y = E.y[E.cur];
  const double4 yp = y[p];
  double a1 = 0.0, a2 = 0.0, a3 = 0.0;
  if (B.deformed) {
    const float4* __restrict__ x = E.x32;
    const int32_t* __restrict__ col = E.col;
    uint8_t* __restrict__ st = E.state;
    const float4 xp = x[p];
    const double3 wp = disp32(B, xp, yp);
    const int d = w.active ? E.deg[p] : 0;
    const int nt = B.ntype;
    const int tp = E.ptype[p];
    const float inv_den = (float)(1.0 / B.density[tp]);
    const uint32_t base = w.ch.slot0 + w.lane;
    float4 ap = make_float4(0.f, 0.f, 0.f, 0.f);
    if (KIND != PD_BB) ap = E.aux32[p];
    double4 ap64 = make_double4(0.0, 0.0, 0.0, 0.0);  // EP: the yield flag and C in FP64
    const double* __restrict__ edp_in = nullptr;
    double* __restrict__ edp_out = nullptr;
    bool yielded_p = false;
    if (KIND == PD_EP) {
      ap64 = E.aux[p];
      yielded_p = pd_has_yielded(ap64);
      edp_in = E.edp[E.edp_cur];
      edp_out = E.edp[E.edp_cur ^ 1];
    }
    const float c11 = (float)B.c11;
    GroupKahan k1, k2, k3;  // groups of kGroupSteps bond steps, as the tiled kernel's
    constexpr int U = 2;  // bonds per batch; column/state loaded one batch ahead
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
      float4 xj[U], aj[U];
      double4 yj[U];
      double eo[U], er[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        xj[u] = gld4(x, jv[u]);
        yj[u] = gld4(y, jv[u]);
        if (KIND != PD_BB) aj[u] = gld4(E.aux32, jv[u]);
        if (KIND == PD_EP) eo[u] = (sv[u] && yielded_p) ? gld(edp_in, base + 64u * (uint32_t)(i0 + u)) : 0.0;
      }
      if (KIND == PD_EP) {
        // the reverse bond's plastic extension, 0 unless j has yielded
#pragma unroll
        for (int u = 0; u < U; ++u) {
          er[u] = 0.0;
          if (sv[u] && pd_has_yielded32(aj[u])) er[u] = gld(edp_in, gld(E.rev, base + 64u * (uint32_t)(i0 + u)));
        }
      }
      if (i0 + U < ni) load_stream(i0 + U);
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const bool grp_end = i0 + u < ni && (((i0 + u) % kGroupSteps) == kGroupSteps - 1 || i0 + u == ni - 1);
        if (!sv[u]) {
          if (grp_end) {
            k1.flush();
            k2.flush();
            k3.flush();
          }
          continue;
        }
        const uint32_t slot = base + 64u * (uint32_t)(i0 + u);
        G32 g;
        geom32(xp, wp, xj[u], disp32(B, xj[u], yj[u]), g);
        const int tj = (nt > 1) ? E.ptype[jv[u]] : 0;
        const int ij = tp + nt * tj;
        float xx = 0.0f;  // (t / |Y|) vol_j / density; 0 for a bond broken here
        if ((double)g.s < B.s0[ij]) {
          if (KIND == PD_BB) {
            const float tt = g.s * g.inv_y;
            xx = tt * cvd32(c11, xj[u].w, inv_den);
          } else if (KIND == PD_OSB) {
            const float Kf = (float)B.K[ij], Gf = (float)B.G[ij];
            const float tt = osb_t32(ap, ap.x * ap.y, aj[u].x, aj[u].y, 3.0f * Kf - 5.0f * Gf, 15.0f * Gf, g) * g.inv_y;
            xx = tt * vd32(xj[u].w, inv_den);
          } else {
            // elasto_plastic.jl:228-255 for both halves of the pair, plastic
            // state in FP64; the own bond's updated extension goes to the
            // other edp buffer
            const double sm = (double)ap.y + (double)aj[u].y;
            const double twm = (double)ap.x * (double)ap.y + (double)aj[u].x * (double)aj[u].y;
            const double ed = (double)g.ext - (twm * pd_rcp(sm)) * (double)g.xij * (1.0 / 3.0);
            const double g15 = 15.0 * B.G[ij] * ((double)g.inv_x * sm);
            double Cp = ap64.w, Cj = (double)aj[u].w;
            if (nt > 1) {
              Cp = __builtin_sqrt((2.0 * B.psi[ij]) / __builtin_fabs(ap64.z));
              Cj = __builtin_sqrt((2.0 * B.psi[ij]) / __builtin_fabs((double)aj[u].z));
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
            const float t = (float)(3.0 * B.K[ij] * twm + 0.5 * (td_own + td_rev));
            xx = (t * g.inv_y) * vd32(xj[u].w, inv_den);
          }
        } else {
          gst(st, slot, (uint8_t)0);
          if (KIND == PD_EP && yielded_p) gst(edp_out, slot, eo[u]);
        }
        k1.add(xx, g.Y1);
        k2.add(xx, g.Y2);
        k3.add(xx, g.Y3);
        if (grp_end) {
          k1.flush();
          k2.flush();
          k3.flush();
        }
      }
    }
    // the tiled kernel runs every packed word of the chunk, empty ones too
    for (int q = group_count(ni); q < group_count(w.ch.width); ++q) {
      k1.flush();
      k2.flush();
      k3.flush();
    }
    a1 = comp_total(k1.s, k1.c, w.C);
    a2 = comp_total(k2.s, k2.c, w.C);
    a3 = comp_total(k3.s, k3.c, w.C);
  }
  if (!w.active || w.g != 0) return;
  add_contacts<GPOW>(E, p, yp, y, a1, a2, a3);
  finish_particle(E, p, E.x[p], yp, a1, a2, a3, flags);
}

// ---- launch wrappers --------------------------------------------------------
static inline int grid32(int nchunks) { return (nchunks + PD_WAVES_PER_BLOCK - 1) / PD_WAVES_PER_BLOCK; }

void pd_launch_theta32(hipStream_t s, const Chunk* chunks, int n, EnvDev* envs) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_theta32<0>, dim3(grid32(n)), dim3(256), 0, s, chunks, n, envs);
}

void pd_launch_td32(hipStream_t s, const Chunk* chunks, int n, EnvDev* envs) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_td32<0>, dim3(grid32(n)), dim3(256), 0, s, chunks, n, envs);
}

void pd_launch_force32(hipStream_t s, int kind, const Chunk* chunks, int n, EnvDev* envs, int flags, bool gpow) {
  if (n <= 0) return;
  const dim3 g(grid32(n)), b(256);
  if (kind == PD_BB) {
    if (gpow) hipLaunchKernelGGL((k_force32<PD_BB, true>), g, b, 0, s, chunks, n, envs, flags);
    else hipLaunchKernelGGL((k_force32<PD_BB, false>), g, b, 0, s, chunks, n, envs, flags);
  } else if (kind == PD_OSB) {
    if (gpow) hipLaunchKernelGGL((k_force32<PD_OSB, true>), g, b, 0, s, chunks, n, envs, flags);
    else hipLaunchKernelGGL((k_force32<PD_OSB, false>), g, b, 0, s, chunks, n, envs, flags);
  } else {
    if (gpow) hipLaunchKernelGGL((k_force32<PD_EP, true>), g, b, 0, s, chunks, n, envs, flags);
    else hipLaunchKernelGGL((k_force32<PD_EP, false>), g, b, 0, s, chunks, n, envs, flags);
  }
}
