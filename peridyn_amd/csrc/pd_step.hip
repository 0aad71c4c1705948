// Per-step kernels of the explicit-dynamics hot path (gfx950, wave64).
//
// One wavefront owns one SELL-64 chunk: lane l handles particle first + l and
// walks its own bond row k = 0..deg-1 (slot = slot0 + 64 k + l), so the column
// and state loads of a wavefront are contiguous and each lane sums its own force
// (a gather: no atomics, deterministic).  Neighbour positions are gathered from
// the particle arrays (L2/MALL resident for lattice-ordered specimens).
//
// Forms (template PREC):
//   PD_PREC_EXACT  reproduces the single-threaded reference's accumulation order
//                  for every particle p (general_material.jl:330-368): reverse
//                  terms from neighbours j < p in ascending j, then p's own row in
//                  descending j (the family order), then reverse terms from j > p
//                  in ascending j; each term rounded as ((xx*Y)*vol)/rho.  theta,
//                  td_norm2 and m are summed in the family's descending order.
//   PD_PREC_FP64   one FP64 pass per bond in ascending neighbour order, both
//                  halves of the pair force combined.
#include "pd_device.h"
#include "pd_kernels.h"

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

// NonLinearSpring force (nonlinearspring.jl:133-143) on dr.
__device__ __forceinline__ void contact_fn(const ContactDev& C, double d1, double d2, double d3,
                                           double& F1, double& F2, double& F3) {
  const double mag = pd_mag(d1, d2, d3);
  const double del = C.equi - mag;
  const double strain = del / C.equi;
  if (del < 0.0) {
    F1 = 0.0; F2 = 0.0; F3 = 0.0;
    return;
  }
  const double pw = C.expo_is_int ? pd_pow_int(strain, (int64_t)C.expo) : pow(strain, C.expo);
  const double s = (C.K * pw) / (mag + 1e-10);
  F1 = s * d1; F2 = s * d2; F3 = s * d3;
}

struct WaveCtx {
  Chunk ch;
  int lane, p;
};

__device__ __forceinline__ bool wave_chunk(const Chunk* __restrict__ chunks, int nchunks, WaveCtx& w) {
  const int lb = pd_xcd_block(gridDim.x);
  const int cid = lb * PD_WAVES_PER_BLOCK + pd_wave_id();
  if (cid >= nchunks) return false;
  w.ch = chunks[cid];
  w.lane = pd_lane();
  w.p = w.ch.first + w.lane;
  return true;
}

}  // namespace

// ---------------------------------------------------------------------------
// Dilatation theta (src/peri.jl:155-172) for OSB and EP blocks.
// ---------------------------------------------------------------------------
template <int PREC>
__global__ __launch_bounds__(256) void k_theta(const Chunk* __restrict__ chunks, int nchunks,
                                               EnvDev* __restrict__ envs) {
  WaveCtx w;
  if (!wave_chunk(chunks, nchunks, w)) return;
  const EnvDev& E = envs[w.ch.env];
  const BlockDev& B = E.blocks[w.ch.block];
  if (B.kind == PD_BB || !B.deformed) return;
  if (w.lane >= w.ch.count) return;
  const int p = w.p;
  const double4* __restrict__ x = E.x;
  const double4* __restrict__ y = E.y[E.cur];
  const int32_t* __restrict__ col = E.col;
  const uint8_t* __restrict__ st = E.state;
  const double4 xp = x[p], yp = y[p];
  const int d = E.deg[p];
  const uint32_t s0 = w.ch.slot0 + w.lane;
  const double c_p = 3.0 / E.m[p];
  double th = 0.0;
  for (int kk = 0; kk < d; ++kk) {
    const int k = (PREC == PD_PREC_EXACT) ? (d - 1 - kk) : kk;
    const uint32_t slot = s0 + 64u * (uint32_t)k;
    if (st[slot]) {
      const int j = col[slot];
      Geom g;
      geom(xp, yp, x[j], y[j], g);
      const double wgt = 1.0 / g.xij;
      th += ((((c_p * wgt) * g.xij) * g.ext) * x[j].w) * 1.0;
    }
  }
  E.theta[p] = th;
}

// ---------------------------------------------------------------------------
// td_norm2 (elasto_plastic.jl:168-172,282-295 through loop_over_neighs!,
// src/peri.jl:411-445); breaks bonds with s >= s0.
// ---------------------------------------------------------------------------
template <int PREC>
__global__ __launch_bounds__(256) void k_td(const Chunk* __restrict__ chunks, int nchunks,
                                            EnvDev* __restrict__ envs) {
  WaveCtx w;
  if (!wave_chunk(chunks, nchunks, w)) return;
  const EnvDev& E = envs[w.ch.env];
  const BlockDev& B = E.blocks[w.ch.block];
  if (B.kind != PD_EP || !B.deformed) return;
  if (w.lane >= w.ch.count) return;
  const int p = w.p;
  const double4* __restrict__ x = E.x;
  const double4* __restrict__ y = E.y[E.cur];
  const int32_t* __restrict__ col = E.col;
  uint8_t* __restrict__ st = E.state;
  const double* __restrict__ edp = E.edp[E.edp_cur];
  const double4 xp = x[p], yp = y[p];
  const int d = E.deg[p];
  const int nt = B.ntype;
  const int tp = E.ptype[p];
  const double m_p = E.m[p], th_p = E.theta[p];
  const uint32_t s0 = w.ch.slot0 + w.lane;
  double acc = 0.0;
  for (int kk = 0; kk < d; ++kk) {
    const int k = (PREC == PD_PREC_EXACT) ? (d - 1 - kk) : kk;
    const uint32_t slot = s0 + 64u * (uint32_t)k;
    if (!st[slot]) continue;
    const int j = col[slot];
    const double4 xj = x[j];
    Geom g;
    geom(xp, yp, xj, y[j], g);
    const int tj = (nt > 1) ? E.ptype[j] : 0;
    if (g.s < B.s0[tp + nt * tj]) {
      const double wij = 1.0 / g.xij;
      const double wji = 1.0 / pd_mag(-g.X1, -g.X2, -g.X3);
      const double Gij = B.G[tp + nt * tj];
      const double a = wij / m_p, b = wji / E.m[j];
      const double w_m = a + b;
      const double tw_m = th_p * a + E.theta[j] * b;
      const double theta_ij = tw_m / w_m;
      const double ed = g.ext - (theta_ij * g.xij) / 3.0;
      const double td = ((15.0 * Gij) * (ed - edp[slot])) * w_m;
      acc += (td * td) * xj.w;
    } else {
      st[slot] = 0;
    }
  }
  E.td[p] = acc;
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
template <int KIND, int PREC>
__global__ __launch_bounds__(256) void k_force(const Chunk* __restrict__ chunks, int nchunks,
                                               EnvDev* __restrict__ envs, int flags) {
  WaveCtx w;
  if (!wave_chunk(chunks, nchunks, w)) return;
  EnvDev& E = envs[w.ch.env];
  const BlockDev& B = E.blocks[w.ch.block];
  if (w.lane >= w.ch.count) return;
  const int p = w.p;
  const double4* __restrict__ x = E.x;
  const double4* __restrict__ y = E.y[E.cur];
  const double4 xp = x[p], yp = y[p];
  const int nt = B.ntype;
  const int tp = E.ptype[p];
  const double den_p = B.density[tp];
  double a1 = 0.0, a2 = 0.0, a3 = 0.0;

  if (B.deformed) {
    const int32_t* __restrict__ col = E.col;
    uint8_t* __restrict__ st = E.state;
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

    // t for owner o (= p when own) and neighbour q, stretch already known
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
      const double t = ep_t(C, B.K[ij], B.G[ij], g.xij, g.ext, wij / m_p, wji / E.m[j], th_p,
                            E.theta[j], edp_in[slot], &en);
      edp_out[slot] = en;
      return t;
    };
    // t_jp seen from neighbour j as owner (EP only differs from t_own)
    auto t_rev = [&](int j, int tj, uint32_t slot, const Geom& g) -> double {
      if (KIND != PD_EP) {
        if (KIND == PD_BB) return g.s * c11;
        const int ji = tj + nt * tp;
        const double wij = 1.0 / g.xij;
        const double wji = 1.0 / pd_mag(-g.X1, -g.X2, -g.X3);
        const double m_j = E.m[j];
        const double a = wij / m_j, b = wji / m_p;  // owner j
        return (3.0 * B.K[ji] - 5.0 * B.G[ji]) * (E.theta[j] * g.xij * a + th_p * g.xij * b) +
               15.0 * B.G[ji] * g.ext * (a + b);
      }
      const int ji = tj + nt * tp;
      const double C = __builtin_sqrt((2.0 * B.psi[ji]) / E.td[j]);
      const double wij = 1.0 / g.xij;
      const double wji = 1.0 / pd_mag(-g.X1, -g.X2, -g.X3);
      double en;
      return ep_t(C, B.K[ji], B.G[ji], g.xij, g.ext, wij / E.m[j], wji / m_p, E.theta[j], th_p,
                  edp_in[E.rev[slot]], &en);
    };

    if (PREC == PD_PREC_EXACT) {
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
        if (!st[slot]) {
          if (KIND == PD_EP) edp_out[slot] = edp_in[slot];
          continue;
        }
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
    } else {
      const double inv_den = 1.0 / den_p;
      for (int k = 0; k < d; ++k) {
        const uint32_t slot = s0 + 64u * (uint32_t)k;
        if (!st[slot]) {
          if (KIND == PD_EP) edp_out[slot] = edp_in[slot];
          continue;
        }
        const int j = col[slot];
        const int tj = (nt > 1) ? E.ptype[j] : 0;
        Geom g;
        const double4 xj = x[j];
        geom(xp, yp, xj, y[j], g);
        if (g.s < B.s0[tp + nt * tj]) {
          double tsum;
          if (KIND == PD_EP) {
            tsum = 0.5 * (t_own(j, tj, slot, g) + t_rev(j, tj, slot, g));
          } else {
            tsum = t_own(j, tj, slot, g);
          }
          const double xx = (tsum / g.yij) * (xj.w * inv_den);
          a1 += xx * g.Y1;
          a2 += xx * g.Y2;
          a3 += xx * g.Y3;
        } else {
          st[slot] = 0;
          if (KIND == PD_EP) edp_out[slot] = edp_in[slot];
        }
      }
    }
  }

  // contact forces, models in order (update_contact_acc!, solvers.jl:257-269)
  if (E.ncontacts > 0) {
    for (int cm = 0; cm < E.ncontacts; ++cm) {
      const ContactDev& C = E.contacts[cm];
      const int i1 = C.inv1[p];
      if (i1 >= 0) {
        const double inv_den_p = E.inv_den[p];
        const int cnt = C.own_cnt[i1];
        for (int k = 0; k < cnt; ++k) {
          const int q = C.idx2[C.own[(size_t)i1 * C.max_neighs + k]];
          const double4 yq = y[q];
          double F1, F2, F3;
          contact_fn(C, yp.x - yq.x, yp.y - yq.y, yp.z - yq.z, F1, F2, F3);
          const double a = x[q].w * inv_den_p;
          a1 += F1 * a; a2 += F2 * a; a3 += F3 * a;
        }
      }
      const int i2 = C.inv2[p];
      if (i2 >= 0) {
        const double inv_den_p = E.inv_den[p];
        const int off = C.rev_off[i2], cnt = C.rev_cnt[i2];
        for (int k = 0; k < cnt; ++k) {
          const int q = C.idx1[C.rev[off + k]];
          const double4 yq = y[q];
          double F1, F2, F3;
          contact_fn(C, yq.x - yp.x, yq.y - yp.y, yq.z - yp.z, F1, F2, F3);
          const double b = x[q].w * inv_den_p;
          a1 -= F1 * b; a2 -= F2 * b; a3 -= F3 * b;
        }
      }
    }
  }

  E.f[p] = make_double4(a1, a2, a3, 0.0);
  if (flags & (PD_F_KICK | PD_F_DRIFT)) {
    const double dt = E.dt, c = dt / 2.0;
    double4 v = E.v[p];
    const int bci = E.bcid[p];
    if (flags & PD_F_KICK) {
      v.x = v.x + c * a1; v.y = v.y + c * a2; v.z = v.z + c * a3;
      if (bci >= 0) {
        const BcDev& bc = E.bcs[bci];
        if (bc.kind == PD_BC_FIX) {
          v.x = 0.0; v.y = 0.0; v.z = 0.0;
        } else {
          const double dir = pd_bc_dir(bc, E.time_step);
          v.x = dir * bc.rate[0]; v.y = dir * bc.rate[1]; v.z = dir * bc.rate[2];
        }
      }
    }
    if (flags & PD_F_DRIFT) {
      v.x = v.x + c * a1; v.y = v.y + c * a2; v.z = v.z + c * a3;
      double4 yn = make_double4(yp.x + dt * v.x, yp.y + dt * v.y, yp.z + dt * v.z, 0.0);
      if (bci >= 0) {
        const BcDev& bc = E.bcs[bci];
        if (bc.kind == PD_BC_FIX) {
          yn = yp;
        } else {
          const double ddir = dt * pd_bc_dir(bc, E.time_step + 1);
          yn.x = yp.x + ddir * bc.rate[0]; yn.y = yp.y + ddir * bc.rate[1]; yn.z = yp.z + ddir * bc.rate[2];
        }
      }
      E.y[E.cur ^ 1][p] = yn;
    }
    E.v[p] = v;
  }
}

// Start of a run call: kick + drift + position BCs of step time_step + 1 from
// the stored force (velocity_verlet.jl:63-69).
__global__ __launch_bounds__(256) void k_prologue(const Chunk* __restrict__ chunks, int nchunks,
                                                  EnvDev* __restrict__ envs) {
  WaveCtx w;
  if (!wave_chunk(chunks, nchunks, w)) return;
  EnvDev& E = envs[w.ch.env];
  if (w.lane >= w.ch.count) return;
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
      const double ddir = dt * pd_bc_dir(bc, E.time_step + 1);
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
// with rtol = sqrt(eps(Float64)); one workgroup per (env, block) entry.
__global__ __launch_bounds__(1024) void k_latch(EnvDev* __restrict__ envs, const int2* __restrict__ items) {
  const int2 it = items[blockIdx.x];
  EnvDev& E = envs[it.x];
  BlockDev& B = E.blocks[it.y];
  if (B.deformed) return;
  const double4* __restrict__ y = E.y[E.cur];
  const double4* __restrict__ x = E.x;
  const double4 y1 = y[B.first], x1 = x[B.first];
  double sd = 0.0, sa = 0.0, sx = 0.0;
  for (int i = threadIdx.x; i < B.n; i += blockDim.x) {
    const double4 yi = y[B.first + i], xi = x[B.first + i];
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
  __shared__ double red[3][1024 / 64];
  for (int o = 32; o > 0; o >>= 1) {
    sd += __shfl_down(sd, o);
    sa += __shfl_down(sa, o);
    sx += __shfl_down(sx, o);
  }
  const int wv = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) { red[0][wv] = sd; red[1][wv] = sa; red[2][wv] = sx; }
  __syncthreads();
  if (threadIdx.x == 0) {
    double D = 0.0, A = 0.0, X = 0.0;
    for (int q = 0; q < (int)(blockDim.x >> 6); ++q) { D += red[0][q]; A += red[1][q]; X += red[2][q]; }
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
}

// ---- launch wrappers --------------------------------------------------------
static inline int grid_for(int nchunks) {
  return (nchunks + PD_WAVES_PER_BLOCK - 1) / PD_WAVES_PER_BLOCK;
}

void pd_launch_theta(hipStream_t s, int prec, const Chunk* chunks, int n, EnvDev* envs) {
  if (n <= 0) return;
  if (prec == PD_PREC_EXACT)
    hipLaunchKernelGGL(k_theta<PD_PREC_EXACT>, dim3(grid_for(n)), dim3(256), 0, s, chunks, n, envs);
  else
    hipLaunchKernelGGL(k_theta<PD_PREC_FP64>, dim3(grid_for(n)), dim3(256), 0, s, chunks, n, envs);
}

void pd_launch_td(hipStream_t s, int prec, const Chunk* chunks, int n, EnvDev* envs) {
  if (n <= 0) return;
  if (prec == PD_PREC_EXACT)
    hipLaunchKernelGGL(k_td<PD_PREC_EXACT>, dim3(grid_for(n)), dim3(256), 0, s, chunks, n, envs);
  else
    hipLaunchKernelGGL(k_td<PD_PREC_FP64>, dim3(grid_for(n)), dim3(256), 0, s, chunks, n, envs);
}

template <int KIND>
static void launch_force_kind(hipStream_t s, int prec, const Chunk* chunks, int n, EnvDev* envs, int flags) {
  if (prec == PD_PREC_EXACT)
    hipLaunchKernelGGL((k_force<KIND, PD_PREC_EXACT>), dim3(grid_for(n)), dim3(256), 0, s, chunks, n, envs, flags);
  else
    hipLaunchKernelGGL((k_force<KIND, PD_PREC_FP64>), dim3(grid_for(n)), dim3(256), 0, s, chunks, n, envs, flags);
}

void pd_launch_force(hipStream_t s, int kind, int prec, const Chunk* chunks, int n, EnvDev* envs, int flags) {
  if (n <= 0) return;
  if (kind == PD_BB) launch_force_kind<PD_BB>(s, prec, chunks, n, envs, flags);
  else if (kind == PD_OSB) launch_force_kind<PD_OSB>(s, prec, chunks, n, envs, flags);
  else launch_force_kind<PD_EP>(s, prec, chunks, n, envs, flags);
}

void pd_launch_prologue(hipStream_t s, const Chunk* chunks, int n, EnvDev* envs) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_prologue, dim3(grid_for(n)), dim3(256), 0, s, chunks, n, envs);
}

void pd_launch_book(hipStream_t s, EnvDev* envs, const int32_t* slots, int n, int flags) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_book, dim3((n + 255) / 256), dim3(256), 0, s, envs, slots, n, flags);
}

void pd_launch_latch(hipStream_t s, EnvDev* envs, const int2* items, int n) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_latch, dim3(n), dim3(1024), 0, s, envs, items);
}
