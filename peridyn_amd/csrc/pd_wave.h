// Wavefront/chunk helpers and the per-particle step epilogue shared by the
// FP64 and FP32-compensated kernels.
#pragma once
#include "pd_device.h"
#include "pd_kernels.h"

struct WaveCtx {
  Chunk ch;
  int cid;  // index of the chunk in the launch's chunk list
  int lane, lgG, G, C, q, g, p;
  bool active;
};

// Chunk of the calling wavefront (XCD-aware block order); false when the
// wavefront has no chunk.  Lane L = g*C + q.
__device__ __forceinline__ bool wave_chunk(const Chunk* __restrict__ chunks, int nchunks,
                                           const EnvDev* __restrict__ envs, WaveCtx& w) {
  const int lb = pd_xcd_block(gridDim.x);
  const int cid = lb * PD_WAVES_PER_BLOCK + pd_wave_id();
  if (cid >= nchunks) return false;
  w.cid = cid;
  w.ch = chunks[cid];
  w.lgG = envs[w.ch.env].lgG;
  w.G = 1 << w.lgG;
  w.C = 64 >> w.lgG;
  w.lane = pd_lane();
  w.q = w.lane & (w.C - 1);
  w.g = w.lane >> (6 - w.lgG);
  w.p = w.ch.first + w.q;
  w.active = w.q < w.ch.count;
  return true;
}

// Sum over the G lanes of a particle (lanes q, q+C, q+2C, ...); every lane
// receives the same total (xor butterfly: a+b == b+a at every level).
__device__ __forceinline__ double group_sum(double v, int C) {
  for (int off = C; off < 64; off <<= 1) v += __shfl_xor(v, off);
  return v;
}

// NonLinearSpring force (nonlinearspring.jl:133-143) on dr.  GPOW admits
// non-integer exponents (Julia's Float64 ^ through the device pow); without it
// every exponent is integer (the host launches the GPOW kernels otherwise) and
// the kernel carries no pow code, whose registers would cap its occupancy.
template <bool GPOW>
__device__ __forceinline__ void contact_fn(const ContactDev& C, double d1, double d2, double d3,
                                           double& F1, double& F2, double& F3) {
  const double mag = pd_mag(d1, d2, d3);
  const double del = C.equi - mag;
  const double strain = del / C.equi;
  if (del < 0.0) {
    F1 = 0.0; F2 = 0.0; F3 = 0.0;
    return;
  }
  double pw;
  if (GPOW && !C.expo_is_int) pw = pow(strain, C.expo);
  else pw = pd_pow_int(strain, (int64_t)C.expo);
  const double s = (C.K * pw) / (mag + 1e-10);
  F1 = s * d1; F2 = s * d2; F3 = s * d3;
}

// Self-contact terms of particle p (short_range_contact! for ContactModel11,
// contacts.jl:357-380), in the order of the reference's loop over members i:
// the terms p receives as neighbour of members i < p, its own term (i = p),
// then those of members i > p.  Own: f_p += F(y_p - y_j) (vol_j / den_p);
// as neighbour of i: f_p -= F(y_i - y_p) (vol_p / den_p) -- the reference uses
// the neighbour's volume on both sides.  den = mass / volume (environment.jl:214).
template <bool GPOW>
__device__ __forceinline__ void add_self_contact(const EnvDev& E, const ContactDev& C, int p, const double4& yp,
                                                 const double4* y, double& a1, double& a2, double& a3) {
  const int i = C.inv1[p];
  if (i < 0) return;
  const double4* __restrict__ x = E.x;
  const double vol_p = x[p].w;
  const double den_p = (vol_p * E.blocks[E.pblock[p]].density[E.ptype[p]]) / vol_p;
  const int off = C.rev_off[i], cnt = C.rev_cnt[i];
  const double b = vol_p / den_p;
  int k = 0;
  auto rev_term = [&](int kk) {
    const int q = C.idx1[C.rev[off + kk]];
    const double4 yq = y[q];
    double F1, F2, F3;
    contact_fn<GPOW>(C, yq.x - yp.x, yq.y - yp.y, yq.z - yp.z, F1, F2, F3);
    a1 -= F1 * b; a2 -= F2 * b; a3 -= F3 * b;
  };
  for (; k < cnt && C.rev[off + k] < i; ++k) rev_term(k);
  if (C.own_cnt[i]) {
    const int q = C.idx1[C.own[(size_t)i * C.max_neighs]];
    const double4 yq = y[q];
    double F1, F2, F3;
    contact_fn<GPOW>(C, yp.x - yq.x, yp.y - yq.y, yp.z - yq.z, F1, F2, F3);
    const double a = x[q].w / den_p;
    a1 += F1 * a; a2 += F2 * a; a3 += F3 * a;
  }
  for (; k < cnt; ++k) rev_term(k);
}

// Terms of one contact model for particle p (short_range_contact! RM12,
// contacts.jl:230-245; RM11, contacts.jl:357-380), added to (a1, a2, a3) in the
// order the reference's loop over block-1 members adds them to f1 / f2.
template <bool GPOW>
__device__ __forceinline__ void add_contact_model(const EnvDev& E, const ContactDev& C, int p, const double4& yp,
                                                  const double4* y, double& a1, double& a2, double& a3) {
  if (C.self_contact) {
    add_self_contact<GPOW>(E, C, p, yp, y, a1, a2, a3);
    return;
  }
  const double4* __restrict__ x = E.x;
  const int i1 = C.inv1[p];
  if (i1 >= 0) {
    const double inv_den_p = E.inv_den[p];
    const int cnt = C.own_cnt[i1];
    for (int k = 0; k < cnt; ++k) {
      const int q = C.idx2[C.own[(size_t)i1 * C.max_neighs + k]];
      const double4 yq = y[q];
      double F1, F2, F3;
      contact_fn<GPOW>(C, yp.x - yq.x, yp.y - yq.y, yp.z - yq.z, F1, F2, F3);
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
      contact_fn<GPOW>(C, yq.x - yp.x, yq.y - yp.y, yq.z - yp.z, F1, F2, F3);
      const double b = x[q].w * inv_den_p;
      a1 -= F1 * b; a2 -= F2 * b; a3 -= F3 * b;
    }
  }
}

// Contact terms of particle p, models in order (update_contact_acc!,
// solvers.jl:257-269), added to the material force as the reference's f1/f2
// updates are.
template <bool GPOW>
__device__ __forceinline__ void add_contacts(const EnvDev& E, int p, const double4& yp, const double4* y,
                                             double& a1, double& a2, double& a3) {
  for (int cm = 0; cm < E.ncontacts; ++cm) add_contact_model<GPOW>(E, E.contacts[cm], p, yp, y, a1, a2, a3);
}

// True when particle p receives contact terms under the current lists (any
// model): the terms add_contacts would add are then not all absent.
__device__ __forceinline__ bool pd_has_contact_terms(const EnvDev& E, int p) {
  for (int cm = 0; cm < E.ncontacts; ++cm) {
    const ContactDev& C = E.contacts[cm];
    const int i1 = C.inv1[p];
    if (C.self_contact) {
      if (i1 >= 0 && (C.own_cnt[i1] > 0 || C.rev_cnt[i1] > 0)) return true;
      continue;
    }
    if (i1 >= 0 && C.own_cnt[i1] > 0) return true;
    const int i2 = C.inv2[p];
    if (i2 >= 0 && C.rev_cnt[i2] > 0) return true;
  }
  return false;
}

// Stores f and runs the FP64 velocity-Verlet part of the fused step:
//   KICK : v += dt/2 f; velocity BCs                  (velocity_verlet.jl:74-79)
//   DRIFT: v += dt/2 f; y' = y + dt v; position BCs of the next step
//          (velocity_verlet.jl:63-69) into the other y buffer.
__device__ __forceinline__ void finish_particle_v(EnvDev& E, int p, const double4& xp, const double4& yp,
                                                  double a1, double a2, double a3, int flags, double4 v,
                                                  const int bci) {
  E.f[p] = make_double4(a1, a2, a3, 0.0);
  if (!(flags & (PD_F_KICK | PD_F_DRIFT))) return;
  const double dt = E.dt, c = dt / 2.0;
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

__device__ __forceinline__ void finish_particle(EnvDev& E, int p, const double4& xp, const double4& yp,
                                                double a1, double a2, double a3, int flags) {
  finish_particle_v(E, p, xp, yp, a1, a2, a3, flags, E.v[p], E.bcid[p]);
}
