// Tiled FP64 passes: one workgroup per tile (up to 4 chunks of one brick of a
// block's internal order).  The workgroup first stages the fields a pass needs
// of every particle in the tile's halo -- the sorted, unique neighbour ids of
// its particles -- into LDS (structure of arrays, one double array per field),
// then runs the bond loop of the untiled FP64 kernels with the neighbour
// records read from LDS: one bulk gather per halo particle instead of one
// gather per bond.  The neighbour of bond step s of a lane comes from the
// packed stream pk (one uint64 per lane and group of 4 bond steps; layout in
// pd_internal.h), so a lane reads its next 4 bonds with one load.
//
// Every bond is evaluated with the same expressions, in the same per-lane order
// and with the same G-lane combination as the untiled FP64 kernels (pd_step.hip),
// so the results are bit-identical to theirs for the same lanes per particle.
// Tiled envs hold single-type blocks only (the host checks), so type pairs
// reduce to index 0 and no neighbour type is read.
#include <type_traits>

#include "pd_fp32c.h"

namespace {

enum { TP_THETA = 0, TP_TD = 1, TP_FORCE = 2 };

// LDS halo records of the FP64 passes: one array per field.  Double fields in
// order: x0 x1 x2 vol y0 y1 y2 | theta 1/m | Cs, where Cs = C with the
// neighbour's "has yielded" flag as its sign (C >= 0).  With XF (every x and
// volume of the set's envs is exactly an FP32 value) x0 x1 x2 vol go to two
// 8-byte fields after the doubles instead, (x0, x1) and (x2, vol) as float
// pairs -- the values read back are the same doubles -- which shrinks the
// records and lets more tiles reside per CU; every field of record j then sits
// at j * 8 plus a constant.  In an integer-lattice block the first pair holds
// the lattice coordinates instead (lat_pack).
template <int PASS, int KIND>
__host__ __device__ constexpr int tile_extra() {  // double fields beyond x, vol, y
  return PASS == TP_THETA ? 0 : PASS == TP_TD ? 2 : (KIND == PD_BB ? 0 : KIND == PD_OSB ? 2 : 3);
}
// LAT (compact lattice records, the force pass of sets whose bond-based
// blocks are all integer lattices): the lattice pair and then the volume as
// one float, without x2: 12 bytes after the doubles instead of 16.  ONEV (the
// same where every such block has one volume, BlockDev::vol1): the lattice
// pair only, 8 bytes.
template <int PASS, int KIND, bool XF, bool LAT = false, bool ONEV = false>
__host__ __device__ constexpr int tile_record_bytes() {
  return LAT ? 8 * (3 + tile_extra<PASS, KIND>()) + (ONEV ? 8 : 12)
             : XF ? 8 * (3 + tile_extra<PASS, KIND>()) + 4 * 4 : 8 * (7 + tile_extra<PASS, KIND>());
}

template <bool XF, int NX, bool LAT = false>  // NX = extra double fields staged
struct Rec {
  const double* sh;
  int hcap;
  __device__ __forceinline__ Rec(const double* s, int h) : sh(s), hcap(h) {}
  __device__ __forceinline__ const float2* xa() const { return (const float2*)(sh + (3 + NX) * hcap); }
  __device__ __forceinline__ const float2* xb() const { return (const float2*)(sh + (4 + NX) * hcap); }
  __device__ __forceinline__ uint2 lat(int j) const { return ((const uint2*)xa())[j]; }
  __device__ __forceinline__ double vol(int j) const {
    return LAT ? (double)((const float*)(sh + (4 + NX) * hcap))[j] : XF ? (double)xb()[j].y : sh[3 * hcap + j];
  }
  __device__ __forceinline__ double y(int k, int j) const { return sh[((XF ? 0 : 4) + k) * hcap + j]; }
  __device__ __forceinline__ double th(int j) const { return sh[(XF ? 3 : 7) * hcap + j]; }
  __device__ __forceinline__ double rm(int j) const { return sh[(XF ? 4 : 8) * hcap + j]; }
  __device__ __forceinline__ double cs(int j) const { return sh[(XF ? 5 : 9) * hcap + j]; }
  // reference position of record j
  __device__ __forceinline__ void x(int j, double& a, double& b, double& c) const {
    if (XF) {
      const float2 p = xa()[j];
      a = p.x; b = p.y; c = xb()[j].x;
    } else {
      a = sh[j]; b = sh[hcap + j]; c = sh[2 * hcap + j];
    }
  }
};

// Lattice coordinates of an integer-lattice position, scaled by 4 and packed as
// 16-bit two's-complement pairs (x0, x1) and (x2, 0), each modulo 2^16: the
// packed difference of two such words is 4 (X0, X1) and 4 (X2, 0) exactly when
// |4 X| < 2^15, and the dot products give 16 |X|^2, the byte offset of |X|^2's
// entry in a table of 16-byte records.
__device__ __forceinline__ uint2 lat_pack(double x0, double x1, double x2) {
  const uint32_t a = (uint32_t)(4 * (int32_t)x0) & 0xffffu, b = (uint32_t)(4 * (int32_t)x1);
  return make_uint2(a | (b << 16), (uint32_t)(4 * (int32_t)x2) & 0xffffu);
}
typedef short pd_short2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ int lat_offset(uint2 a, uint2 b) {
  const pd_short2 d01 = __builtin_bit_cast(pd_short2, a.x) - __builtin_bit_cast(pd_short2, b.x);
  const pd_short2 d2 = __builtin_bit_cast(pd_short2, a.y) - __builtin_bit_cast(pd_short2, b.y);
  return __builtin_amdgcn_sdot2(d01, d01, __builtin_amdgcn_sdot2(d2, d2, 0, false), false);
}

struct TileWave {
  Chunk ch;
  int lane, q, g, p;
  bool active;
};

// Stages the halo of tile T into LDS; every thread of the workgroup takes part,
// R records per thread with all their loads in flight together (halo ids,
// then the records, then the LDS stores).
constexpr int kStageR = 4;

template <bool XF, int NX, bool LAT = false, bool ONEV = false>
__device__ __forceinline__ void stage_halo(const EnvDev& E, const Tile& T, double* __restrict__ sh, int hcap,
                                           bool lat) {
  const double4* __restrict__ x = E.x;
  const double4* __restrict__ y = E.y[E.cur];
  const int32_t* __restrict__ halo = E.halo;
  constexpr int Y0 = XF ? 0 : 4, T0 = XF ? 3 : 7;
  float2* __restrict__ xa = (float2*)(sh + (3 + NX) * hcap);
  float2* __restrict__ xb = (float2*)(sh + (4 + NX) * hcap);
  for (int i0 = threadIdx.x; i0 < T.hn; i0 += 256 * kStageR) {
    uint32_t id[kStageR];
#pragma unroll
    for (int r = 0; r < kStageR; ++r) {
      const int i = i0 + 256 * r;
      id[r] = i < T.hn ? (uint32_t)gld(halo, T.hoff + (uint32_t)i) : 0u;
    }
    double4 a[kStageR], b[kStageR], c[kStageR];
#pragma unroll
    for (int r = 0; r < kStageR; ++r) {
      if (i0 + 256 * r < T.hn) {
        a[r] = gld4(x, id[r]);
        b[r] = gld4(y, id[r]);
        if (NX > 0) c[r] = gld4(E.aux, id[r]);
      }
    }
#pragma unroll
    for (int r = 0; r < kStageR; ++r) {
      const int i = i0 + 256 * r;
      if (i < T.hn) {
        if (LAT) {
          ((uint2*)xa)[i] = lat_pack(a[r].x, a[r].y, a[r].z);
          if (!ONEV) ((float*)xb)[i] = (float)a[r].w;
        } else if (XF) {
          if (lat) ((uint2*)xa)[i] = lat_pack(a[r].x, a[r].y, a[r].z);
          else xa[i] = make_float2((float)a[r].x, (float)a[r].y);
          xb[i] = make_float2((float)a[r].z, (float)a[r].w);
        } else {
          sh[i] = a[r].x;
          sh[hcap + i] = a[r].y;
          sh[2 * hcap + i] = a[r].z;
          sh[3 * hcap + i] = a[r].w;
        }
        sh[Y0 * hcap + i] = b[r].x;
        sh[(Y0 + 1) * hcap + i] = b[r].y;
        sh[(Y0 + 2) * hcap + i] = b[r].z;
        if (NX > 0) {
          sh[T0 * hcap + i] = c[r].x;
          sh[(T0 + 1) * hcap + i] = c[r].y;
          if (NX > 2) sh[(T0 + 2) * hcap + i] = __builtin_copysign(c[r].w, c[r].z);
        }
      }
    }
  }
}

template <int LG>
__device__ __forceinline__ bool tile_wave(const Tile& T, TileWave& w) {
  constexpr int C = 64 >> LG;
  const int wv = pd_wave_id();
  if (wv >= T.nch) return false;
  w.ch = T.ch[wv];
  w.lane = pd_lane();
  w.q = w.lane & (C - 1);
  w.g = w.lane >> (6 - LG);
  w.p = w.ch.first + w.q;
  w.active = w.q < w.ch.count;
  return true;
}

// Packed words of a lane's bond groups, the next D groups' words in flight.
template <int D>
struct PkStream {
  const uint64_t* pk;
  uint32_t wb;
  int nt4;
  uint64_t w[D];
  __device__ __forceinline__ uint64_t word(int g) const {
    return g < nt4 ? gld(pk, wb + 64u * (uint32_t)g) : 0ull;
  }
  __device__ __forceinline__ PkStream(const uint64_t* p, uint32_t b, int n) : pk(p), wb(b), nt4(n) {
#pragma unroll
    for (int d = 0; d < D; ++d) w[d] = word(d);
  }
  // word of group g (called for g = 0, 1, 2, ... in order)
  __device__ __forceinline__ uint64_t next(int g) {
    const uint64_t cw = w[0];
#pragma unroll
    for (int d = 0; d + 1 < D; ++d) w[d] = w[d + 1];
    w[D - 1] = word(g + D);
    return cw;
  }
};

// Runs body(t4, word) over a lane's bond groups t4 = 0 .. nt4-1 in order, the
// next group's packed word in flight while a group is computed.  The words
// alternate between two variables, each reloaded only after its group ran, and
// every iteration issues both loads (beyond the last group they re-read it and
// the value goes unused), so the wait before a group's first use leaves exactly
// one load in flight.  body must issue no global memory instruction: loads and
// stores share one counter, and any would make that wait drain everything.
template <class F>
__device__ __forceinline__ void pk_groups(const uint64_t* __restrict__ pk, uint32_t wb, int nt4, F&& body) {
  if (nt4 <= 0) return;
  const int last = nt4 - 1;
  auto word = [&](int g) { return gld(pk, wb + 64u * (uint32_t)(g < last ? g : last)); };
  uint64_t w0 = word(0), w1 = word(1);
  for (int t4 = 0; t4 < nt4; t4 += 2) {
    body(t4, w0);
    w0 = word(t4 + 2);
    if (t4 + 1 < nt4) body(t4 + 1, w1);
    w1 = word(t4 + 3);
  }
}

// Records the bonds a lane broke in the bond loop just run: every intact bond
// for which broken(j) holds (the loop's own test, recomputed from the same
// staged records) loses its state byte and its pk intact bit.  Called after
// the loop by the lanes that saw a break, so that the loop stores nothing.
template <class F>
__device__ __forceinline__ void pk_record_breaks(const uint64_t* __restrict__ pk, uint16_t* __restrict__ pk16,
                                                 uint8_t* __restrict__ st, uint32_t wb, uint32_t base, int nt4,
                                                 F&& broken) {
  for (int t4 = 0; t4 < nt4; ++t4) {
    const uint64_t cw = gld(pk, wb + 64u * (uint32_t)t4);
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const uint32_t e = (uint32_t)(cw >> (16 * u)) & 0xffffu;
      if (!(e & 0x8000u) || !broken((int)(e & 0x7fffu))) continue;
      gst(st, base + 64u * (uint32_t)(4 * t4 + u), (uint8_t)0);
      gst(pk16, 4u * (wb + 64u * (uint32_t)t4) + (uint32_t)u, (uint16_t)(e & 0x7fffu));
    }
  }
}

// Epilogue of the tiled force kernels for particle p with bond sum a: a
// particle with contact terms stores a in f and is finished by
// k_contact_finish (contact terms in list order, then the Verlet update, as
// the untiled kernels do in their epilogue); the others are finished here.
__device__ __forceinline__ void finish_tiled(EnvDev& E, int p, const double4& xp, const double4& yp, double a1,
                                             double a2, double a3, int flags) {
  if (E.ncontacts > 0 && E.cflag[p]) {
    E.f[p] = make_double4(a1, a2, a3, 0.0);
    return;
  }
  finish_particle_v(E, p, xp, yp, a1, a2, a3, flags, gld4(E.v, p), E.bcid[p]);
}

// deformed bond vector to the staged neighbour j (expressions of the untiled kernels)
template <class R>
__device__ __forceinline__ void tile_y(const R& rc, int j, const double4& yp, double& Y1, double& Y2, double& Y3) {
  Y1 = rc.y(0, j) - yp.x; Y2 = rc.y(1, j) - yp.y; Y3 = rc.y(2, j) - yp.z;
}

// Reference bond lengths of integer lattices.  In a block whose reference
// positions are all integers (BlockDev::xint) every bond vector X is integral,
// so |X|^2 is an integer below kXTab and |X|, 1/|X| are read from a table the
// workgroup fills first, with the very expressions the general path evaluates
// (xij = xx2 * pd_rsqrt(xx2)): both paths give the same bits.
constexpr int kXTab = 32;
__device__ __forceinline__ void fill_xtab64(double2* __restrict__ t) {
  if (threadIdx.x < kXTab) {
    const double a = (double)threadIdx.x;
    const double r = pd_rsqrt(a);
    t[threadIdx.x] = make_double2(a * r, r);
  }
}

// The owner's reference position and where |X| comes from (FP64 kernels)
struct XGeo {
  double4 xp;
  uint2 lp;  // the owner's lat_pack (when the table is used)
  bool lat;  // the integer-lattice table applies (uniform over the workgroup)
  const double2* tab;
  __device__ __forceinline__ XGeo(const double4& x, bool l, const double2* t)
      : xp(x), lp(l ? lat_pack(x.x, x.y, x.z) : make_uint2(0u, 0u)), lat(l), tab(t) {}
  // |X| (xij) and 1/|X| of the bond to staged record j; LAT = lat, a
  // compile-time copy so that the bond loops carry no branch on it
  template <bool LAT, class R>
  __device__ __forceinline__ void xlen(const R& rc, int j, double& xij, double& inv_x) const {
    if constexpr (LAT) {
      const double2 v = *(const double2*)((const char*)tab + lat_offset(rc.lat(j), lp));
      xij = v.x;
      inv_x = v.y;
    } else {
      double a, b, c;
      rc.x(j, a, b, c);
      const double X1 = a - xp.x, X2 = b - xp.y, X3 = c - xp.z;
      const double xx2 = pd_norm2(X1, X2, X3);
      inv_x = pd_rsqrt(xx2);
      xij = xx2 * inv_x;
    }
  }
};

// Runs f(std::integral_constant<bool, lat>) -- a bond loop specialised for the
// lattice table or the general |X| -- where lat is uniform over the workgroup.
template <bool XF, class F, bool LAT = false>
__device__ __forceinline__ void with_lat(bool lat, F&& f, std::integral_constant<bool, LAT> = {}) {
  if constexpr (LAT) {  // compact lattice records: the table only
    f(std::true_type{});
    return;
  }
  if constexpr (XF) {
    if (lat) {
      f(std::true_type{});
      return;
    }
  }
  f(std::false_type{});
}

// ---------------------------------------------------------------------------
// dilatation theta (src/peri.jl:155-172), OSB and EP blocks
template <int LG, bool XF, int S>
__global__ __launch_bounds__(256) void k_tile_theta(const Tile* __restrict__ tiles, int ntiles,
                                                    EnvDev* __restrict__ envs, int hcap) {
  const int hs = S > 0 ? S : hcap;  // LDS stride of the record fields
  extern __shared__ double sh[];
  constexpr int C = 64 >> LG;
  const int t = pd_xcd_block(gridDim.x);
  if (t >= ntiles) return;
  const Tile& T = tiles[t];
  EnvDev& E = envs[T.env];
  const BlockDev& B = E.blocks[T.block];
  if (B.kind == PD_BB || !B.deformed) return;
  __shared__ double2 xtab[XF ? kXTab : 1];
  if (XF) fill_xtab64(xtab);
  const bool lat = XF && B.xint;
  stage_halo<XF, 0>(E, T, sh, hs, lat);
  __syncthreads();
  TileWave w;
  if (!tile_wave<LG>(T, w)) return;
  const Rec<XF, 0> rc(sh, hs);
  const uint64_t* __restrict__ pk = E.pk;
  const int p = w.active ? w.p : w.ch.first;
  const double4 xp = gld4(E.x, p), yp = gld4(E.y[E.cur], p);
  const XGeo xg(xp, lat, xtab);
  const int nt4 = (w.ch.width + 3) >> 2;
  const uint32_t wb = w.ch.pk0 + (uint32_t)w.lane;
  double acc = 0.0;
  with_lat<XF>(lat, [&](auto LT) {
    pk_groups(pk, wb, nt4, [&](int, uint64_t cw) {
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const uint32_t e = (uint32_t)(cw >> (16 * u)) & 0xffffu;
        if (!(e & 0x8000u)) continue;
        const int j = (int)(e & 0x7fffu);
        double Y1, Y2, Y3, xij, inv_x;
        tile_y(rc, j, yp, Y1, Y2, Y3);
        xg.template xlen<decltype(LT)::value>(rc, j, xij, inv_x);
        const double yy2 = pd_norm2(Y1, Y2, Y3);
        const double ext = yy2 * pd_rsqrt(yy2) - xij;
        acc = __builtin_fma(ext, rc.vol(j), acc);
      }
    });
  });
  acc = group_sum(acc, C);
  if (w.active && w.g == 0) {
    const double th = (3.0 * E.aux[p].y) * acc;
    E.theta[p] = th;
    E.aux[p].x = th;
  }
}

// ---------------------------------------------------------------------------
// td_norm2 (elasto_plastic.jl:168-172,282-295; src/peri.jl:411-445), EP
// blocks; breaks bonds with s >= s0 (state and the pk intact bit)
template <int LG, bool XF, int S>
__global__ __launch_bounds__(256) void k_tile_td(const Tile* __restrict__ tiles, int ntiles,
                                                 EnvDev* __restrict__ envs, int hcap) {
  const int hs = S > 0 ? S : hcap;
  extern __shared__ double sh[];
  constexpr int C = 64 >> LG;
  const int t = pd_xcd_block(gridDim.x);
  if (t >= ntiles) return;
  const Tile& T = tiles[t];
  EnvDev& E = envs[T.env];
  const BlockDev& B = E.blocks[T.block];
  if (B.kind != PD_EP || !B.deformed) return;
  __shared__ double2 xtab[XF ? kXTab : 1];
  if (XF) fill_xtab64(xtab);
  const bool lat = XF && B.xint;
  stage_halo<XF, 2>(E, T, sh, hs, lat);
  __syncthreads();
  TileWave w;
  if (!tile_wave<LG>(T, w)) return;
  const Rec<XF, 2> rc(sh, hs);
  const uint64_t* __restrict__ pk = E.pk;
  uint16_t* __restrict__ pk16 = (uint16_t*)E.pk;
  uint8_t* __restrict__ st = E.state;
  const double* __restrict__ edp = E.edp[E.edp_cur];
  double* __restrict__ edp_o = E.edp[E.edp_cur ^ 1];
  const int p = w.active ? w.p : w.ch.first;
  const double4 xp = gld4(E.x, p), yp = gld4(E.y[E.cur], p);
  const XGeo xg(xp, lat, xtab);
  const double4 ap = E.aux[p];  // (theta, 1/m, +-td_norm2 of the last step, C)
  const bool yielded = pd_has_yielded(ap);
  const double s0 = B.s0[0];
  const double g15 = 15.0 * B.G[0];
  const int nt4 = (w.ch.width + 3) >> 2;
  const uint32_t wb = w.ch.pk0 + (uint32_t)w.lane;
  const uint32_t base = w.ch.slot0 + (uint32_t)w.lane;
  double acc = 0.0;
  // own plastic extensions of the next group loaded while this one is computed
  auto load_eo = [&](uint64_t cw, int g, double* eo) {
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const bool on = (cw >> (16 * u + 15)) & 1ull;
      eo[u] = (on && yielded) ? gld(edp, base + 64u * (uint32_t)(4 * g + u)) : 0.0;
    }
  };
  with_lat<XF>(lat, [&](auto LT) {
    PkStream<3> ps(pk, wb, nt4);
    uint64_t cn = ps.next(0);
    double en[4];
    load_eo(cn, 0, en);
    for (int t4 = 0; t4 < nt4; ++t4) {
      const uint64_t cw = cn;
      double eo[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) eo[u] = en[u];
      cn = ps.next(t4 + 1);
      load_eo(cn, t4 + 1, en);
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const uint32_t e = (uint32_t)(cw >> (16 * u)) & 0xffffu;
        if (!(e & 0x8000u)) continue;
        const int j = (int)(e & 0x7fffu);
        const uint32_t slot = base + 64u * (uint32_t)(4 * t4 + u);
        double Y1, Y2, Y3, xij, inv_x;
        tile_y(rc, j, yp, Y1, Y2, Y3);
        xg.template xlen<decltype(LT)::value>(rc, j, xij, inv_x);
        const double yy2 = pd_norm2(Y1, Y2, Y3);
        const double ext = yy2 * pd_rsqrt(yy2) - xij;
        if (ext * inv_x < s0) {
          const double th_j = rc.th(j), rm_j = rc.rm(j);
          const double sm = ap.y + rm_j;
          const double twm = ap.x * ap.y + th_j * rm_j;
          const double ed = ext - (twm * pd_rcp(sm)) * xij * (1.0 / 3.0);
          const double td = g15 * (ed - eo[u]) * (inv_x * sm);
          acc = __builtin_fma(td * td, rc.vol(j), acc);
        } else {
          gst(st, slot, (uint8_t)0);
          gst(pk16, 4u * (wb + 64u * (uint32_t)t4) + (uint32_t)u, (uint16_t)(e & 0x7fffu));
          if (yielded) gst(edp_o, slot, eo[u]);  // else 0 in both buffers already
        }
      }
    }
  });
  acc = group_sum(acc, C);
  if (w.active && w.g == 0) {
    const double Cp = __builtin_sqrt((2.0 * B.psi[0]) / acc);
    const bool flag = yielded || !(Cp > 1.0);
    E.td[p] = acc;
    E.aux[p].z = flag ? -acc : acc;
    E.aux[p].w = Cp;
  }
}

// ---------------------------------------------------------------------------
// Elastoplastic bond loop of the tiled force pass, software-pipelined over
// groups of 4 bond steps: while group g is computed, the reverse bonds' plastic
// extensions of group g+1, the own plastic extensions and reverse slots of
// group g+2 and the packed words of groups g+3 and g+4 are in flight (126
// VGPRs: 4 waves per SIMD, as many as the compact LDS records allow).  Bonds are
// evaluated in the same order with the same expressions as the untiled kernel.
struct EpGrp {
  uint64_t cw;
  double eo[4], er[4];
  uint32_t rv[4];
  uint32_t need;  // bit u: the reverse bond's plastic extension is needed
};

template <class R>
__device__ __forceinline__ void ep_stage_a(EpGrp& G, uint64_t cw, uint32_t base, int t4, bool yielded_p,
                                           const double* __restrict__ edp_in, const uint32_t* __restrict__ rev,
                                           const R& rc) {
  G.cw = cw;
  G.need = 0;
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    const uint32_t e = (uint32_t)(cw >> (16 * u)) & 0xffffu;
    const bool on = (e & 0x8000u) != 0;
    const uint32_t slot = base + 64u * (uint32_t)(4 * t4 + u);
    G.eo[u] = (on && yielded_p) ? gld(edp_in, slot) : 0.0;
    const bool nd = on && __builtin_signbit(rc.cs((int)(e & 0x7fffu)));
    G.rv[u] = nd ? gld(rev, slot) : 0u;
    G.need |= (nd ? 1u : 0u) << u;
  }
}

__device__ __forceinline__ void ep_stage_b(EpGrp& G, const double* __restrict__ edp_in) {
#pragma unroll
  for (int u = 0; u < 4; ++u) G.er[u] = ((G.need >> u) & 1u) ? gld(edp_in, G.rv[u]) : 0.0;
}

template <bool LAT, class R>
__device__ __forceinline__ void ep_group(const EpGrp& G, int t4, uint32_t base, uint32_t wb, const XGeo& xg,
                                         const double4& yp, const double4& ap, bool yielded_p, double s0_pp,
                                         double K0, double G0, double inv_den, const R& rc,
                                         double* __restrict__ edp_out, uint8_t* __restrict__ st,
                                         uint16_t* __restrict__ pk16, double& a1, double& a2, double& a3) {
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    const uint32_t e = (uint32_t)(G.cw >> (16 * u)) & 0xffffu;
    if (!(e & 0x8000u)) continue;
    const int j = (int)(e & 0x7fffu);
    const uint32_t slot = base + 64u * (uint32_t)(4 * t4 + u);
    double Y1, Y2, Y3, xij, inv_x;
    tile_y(rc, j, yp, Y1, Y2, Y3);
    xg.template xlen<LAT>(rc, j, xij, inv_x);
    const double yy2 = pd_norm2(Y1, Y2, Y3);
    const double inv_y = pd_rsqrt(yy2);
    const double yij = yy2 * inv_y;
    const double ext = yij - xij;
    if (ext * inv_x < s0_pp) {
      const double th_j = rc.th(j), rm_j = rc.rm(j);
      const double sm = ap.y + rm_j;
      const double twm = ap.x * ap.y + th_j * rm_j;
      const double ed = ext - (twm * pd_rcp(sm)) * xij * (1.0 / 3.0);
      const double g15 = 15.0 * G0 * (inv_x * sm);
      const double Cp = ap.w, Cj = __builtin_fabs(rc.cs(j));
      const double ede = ed - G.eo[u];
      double td_own = g15 * ede;
      double e_new = G.eo[u];
      if (!(Cp > 1.0)) {
        td_own *= Cp;
        e_new = G.eo[u] + ede * (1.0 - Cp);
      }
      if (yielded_p) gst(edp_out, slot, e_new);  // else e_new = 0 = stored
      double td_rev = g15 * (ed - G.er[u]);
      if (!(Cj > 1.0)) td_rev *= Cj;
      const double t = 3.0 * K0 * twm + 0.5 * (td_own + td_rev);
      const double xx = t * inv_y * (rc.vol(j) * inv_den);
      a1 = __builtin_fma(xx, Y1, a1);
      a2 = __builtin_fma(xx, Y2, a2);
      a3 = __builtin_fma(xx, Y3, a3);
    } else {
      gst(st, slot, (uint8_t)0);
      gst(pk16, 4u * (wb + 64u * (uint32_t)t4) + (uint32_t)u, (uint16_t)(e & 0x7fffu));
      if (yielded_p) gst(edp_out, slot, G.eo[u]);
    }
  }
}

// ---------------------------------------------------------------------------
// fused force density + velocity-Verlet (material part of
// general_material.jl:296-369 with the plugins' t_ij, then contacts and the
// step epilogue of pd_wave.h)
// Waves per SIMD the register allocation aims for (measured best with the
// compact LDS records, which leave LDS for as many tiles per CU): BB 6 (80
// VGPRs), OSB 5 (96), EP 4 (126); the BB and OSB targets spill 20-28 B per lane
// outside the bond loop.
// LAT: every block of the launch is an integer lattice and its records are
// the compact lattice ones (bond-based blocks only); 7 workgroups of them fit
// a CU's LDS in the densest class, so the kernel aims at 7 waves per SIMD.
// ONEV: every block of the launch also has one volume; its records are 32
// bytes, 8 workgroups fit, and the kernel aims at 8 waves per SIMD (64 VGPRs:
// the compiler spills a few of the halo staging's registers, none of the bond
// loop's).
template <int KIND, int LG, bool XF, int S, bool LAT = false, bool ONEV = false>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(KIND == PD_BB ? (LAT ? (ONEV ? 8 : 7) : 6) : KIND == PD_OSB ? 5 : 4)))
void k_tile_force(const Tile* __restrict__ tiles, int ntiles,
                                                    EnvDev* __restrict__ envs, int flags, int hcap) {
  const int hs = S > 0 ? S : hcap;
  extern __shared__ double sh[];
  constexpr int C = 64 >> LG;
  constexpr int NX = tile_extra<TP_FORCE, KIND>();
  const int t = pd_xcd_block(gridDim.x);
  if (t >= ntiles) return;
  const Tile& T = tiles[t];
  EnvDev& E = envs[T.env];
  const BlockDev& B = E.blocks[T.block];
  const bool deformed = B.deformed != 0;  // uniform over the workgroup
  TileWave w;
  const bool has = tile_wave<LG>(T, w);
  const int p = has ? (w.active ? w.p : w.ch.first) : T.ch[0].first;
  __shared__ double2 xtab[XF ? kXTab : 1];
  if (XF) fill_xtab64(xtab);
  const bool lat = LAT || (XF && B.xint);
  if (deformed) stage_halo<XF, NX, LAT, ONEV>(E, T, sh, hs, lat);
  __syncthreads();
  if (!has) return;
  const Rec<XF, NX, LAT> rc(sh, hs);
  const double4* __restrict__ y = E.y[E.cur];
  const double4 xp = gld4(E.x, p), yp = gld4(y, p);
  const XGeo xg(xp, lat, xtab);
  double a1 = 0.0, a2 = 0.0, a3 = 0.0;
  if (deformed) {
    const uint64_t* __restrict__ pk = E.pk;
    uint16_t* __restrict__ pk16 = (uint16_t*)E.pk;
    uint8_t* __restrict__ st = E.state;
    const double inv_den = 1.0 / B.density[0];
    const int nt4 = (w.ch.width + 3) >> 2;
    const uint32_t wb = w.ch.pk0 + (uint32_t)w.lane;
    const uint32_t base = w.ch.slot0 + (uint32_t)w.lane;
    const double s0_pp = B.s0[0], K0 = B.K[0], G0 = B.G[0];
    if constexpr (KIND == PD_EP) {
      const double4 ap = E.aux[p];
      const double* __restrict__ edp_in = E.edp[E.edp_cur];
      double* __restrict__ edp_out = E.edp[E.edp_cur ^ 1];
      const bool yielded_p = pd_has_yielded(ap);
      with_lat<XF>(lat, [&](auto LT) {
        PkStream<2> ps(pk, wb, nt4);
        EpGrp g0, g1, g2;
        ep_stage_a(g0, ps.next(0), base, 0, yielded_p, edp_in, E.rev, rc);
        ep_stage_b(g0, edp_in);
        ep_stage_a(g1, ps.next(1), base, 1, yielded_p, edp_in, E.rev, rc);
        for (int t4 = 0; t4 < nt4; ++t4) {
          ep_stage_b(g1, edp_in);
          ep_stage_a(g2, ps.next(t4 + 2), base, t4 + 2, yielded_p, edp_in, E.rev, rc);
          ep_group<decltype(LT)::value>(g0, t4, base, wb, xg, yp, ap, yielded_p, s0_pp, K0, G0, inv_den, rc,
                                        edp_out, st, pk16, a1, a2, a3);
          g0 = g1;
          g1 = g2;
        }
      }, std::integral_constant<bool, LAT>{});
    } else {
      // bond-based and ordinary state-based blocks
      double4 ap = make_double4(0.0, 0.0, 0.0, 0.0);
      if (KIND != PD_BB) ap = E.aux[p];
      const double c11 = B.c11;
      bool brk = false;  // this lane broke a bond
      // a lattice block of one volume: vol_j / density is the same for every
      // bond (the value the records give, formed once)
      const double vd0 = B.vol0 * inv_den;
      with_lat<XF>(lat, [&](auto LT) {
        auto bonds = [&](auto U) {
        pk_groups(pk, wb, nt4, [&](int, uint64_t cw) {
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            const uint32_t e = (uint32_t)(cw >> (16 * u)) & 0xffffu;
            if (!(e & 0x8000u)) continue;
            const int j = (int)(e & 0x7fffu);
            double Y1, Y2, Y3, xij, inv_x;
            tile_y(rc, j, yp, Y1, Y2, Y3);
            xg.template xlen<decltype(LT)::value>(rc, j, xij, inv_x);
            const double yy2 = pd_norm2(Y1, Y2, Y3);
            const double inv_y = pd_rsqrt(yy2);
            const double yij = yy2 * inv_y;
            const double ext = yij - xij;
            // the bond's term, dropped by the select when it breaks (adding
            // fma(0, Y, a) leaves a unchanged: Y is finite and a never -0),
            // so that the loop carries no branch but the intact test
            const bool ok = ext * inv_x < s0_pp;
            brk = brk || !ok;
            double tt;
            if (KIND == PD_BB) {
              tt = ext * inv_x * c11;
            } else {
              const double th_j = rc.th(j), rm_j = rc.rm(j);
              const double sm = ap.y + rm_j;
              const double twm = ap.x * ap.y + th_j * rm_j;
              tt = (3.0 * K0 - 5.0 * G0) * twm + 15.0 * G0 * ext * (inv_x * sm);
            }
            const double xt = tt * inv_y * (decltype(U)::value ? vd0 : rc.vol(j) * inv_den);
            const double xx = ok ? xt : 0.0;
            a1 = __builtin_fma(xx, Y1, a1);
            a2 = __builtin_fma(xx, Y2, a2);
            a3 = __builtin_fma(xx, Y3, a3);
          }
        });
        };
        if constexpr (ONEV) {
          bonds(std::true_type{});
        } else if constexpr (LAT) {
          if (B.vol1) bonds(std::true_type{});
          else bonds(std::false_type{});
        } else {
          bonds(std::false_type{});
        }
        if (brk)
          pk_record_breaks(pk, pk16, st, wb, base, nt4, [&](int j) {
            double Y1, Y2, Y3, xij, inv_x;
            tile_y(rc, j, yp, Y1, Y2, Y3);
            xg.template xlen<decltype(LT)::value>(rc, j, xij, inv_x);
            const double yy2 = pd_norm2(Y1, Y2, Y3);
            const double inv_y = pd_rsqrt(yy2);
            const double yij = yy2 * inv_y;
            const double ext = yij - xij;
            return !(ext * inv_x < s0_pp);
          });
      }, std::integral_constant<bool, LAT>{});
    }
    a1 = group_sum(a1, C);
    a2 = group_sum(a2, C);
    a3 = group_sum(a3, C);
  }
  if (!w.active || w.g != 0) return;
  finish_tiled(E, p, xp, yp, a1, a2, a3, flags);
}

// ---------------------------------------------------------------------------
// FP32-compensated form (BB / OSB blocks of PD_PREC_FP32C envs; the bond
// arithmetic of pd_step32.hip).  LDS record, every field 8 bytes wide so that one
// address j * 8 reaches them all: the FP64 displacement w = y - (corner + x32)
// (3 doubles, formed here once per halo record), then float pairs (x32.x,
// x32.y), (x32.z, v) and for OSB forces (theta, 1/m).  v is the volume for
// theta, vol / density for OSB forces and c11 vol / density for BB forces (the
// neighbour factors the untiled kernels form per bond, with the same roundings).
enum { ST_THETA = 0, ST_BB = 1, ST_OSB = 2 };

template <int MODE>
__device__ __forceinline__ void stage_halo32(const EnvDev& E, const BlockDev& B, const Tile& T,
                                             double* __restrict__ sh, int hcap) {
  const float4* __restrict__ x = E.x32;
  const double4* __restrict__ y = E.y[E.cur];
  const int32_t* __restrict__ halo = E.halo;
  float2* __restrict__ f2 = (float2*)(sh + 3 * hcap);
  const float inv_den = (float)(1.0 / B.density[0]);
  const float c11 = (float)B.c11;
  const double3 corner = corner_of(B);  // read once: the loop's LDS stores could alias B for the compiler
  for (int i0 = threadIdx.x; i0 < T.hn; i0 += 256 * kStageR) {
    uint32_t id[kStageR];
#pragma unroll
    for (int r = 0; r < kStageR; ++r) {
      const int i = i0 + 256 * r;
      id[r] = i < T.hn ? (uint32_t)gld(halo, T.hoff + (uint32_t)i) : 0u;
    }
    float4 a[kStageR], c[kStageR];
    double4 b[kStageR];
#pragma unroll
    for (int r = 0; r < kStageR; ++r) {
      if (i0 + 256 * r < T.hn) {
        a[r] = gld4(x, id[r]);
        b[r] = gld4(y, id[r]);
        if (MODE == ST_OSB) c[r] = gld4(E.aux32, id[r]);
      }
    }
#pragma unroll
    for (int r = 0; r < kStageR; ++r) {
      const int i = i0 + 256 * r;
      if (i < T.hn) {
        const double3 wd = disp32(corner, a[r], b[r]);
        sh[i] = wd.x;
        sh[hcap + i] = wd.y;
        sh[2 * hcap + i] = wd.z;
        const float v = MODE == ST_THETA ? a[r].w : MODE == ST_BB ? cvd32(c11, a[r].w, inv_den) : vd32(a[r].w, inv_den);
        f2[i] = make_float2(a[r].x, a[r].y);
        f2[hcap + i] = make_float2(a[r].z, v);
        if (MODE == ST_OSB) f2[2 * hcap + i] = make_float2(c[r].x, c[r].y);
      }
    }
  }
}

// staged record j: reference offsets (xyz) with v in w, and the displacement
struct Rec32 {
  float4 x;
  double3 w;
};
__device__ __forceinline__ Rec32 rec32(const double* __restrict__ sh, const float2* __restrict__ f2, int hcap,
                                       int j) {
  const float2 a = f2[j], b = f2[hcap + j];
  Rec32 r;
  r.x = make_float4(a.x, a.y, b.x, b.y);
  r.w = make_double3(sh[j], sh[hcap + j], sh[2 * hcap + j]);
  return r;
}

// dilatation theta (src/peri.jl:155-172) of OSB blocks; S as in the FP64
// kernels (compile-time LDS stride, 0: the run-time capacity)
template <int LG, int S>
__global__ __launch_bounds__(256) void k_tile_theta32(const Tile* __restrict__ tiles, int ntiles,
                                                      EnvDev* __restrict__ envs, int hcap_rt) {
  const int hcap = S > 0 ? S : hcap_rt;
  extern __shared__ double sh[];
  constexpr int C = 64 >> LG;
  const int t = pd_xcd_block(gridDim.x);
  if (t >= ntiles) return;
  const Tile& T = tiles[t];
  EnvDev& E = envs[T.env];
  const BlockDev& B = E.blocks[T.block];
  if (B.kind != PD_OSB || !B.deformed) return;
  stage_halo32<ST_THETA>(E, B, T, sh, hcap);
  __syncthreads();
  TileWave w;
  if (!tile_wave<LG>(T, w)) return;
  const float2* __restrict__ f2 = (const float2*)(sh + 3 * hcap);
  const uint64_t* __restrict__ pk = E.pk;
  const int p = w.active ? w.p : w.ch.first;
  const float4 xp = gld4(E.x32, p);
  const double3 wp = disp32(B, xp, gld4(E.y[E.cur], p));
  const int nt4 = (w.ch.width + 3) >> 2;
  const uint32_t wb = w.ch.pk0 + (uint32_t)w.lane;
  GroupKahan acc;
  pk_groups(pk, wb, nt4, [&](int t4, uint64_t cw) {
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const uint32_t e = (uint32_t)(cw >> (16 * u)) & 0xffffu;
      if (!(e & 0x8000u)) continue;
      const int j = (int)(e & 0x7fffu);
      const Rec32 r = rec32(sh, f2, hcap, j);
      G32 g;
      geom32(xp, wp, r.x, r.w, g);
      acc.add(g.ext, r.x.w);
    }
    if (group_end_word(t4, nt4)) acc.flush();
  });
  const double tot = comp_total(acc.s, acc.c, C);
  if (w.active && w.g == 0) {
    const float th = (float)((3.0 * (double)E.aux32[p].y) * tot);
    E.aux32[p].x = th;
    E.theta[p] = (double)th;
  }
}

// fused force density (BB / OSB) + FP64 velocity-Verlet
template <int KIND, int LG, int S>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(KIND == PD_BB ? 6 : 5)))
void k_tile_force32(const Tile* __restrict__ tiles, int ntiles,
                                                      EnvDev* __restrict__ envs, int flags, int hcap_rt) {
  const int hcap = S > 0 ? S : hcap_rt;
  extern __shared__ double sh[];
  constexpr int C = 64 >> LG;
  const int t = pd_xcd_block(gridDim.x);
  if (t >= ntiles) return;
  const Tile& T = tiles[t];
  EnvDev& E = envs[T.env];
  const BlockDev& B = E.blocks[T.block];
  const bool deformed = B.deformed != 0;  // uniform over the workgroup
  TileWave w;
  const bool has = tile_wave<LG>(T, w);
  const int p = has ? (w.active ? w.p : w.ch.first) : T.ch[0].first;
  if (deformed) stage_halo32<KIND == PD_OSB ? ST_OSB : ST_BB>(E, B, T, sh, hcap);
  __syncthreads();
  if (!has) return;
  const double4* __restrict__ y = E.y[E.cur];
  const double4 yp = gld4(y, p);
  double a1 = 0.0, a2 = 0.0, a3 = 0.0;
  if (deformed) {
    const float2* __restrict__ f2 = (const float2*)(sh + 3 * hcap);
    const uint64_t* __restrict__ pk = E.pk;
    uint16_t* __restrict__ pk16 = (uint16_t*)E.pk;
    uint8_t* __restrict__ st = E.state;
    const float4 xp = gld4(E.x32, p);
    const double3 wp = disp32(B, xp, yp);
    float4 ap = make_float4(0.f, 0.f, 0.f, 0.f);
    if (KIND == PD_OSB) ap = E.aux32[p];
    const float apxy = ap.x * ap.y;
    const float s0f = B.s0f_up;  // (double)s < s0 exactly when s < s0f
    const float Kf = (float)B.K[0], Gf = (float)B.G[0];
    const float A = 3.0f * Kf - 5.0f * Gf, B15 = 15.0f * Gf;
    GroupKahan k1, k2, k3;
    const int nt4 = (w.ch.width + 3) >> 2;
    const uint32_t wb = w.ch.pk0 + (uint32_t)w.lane;
    const uint32_t base = w.ch.slot0 + (uint32_t)w.lane;
    bool brk = false;  // this lane broke a bond
    pk_groups(pk, wb, nt4, [&](int t4, uint64_t cw) {
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const uint32_t e = (uint32_t)(cw >> (16 * u)) & 0xffffu;
        if (!(e & 0x8000u)) continue;
        const int j = (int)(e & 0x7fffu);
        const Rec32 r = rec32(sh, f2, hcap, j);
        float2 am = make_float2(0.f, 0.f);
        if (KIND == PD_OSB) am = f2[2 * hcap + j];
        G32 g;
        geom32(xp, wp, r.x, r.w, g);
        // (t / |Y|) vol_j / density, formed for every bond; 0 for one broken here
        const float tt = (KIND == PD_BB ? g.s : osb_t32(ap, apxy, am.x, am.y, A, B15, g)) * g.inv_y;
        const bool intact = g.s < s0f;
        brk = brk || !intact;
        const float xx = intact ? tt * r.x.w : 0.0f;
        k1.add(xx, g.Y1);
        k2.add(xx, g.Y2);
        k3.add(xx, g.Y3);
      }
      if (group_end_word(t4, nt4)) {
        k1.flush();
        k2.flush();
        k3.flush();
      }
    });
    if (brk)
      pk_record_breaks(pk, pk16, st, wb, base, nt4, [&](int j) {
        const Rec32 r = rec32(sh, f2, hcap, j);
        G32 g;
        geom32(xp, wp, r.x, r.w, g);
        return !(g.s < s0f);
      });
    a1 = comp_total(k1.s, k1.c, C);
    a2 = comp_total(k2.s, k2.c, C);
    a3 = comp_total(k3.s, k3.c, C);
  }
  if (!w.active || w.g != 0) return;
  finish_tiled(E, p, gld4(E.x, p), yp, a1, a2, a3, flags);
}

// ---------------------------------------------------------------------------
// Device build of the tiles' halos and packed streams (one workgroup per
// tile).  The neighbour ids of the tile's bond slots are inserted into an LDS
// hash set of hc slots (open addressing, hc >= twice the slot count, so every
// probe sequence ends), the distinct ids are compacted, sorted (bitonic) and
// each bond's index in the sorted halo is found by bisection.  Every bond of a
// freshly built family is intact.

__device__ __forceinline__ void lds_bitonic_sort(int32_t* __restrict__ buf, int n) {
  for (int k = 2; k <= n; k <<= 1)
    for (int j = k >> 1; j > 0; j >>= 1) {
      for (int i = threadIdx.x; i < n; i += blockDim.x) {
        const int ixj = i ^ j;
        if (ixj > i) {
          const int32_t a = buf[i], b = buf[ixj];
          const bool up = (i & k) == 0;
          if ((a > b) == up) {
            buf[i] = b;
            buf[ixj] = a;
          }
        }
      }
      __syncthreads();
    }
}

// inserts every neighbour id of tile T's slots into table[hc] (-1 = empty)
__device__ void tile_hash_ids(const Tile& T, const int32_t* __restrict__ col, int32_t* __restrict__ table, int hc) {
  for (int i = threadIdx.x; i < hc; i += blockDim.x) table[i] = -1;
  __syncthreads();
  const int lg = __builtin_ctz(hc);
  for (int k = 0; k < T.nch; ++k) {
    const Chunk& ch = T.ch[k];
    const int n = PD_WAVE * ch.width;
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
      const int32_t j = col[ch.slot0 + (uint32_t)i];
      if (j < 0) continue;
      uint32_t h = ((uint32_t)j * 2654435761u) >> (32 - lg);
      while (true) {
        const int32_t old = atomicCAS(&table[h], -1, j);
        if (old == -1 || old == j) break;
        h = (h + 1) & (uint32_t)(hc - 1);
      }
    }
  }
  __syncthreads();
}

// exclusive positions of the occupied table slots over per-thread segments;
// part[blockDim.x] receives the total
__device__ int occupied_scan(const int32_t* table, int hc, int* part, int& seg0, int& seg1) {
  const int per = hc / (int)blockDim.x > 0 ? hc / (int)blockDim.x : 1;
  seg0 = min((int)threadIdx.x * per, hc);
  seg1 = (threadIdx.x + 1 == blockDim.x) ? hc : min(seg0 + per, hc);
  int cnt = 0;
  for (int i = seg0; i < seg1; ++i) cnt += table[i] >= 0 ? 1 : 0;
  part[threadIdx.x] = cnt;
  __syncthreads();
  if (threadIdx.x == 0) {
    int acc = 0;
    for (int t = 0; t < (int)blockDim.x; ++t) {
      const int c = part[t];
      part[t] = acc;
      acc += c;
    }
    part[blockDim.x] = acc;
  }
  __syncthreads();
  return part[threadIdx.x];
}

__global__ __launch_bounds__(256) void k_halo_count(const Tile* __restrict__ tiles, int ntiles,
                                                    const int32_t* __restrict__ col, int32_t* __restrict__ counts,
                                                    int hc) {
  extern __shared__ int32_t sbuf[];
  if ((int)blockIdx.x >= ntiles) return;
  const Tile T = tiles[blockIdx.x];
  int32_t* table = sbuf;
  int* part = sbuf + hc;
  tile_hash_ids(T, col, table, hc);
  int s0, s1;
  occupied_scan(table, hc, part, s0, s1);
  if (threadIdx.x == 0) counts[blockIdx.x] = part[blockDim.x];
}

__global__ __launch_bounds__(256) void k_halo_fill(const Tile* __restrict__ tiles, int ntiles,
                                                   const int32_t* __restrict__ col, int32_t* __restrict__ halo,
                                                   uint64_t* __restrict__ pk, int hc) {
  extern __shared__ int32_t sbuf[];
  if ((int)blockIdx.x >= ntiles) return;
  const Tile T = tiles[blockIdx.x];
  int32_t* buf = sbuf;
  int* part = sbuf + hc;
  tile_hash_ids(T, col, buf, hc);
  int s0, s1;
  int pos = occupied_scan(buf, hc, part, s0, s1);
  for (int i = s0; i < s1; ++i)
    if (buf[i] >= 0) halo[T.hoff + (uint32_t)(pos++)] = buf[i];
  __syncthreads();
  // the distinct ids back into LDS, sorted
  const int hn = T.hn;
  int P = 1;
  while (P < hn) P <<= 1;
  for (int i = threadIdx.x; i < P; i += blockDim.x) buf[i] = i < hn ? halo[T.hoff + (uint32_t)i] : INT32_MAX;
  __syncthreads();
  lds_bitonic_sort(buf, P);
  for (int i = threadIdx.x; i < hn; i += blockDim.x) halo[T.hoff + (uint32_t)i] = buf[i];
  for (int k = 0; k < T.nch; ++k) {
    const Chunk& ch = T.ch[k];
    const int nt4 = (ch.width + 3) >> 2;
    for (int w = threadIdx.x; w < PD_WAVE * nt4; w += blockDim.x) {
      const int t4 = w / PD_WAVE, lane = w % PD_WAVE;
      uint64_t word = 0;
      for (int u = 0; u < 4; ++u) {
        const int st = 4 * t4 + u;
        if (st >= ch.width) break;
        const int32_t j = col[ch.slot0 + (uint32_t)(PD_WAVE * st + lane)];
        if (j < 0) continue;
        int lo = 0, hi = hn;  // lower_bound of j in buf[0, hn)
        while (lo < hi) {
          const int mid = (lo + hi) >> 1;
          if (buf[mid] < j) lo = mid + 1;
          else hi = mid;
        }
        word |= (uint64_t)(0x8000u | (uint32_t)lo) << (16 * u);
      }
      pk[ch.pk0 + (uint32_t)w] = word;
    }
  }
}

// pk intact bits := state (after bonds were broken by the untiled passes or the
// state was loaded); one wavefront per chunk
__global__ __launch_bounds__(256) void k_pk_sync(const Chunk* __restrict__ chunks, int nchunks,
                                                 EnvDev* __restrict__ envs) {
  const int cid = blockIdx.x * PD_WAVES_PER_BLOCK + pd_wave_id();
  if (cid >= nchunks) return;
  const Chunk ch = chunks[cid];
  EnvDev& E = envs[ch.env];
  const int lane = pd_lane();
  const int nt4 = (ch.width + 3) >> 2;
  for (int t4 = 0; t4 < nt4; ++t4) {
    const uint32_t wi = ch.pk0 + 64u * (uint32_t)t4 + (uint32_t)lane;
    uint64_t wd = gld(E.pk, wi);
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int s = 4 * t4 + u;
      const uint64_t bit = 1ull << (16 * u + 15);
      const bool on = s < ch.width && gld(E.state, ch.slot0 + 64u * (uint32_t)s + (uint32_t)lane);
      wd = on ? (wd | bit) : (wd & ~bit);
    }
    gst(E.pk, wi, wd);
  }
}

// LDS stride of the record fields (records per field): compile-time strides
// (the field offsets then fold into the LDS instructions' immediate offsets)
// sized so that 6, 5, 4 or 3 workgroups fit a CU's 160 KiB for the kernel's own
// record size, else the halo capacity at run time.  A launch takes the smallest
// stride its largest halo fits.  The byte budgets are multiples of 1280 (the
// coarsest LDS allocation granule, 160 KiB / 128) whose multiples by 6, 5, 4, 3
// stay within 160 KiB also when rounded up to 512-byte granules.
// The FP64 kernels' records leave kXTabLds bytes for their static lattice
// table (RES); the FP32C kernels have none (RES = 0).
constexpr int kLdsClassBytes[4] = {26880, 32000, 40960, 53760};
constexpr int kXTabLds = 512;
constexpr int halo_stride(int c, int rec, int reserve) { return (kLdsClassBytes[c] - reserve) / rec; }
// The class of 7 workgroups per CU (45 allocation granules of 512 bytes),
// for the compact lattice records of the bond-based force pass only: its
// kernel fits 7 waves per SIMD in registers, the other FP64 kernels do not.
constexpr int kLdsClass7Bytes = 23040;
constexpr int halo_stride7(int rec, int reserve) { return (kLdsClass7Bytes - reserve) / rec; }
// and of 8 (40 granules) for the one-volume lattice records of the same pass
constexpr int kLdsClass8Bytes = 20480;
constexpr int halo_stride8(int rec, int reserve) { return (kLdsClass8Bytes - reserve) / rec; }

template <int REC, int RES = kXTabLds, class F>
void with_stride(int hcap, F&& f) {
  constexpr int s0 = halo_stride(0, REC, RES), s1 = halo_stride(1, REC, RES);
  constexpr int s2 = halo_stride(2, REC, RES), s3 = halo_stride(3, REC, RES);
  if (hcap <= s0) f(std::integral_constant<int, s0>{});
  else if (hcap <= s1) f(std::integral_constant<int, s1>{});
  else if (hcap <= s2) f(std::integral_constant<int, s2>{});
  else if (hcap <= s3) f(std::integral_constant<int, s3>{});
  else f(std::integral_constant<int, 0>{});
}
template <int REC, int RES = kXTabLds, class F>
void with_stride7(int hcap, F&& f) {
  constexpr int s7 = halo_stride7(REC, RES);
  if (hcap <= s7) f(std::integral_constant<int, s7>{});
  else with_stride<REC, RES>(hcap, f);
}
template <int REC, int RES = kXTabLds, class F>
void with_stride8(int hcap, F&& f) {
  constexpr int s8 = halo_stride8(REC, RES);
  if (hcap <= s8) f(std::integral_constant<int, s8>{});
  else with_stride<REC, RES>(hcap, f);
}
template <int REC, int RES = kXTabLds, class F>
void each_stride(F&& f) {
  f(std::integral_constant<int, halo_stride(0, REC, RES)>{});
  f(std::integral_constant<int, halo_stride(1, REC, RES)>{});
  f(std::integral_constant<int, halo_stride(2, REC, RES)>{});
  f(std::integral_constant<int, halo_stride(3, REC, RES)>{});
  f(std::integral_constant<int, 0>{});
}

constexpr int rec32_bytes(int pass, int kind) {
  return (int)(3 * sizeof(double) + ((pass == TP_FORCE && kind == PD_OSB) ? 6 : 4) * sizeof(float));
}

template <int LG, bool XF, int S>
void launch_theta_s(hipStream_t s, const Tile* tiles, int n, EnvDev* envs, int hcap) {
  const size_t lds = (size_t)tile_record_bytes<TP_THETA, 0, XF>() * (size_t)(S > 0 ? S : hcap);
  hipLaunchKernelGGL((k_tile_theta<LG, XF, S>), dim3(n), dim3(256), lds, s, tiles, n, envs, hcap);
}
template <int LG, bool XF, int S>
void launch_td_s(hipStream_t s, const Tile* tiles, int n, EnvDev* envs, int hcap) {
  const size_t lds = (size_t)tile_record_bytes<TP_TD, 0, XF>() * (size_t)(S > 0 ? S : hcap);
  hipLaunchKernelGGL((k_tile_td<LG, XF, S>), dim3(n), dim3(256), lds, s, tiles, n, envs, hcap);
}
template <int KIND, int LG, bool XF, int S, bool LAT = false, bool ONEV = false>
void launch_force_s(hipStream_t s, const Tile* tiles, int n, EnvDev* envs, int flags, int hcap) {
  const size_t lds = (size_t)tile_record_bytes<TP_FORCE, KIND, XF, LAT, ONEV>() * (size_t)(S > 0 ? S : hcap);
  hipLaunchKernelGGL((k_tile_force<KIND, LG, XF, S, LAT, ONEV>), dim3(n), dim3(256), lds, s, tiles, n, envs, flags,
                     hcap);
}

template <int LG, bool XF>
void launch_theta_lg(hipStream_t s, const Tile* tiles, int n, EnvDev* envs, int hcap) {
  with_stride<tile_record_bytes<TP_THETA, 0, XF>()>(
      hcap, [&](auto S) { launch_theta_s<LG, XF, decltype(S)::value>(s, tiles, n, envs, hcap); });
}
template <int LG, bool XF>
void launch_td_lg(hipStream_t s, const Tile* tiles, int n, EnvDev* envs, int hcap) {
  with_stride<tile_record_bytes<TP_TD, 0, XF>()>(
      hcap, [&](auto S) { launch_td_s<LG, XF, decltype(S)::value>(s, tiles, n, envs, hcap); });
}
template <int KIND, int LG, bool XF>
void launch_force_lg(hipStream_t s, const Tile* tiles, int n, EnvDev* envs, int flags, int hcap) {
  with_stride<tile_record_bytes<TP_FORCE, KIND, XF>()>(
      hcap, [&](auto S) { launch_force_s<KIND, LG, XF, decltype(S)::value>(s, tiles, n, envs, flags, hcap); });
}
template <int LG>
void launch_force_lat_lg(hipStream_t s, const Tile* tiles, int n, EnvDev* envs, int flags, int hcap) {
  with_stride7<tile_record_bytes<TP_FORCE, PD_BB, true, true>()>(hcap, [&](auto S) {
    launch_force_s<PD_BB, LG, true, decltype(S)::value, true>(s, tiles, n, envs, flags, hcap);
  });
}
template <int LG>
void launch_force_onev_lg(hipStream_t s, const Tile* tiles, int n, EnvDev* envs, int flags, int hcap) {
  with_stride8<tile_record_bytes<TP_FORCE, PD_BB, true, true, true>()>(hcap, [&](auto S) {
    launch_force_s<PD_BB, LG, true, decltype(S)::value, true, true>(s, tiles, n, envs, flags, hcap);
  });
}

template <int LG, int S>
void launch_theta32_s(hipStream_t s, const Tile* tiles, int n, EnvDev* envs, int hcap) {
  const size_t lds = (size_t)rec32_bytes(TP_THETA, PD_OSB) * (size_t)(S > 0 ? S : hcap);
  hipLaunchKernelGGL((k_tile_theta32<LG, S>), dim3(n), dim3(256), lds, s, tiles, n, envs, hcap);
}
template <int LG>
void launch_theta32_lg(hipStream_t s, const Tile* tiles, int n, EnvDev* envs, int hcap) {
  with_stride<rec32_bytes(TP_THETA, PD_OSB), 0>(
      hcap, [&](auto S) { launch_theta32_s<LG, decltype(S)::value>(s, tiles, n, envs, hcap); });
}
template <int KIND, int LG, int S>
void launch_force32_s(hipStream_t s, const Tile* tiles, int n, EnvDev* envs, int flags, int hcap) {
  const size_t lds = (size_t)rec32_bytes(TP_FORCE, KIND) * (size_t)(S > 0 ? S : hcap);
  hipLaunchKernelGGL((k_tile_force32<KIND, LG, S>), dim3(n), dim3(256), lds, s, tiles, n, envs, flags, hcap);
}
template <int KIND, int LG>
void launch_force32_lg(hipStream_t s, const Tile* tiles, int n, EnvDev* envs, int flags, int hcap) {
  with_stride<rec32_bytes(TP_FORCE, KIND), 0>(
      hcap, [&](auto S) { launch_force32_s<KIND, LG, decltype(S)::value>(s, tiles, n, envs, flags, hcap); });
}

// dynamic LDS beyond the default 64 KiB per workgroup must be enabled per kernel
template <class K>
void allow_lds(K* k, size_t bytes) {
  (void)hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
}
template <int LG>
void allow_lds_fp32c(size_t bytes) {
  each_stride<rec32_bytes(TP_THETA, PD_OSB), 0>([&](auto S) { allow_lds(k_tile_theta32<LG, decltype(S)::value>, bytes); });
  each_stride<rec32_bytes(TP_FORCE, PD_BB), 0>(
      [&](auto S) { allow_lds(k_tile_force32<PD_BB, LG, decltype(S)::value>, bytes); });
  each_stride<rec32_bytes(TP_FORCE, PD_OSB), 0>(
      [&](auto S) { allow_lds(k_tile_force32<PD_OSB, LG, decltype(S)::value>, bytes); });
}
template <int LG, bool XF>
void allow_lds_fp64(size_t bytes) {
  each_stride<tile_record_bytes<TP_THETA, 0, XF>()>(
      [&](auto S) { allow_lds(k_tile_theta<LG, XF, decltype(S)::value>, bytes); });
  each_stride<tile_record_bytes<TP_TD, 0, XF>()>(
      [&](auto S) { allow_lds(k_tile_td<LG, XF, decltype(S)::value>, bytes); });
  each_stride<tile_record_bytes<TP_FORCE, PD_BB, XF>()>(
      [&](auto S) { allow_lds(k_tile_force<PD_BB, LG, XF, decltype(S)::value>, bytes); });
  each_stride<tile_record_bytes<TP_FORCE, PD_OSB, XF>()>(
      [&](auto S) { allow_lds(k_tile_force<PD_OSB, LG, XF, decltype(S)::value>, bytes); });
  each_stride<tile_record_bytes<TP_FORCE, PD_EP, XF>()>(
      [&](auto S) { allow_lds(k_tile_force<PD_EP, LG, XF, decltype(S)::value>, bytes); });
  if constexpr (XF) {
    constexpr int rec = tile_record_bytes<TP_FORCE, PD_BB, true, true>();
    allow_lds(k_tile_force<PD_BB, LG, true, halo_stride7(rec, kXTabLds), true>, bytes);
    each_stride<rec>([&](auto S) { allow_lds(k_tile_force<PD_BB, LG, true, decltype(S)::value, true>, bytes); });
    constexpr int rec1 = tile_record_bytes<TP_FORCE, PD_BB, true, true, true>();
    allow_lds(k_tile_force<PD_BB, LG, true, halo_stride8(rec1, kXTabLds), true, true>, bytes);
    each_stride<rec1>(
        [&](auto S) { allow_lds(k_tile_force<PD_BB, LG, true, decltype(S)::value, true, true>, bytes); });
  }
}
template <int LG>
void allow_lds_lg(size_t bytes) {
  allow_lds_fp32c<LG>(bytes);
  allow_lds_fp64<LG, false>(bytes);
  allow_lds_fp64<LG, true>(bytes);
}

template <int PASS, bool XF>
int fp64_record_bytes(int kind) {
  return kind == PD_BB ? tile_record_bytes<PASS, PD_BB, XF>()
                       : kind == PD_OSB ? tile_record_bytes<PASS, PD_OSB, XF>() : tile_record_bytes<PASS, PD_EP, XF>();
}

}  // namespace

int pd_tile_record_bytes(int pass, int kind, int prec, int xf) {
  if (prec == PD_PREC_FP32C && kind != PD_EP && pass != TP_TD) return rec32_bytes(pass, kind);
  if (xf == 2 && pass == TP_FORCE && kind == PD_BB) return tile_record_bytes<TP_FORCE, PD_BB, true, true>();
  if (xf == 3 && pass == TP_FORCE && kind == PD_BB) return tile_record_bytes<TP_FORCE, PD_BB, true, true, true>();
  if (pass == 0) return xf ? fp64_record_bytes<TP_THETA, true>(kind) : fp64_record_bytes<TP_THETA, false>(kind);
  if (pass == 1) return xf ? fp64_record_bytes<TP_TD, true>(kind) : fp64_record_bytes<TP_TD, false>(kind);
  return xf ? fp64_record_bytes<TP_FORCE, true>(kind) : fp64_record_bytes<TP_FORCE, false>(kind);
}

int pd_tile_class(int pass, int kind, int prec, int xf, int hmax) {
  const int rec = pd_tile_record_bytes(pass, kind, prec, xf);
  const int res = (prec == PD_PREC_FP32C && kind != PD_EP && pass != TP_TD) ? 0 : kXTabLds;
  const bool lat = (xf == 2 || xf == 3) && pass == TP_FORCE && kind == PD_BB && prec != PD_PREC_FP32C;
  if (lat && hmax <= (xf == 3 ? halo_stride8(rec, res) : halo_stride7(rec, res))) return 0;
  for (int c = 0; c < 4; ++c)
    if (hmax <= halo_stride(c, rec, res)) return c + 1;
  return 5;
}

void pd_tile_allow_lds(size_t bytes) {
  (void)hipFuncSetAttribute((const void*)k_halo_count, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
  (void)hipFuncSetAttribute((const void*)k_halo_fill, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
  allow_lds_lg<1>(bytes);
  allow_lds_lg<2>(bytes);
  allow_lds_lg<3>(bytes);
}

#define PD_LG_XF(F, ...)                                      \
  do {                                                        \
    if (xf) {                                                 \
      if (lgG == 1) F<1, true>(__VA_ARGS__);                  \
      else if (lgG == 2) F<2, true>(__VA_ARGS__);             \
      else F<3, true>(__VA_ARGS__);                           \
    } else {                                                  \
      if (lgG == 1) F<1, false>(__VA_ARGS__);                 \
      else if (lgG == 2) F<2, false>(__VA_ARGS__);            \
      else F<3, false>(__VA_ARGS__);                          \
    }                                                         \
  } while (0)

void pd_launch_tile_theta(hipStream_t s, int lgG, int xf, const Tile* tiles, int n, EnvDev* envs, int hcap) {
  if (n <= 0) return;
  PD_LG_XF(launch_theta_lg, s, tiles, n, envs, hcap);
}

void pd_launch_tile_td(hipStream_t s, int lgG, int xf, const Tile* tiles, int n, EnvDev* envs, int hcap) {
  if (n <= 0) return;
  PD_LG_XF(launch_td_lg, s, tiles, n, envs, hcap);
}

template <int LG, bool XF>
void launch_force_any(int kind, hipStream_t s, const Tile* tiles, int n, EnvDev* envs, int flags, int hcap) {
  if (kind == PD_BB) launch_force_lg<PD_BB, LG, XF>(s, tiles, n, envs, flags, hcap);
  else if (kind == PD_OSB) launch_force_lg<PD_OSB, LG, XF>(s, tiles, n, envs, flags, hcap);
  else launch_force_lg<PD_EP, LG, XF>(s, tiles, n, envs, flags, hcap);
}

void pd_launch_tile_force(hipStream_t s, int kind, int lgG, int xf, const Tile* tiles, int n, EnvDev* envs,
                          int flags, int hcap) {
  if (n <= 0) return;
  if (xf == 2 && kind == PD_BB) {
    if (lgG == 1) launch_force_lat_lg<1>(s, tiles, n, envs, flags, hcap);
    else if (lgG == 2) launch_force_lat_lg<2>(s, tiles, n, envs, flags, hcap);
    else launch_force_lat_lg<3>(s, tiles, n, envs, flags, hcap);
    return;
  }
  if (xf == 3 && kind == PD_BB) {
    if (lgG == 1) launch_force_onev_lg<1>(s, tiles, n, envs, flags, hcap);
    else if (lgG == 2) launch_force_onev_lg<2>(s, tiles, n, envs, flags, hcap);
    else launch_force_onev_lg<3>(s, tiles, n, envs, flags, hcap);
    return;
  }
  PD_LG_XF(launch_force_any, kind, s, tiles, n, envs, flags, hcap);
}
#undef PD_LG_XF

void pd_launch_pk_sync(hipStream_t s, const Chunk* chunks, int n, EnvDev* envs) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_pk_sync, dim3((n + PD_WAVES_PER_BLOCK - 1) / PD_WAVES_PER_BLOCK), dim3(256), 0, s, chunks,
                     n, envs);
}

void pd_launch_tile_theta32(hipStream_t s, int lgG, const Tile* tiles, int n, EnvDev* envs, int hcap) {
  if (n <= 0) return;
  if (lgG == 1) launch_theta32_lg<1>(s, tiles, n, envs, hcap);
  else if (lgG == 2) launch_theta32_lg<2>(s, tiles, n, envs, hcap);
  else launch_theta32_lg<3>(s, tiles, n, envs, hcap);
}

void pd_launch_tile_force32(hipStream_t s, int kind, int lgG, const Tile* tiles, int n, EnvDev* envs, int flags,
                            int hcap) {
  if (n <= 0) return;
  if (kind == PD_BB) {
    if (lgG == 1) launch_force32_lg<PD_BB, 1>(s, tiles, n, envs, flags, hcap);
    else if (lgG == 2) launch_force32_lg<PD_BB, 2>(s, tiles, n, envs, flags, hcap);
    else launch_force32_lg<PD_BB, 3>(s, tiles, n, envs, flags, hcap);
  } else {
    if (lgG == 1) launch_force32_lg<PD_OSB, 1>(s, tiles, n, envs, flags, hcap);
    else if (lgG == 2) launch_force32_lg<PD_OSB, 2>(s, tiles, n, envs, flags, hcap);
    else launch_force32_lg<PD_OSB, 3>(s, tiles, n, envs, flags, hcap);
  }
}

int pd_tile_build_lds(int hc) { return (int)sizeof(int32_t) * (hc + 257); }

void pd_launch_halo_count(hipStream_t s, const Tile* tiles, int n, const int32_t* col, int32_t* counts, int hc) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_halo_count, dim3(n), dim3(256), (size_t)pd_tile_build_lds(hc), s, tiles, n, col, counts, hc);
}

void pd_launch_halo_fill(hipStream_t s, const Tile* tiles, int n, const int32_t* col, int32_t* halo, uint64_t* pk,
                         int hc) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_halo_fill, dim3(n), dim3(256), (size_t)pd_tile_build_lds(hc), s, tiles, n, col, halo, pk, hc);
}
