// Internal device/host structures of libperidyn_hip (gfx950, wave64).
//
// HBM layout (per env, all arrays device-resident, env-local particle ids):
//   x      double4[N]  reference position xyz + particle volume in .w
//   y[2]   double4[N]  deformed positions, double-buffered across the fused
//                      force->Verlet kernel (read y[cur], write y[cur^1])
//   v, f   double4[N]  velocity, force density (acceleration) — .w unused
//   bonds  SELL-64 with G lanes per particle (G = 2^lgG, C = 64/G particles
//          per wavefront chunk of consecutive ids of one block).  Lane
//          L = g*C + q of a chunk serves particle first+q and its bonds
//          k = g, g+G, g+2G, ...; bond k of that particle sits at
//              slot0 + 64*(k/G) + (k%G)*C + q,
//          so at every step of the bond loop the 64 lanes of a wavefront read 64
//          consecutive slots.  col int32 (neighbour id; each particle's row is
//          sorted ascending in k), state uint8 (1 intact).  G = 1 is one lane per
//          particle (used by the reference-order form); G = 64 is one wavefront
//          per particle.
//   aux    double4[N] per-particle gather record of the fast forms:
//          (theta, 1/m, td_norm2, C = sqrt(2 psi / td_norm2))
//   EP     rev uint32 slot of the reverse bond, edp[2] double per slot.
//   tiles  (FP64 form, optional) a tile is up to 4 chunks of one brick of the
//          block's internal (brick) order, run by one workgroup that first
//          stages its halo -- the sorted, unique neighbour ids of its particles
//          (halo[hoff .. hoff+hn)) -- into LDS.  pk holds, per chunk, per group
//          of 4 bond steps t and lane L, one uint64 at pk0 + 64 t + L: bond step
//          4t+u in bits 16u..16u+15 as (intact << 15) | halo-local index.  The
//          intact bits mirror `state` (the master copy).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/peridyn.h"

#define PD_WAVE 64
#define PD_WAVES_PER_BLOCK 4
#define PD_MAX_TYPES 8
#define PD_MAX_ARRAY_BYTES (1ull << 32)  // per device array of an env (32-bit offsets)
#define PD_MAX_CONTACTS 8

struct BlockDev {
  int32_t kind, ntype;
  int32_t first, n;          // env-local particle range
  int32_t chunk0, nchunks;   // chunk range (env-local chunk ids)
  int32_t deformed;          // latch (general_material.jl:243-255)
  // every reference position of the block is an integer (|x| < 2^22) and
  // horizon^2 <= 32: each bond's |X|^2 is an integer 1..31, and the tiled
  // kernels take |X| and 1/|X| from a table (pd_xtab_size entries)
  int32_t xint;
  double c11;                // BB bond stiffness c[1,1]
  double horizon;
  double density[PD_MAX_TYPES];
  double s0[PD_MAX_TYPES * PD_MAX_TYPES];   // [t1 + nt*t2] (symmetric)
  double K[PD_MAX_TYPES * PD_MAX_TYPES];
  double G[PD_MAX_TYPES * PD_MAX_TYPES];
  double psi[PD_MAX_TYPES * PD_MAX_TYPES];
  // FP32C form: origin of the FP32 reference offsets x32 (the block's minimum
  // corner), and the smallest FP32 value >= s0[0]: for an FP32 stretch s,
  // s < s0f_up exactly when (double)s < s0[0]
  double corner[3];
  float s0f_up;
  // every particle of the block has the volume vol0 (the tiled bond-based
  // kernels of lattice records then take it from here, not per record)
  int32_t vol1;
  double vol0;
};

struct BcDev {
  int32_t kind;
  int32_t pad;
  double rate[3];
  double freq;               // integer-valued or +Inf
  int64_t applyafter;
};

struct ContactDev {
  int32_t n1, n2;            // members of pair[1] / pair[2] (self contact: n2 = n1)
  int32_t max_neighs;
  int32_t expo_is_int;
  int32_t self_contact;      // ContactModel11 of one block (idx2 = idx1)
  int32_t pad_;
  double expo, K, equi, distance;
  const int32_t* idx1;       // member -> env particle id
  const int32_t* idx2;
  const int32_t* inv1;       // env particle id -> member or -1
  const int32_t* inv2;
  int32_t* own_cnt;          // n1
  int32_t* own;              // n1 * max_neighs, block-2 member ids, descending
  int32_t* rev_cnt;          // n2
  int32_t* rev_off;          // n2 (exclusive offsets)
  int32_t* rev;              // block-1 member ids, ascending per block-2 member
  int32_t* rev_cursor;       // n2 scratch
  int32_t* cell_start;       // cell list over in-box block-2 members
  int32_t* cell_count;
  int32_t* cell_items;       // n2
  int32_t* cell_of;          // n2 (cell id or -1 outside the box)
  double* box;               // [0..2] min, [3..5] max, [6] overlap flag (1/0), [7..] grid
  int32_t* wlist;            // max(n1, n2): members with work in a rebuild (search, then sort)
  int32_t* wcount;           // 2: lengths of the two lists
  int32_t ncell_cap;
  int32_t rev_cap;
};

struct EnvDev {
  int32_t N, nblocks, ncontacts, nbc;
  int32_t precision, cur, edp_cur, has_ep;
  int32_t lgG, tiled;        // log2(lanes per particle) of the bond layout; tiles built
  int32_t xf32, pad1;        // every x and volume is exactly an FP32 value (compact tile records)
  double dt;
  int64_t time_step;         // completed steps (env.time_step)
  double4* x;
  double4* y[2];
  double4* v;
  double4* f;
  // FP32C form (BB / OSB blocks): x32 = float(x - block minimum corner) with the
  // volume in .w; aux32 = (theta, 1/m, 0, 0)
  float4* x32;
  float4* aux32;
  int32_t* pblock;           // block of each particle
  uint8_t* ptype;            // 0-based type-local index
  double* inv_den;           // 1 / env density (mass/volume), contacts.jl:231,235
  int32_t* bcid;             // effective BC per particle (last in list order) or -1
  uint8_t* qheld;            // minimize!'s held particles when some BC has onlyatstart, else null (bcid >= 0)
  int32_t* deg;              // degree per particle
  uint32_t* pslot0;          // slot of (p, k = 0); slot(p, k) = pslot0[p] + 64 k
  double* m;                 // weighted volume
  double4* aux;              // (theta, 1/m, td_norm2, C) for the fast forms
  double* theta;
  double* td;                // EP td_norm2
  int32_t* col;
  uint8_t* state;
  uint32_t* rev;             // EP only
  double* edp[2];            // EP only
  int32_t* halo;             // tiled: halo lists of all tiles
  uint64_t* pk;              // tiled: packed local indices + intact bits
  BlockDev* blocks;
  BcDev* bcs;
  ContactDev* contacts;
  // tiled envs with contacts: the particles with contact terms under the
  // current lists (cflag), compacted into cpart[0 .. *ccount) in no particular
  // order; their tiled force kernels store the bond sum in f and leave the
  // contact terms and the Verlet update to k_contact_finish
  uint8_t* cflag;
  int32_t* cpart;
  int32_t* ccount;
  // envs with one pair contact model: per cpart entry (p, i1, n1, i2, n2,
  // rev_off, 0, 0): the member indices, list lengths and transposed-list
  // offset that the finish pass would otherwise look up one after another
  int32_t* crec;
};

struct Chunk {
  int32_t env, block;
  int32_t first, count;      // env-local particle range of the chunk
  uint32_t slot0;
  int32_t width;             // bond steps per lane
  uint32_t pk0;              // tiled: first pk word
  int32_t pad;
};

#define PD_TILE_CHUNKS 4
struct Tile {
  int32_t env, block;
  int32_t nch, hn;           // chunks (<= PD_TILE_CHUNKS), halo records
  uint32_t hoff;             // first halo entry
  int32_t pad;
  Chunk ch[PD_TILE_CHUNKS];
};

// ---- host-side helpers ----------------------------------------------------
#define PD_HIP(call)                                                        \
  do {                                                                      \
    hipError_t e_ = (call);                                                 \
    if (e_ != hipSuccess) {                                                 \
      throw PdError(PD_E_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    }                                                                       \
  } while (0)

#include <string>
struct PdError {
  int code;
  std::string msg;
  PdError(int c, std::string m) : code(c), msg(std::move(m)) {}
};
