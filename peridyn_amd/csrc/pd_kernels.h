// Host-side launch wrappers of the device kernels (all asynchronous on `s`).
#pragma once
#include "pd_internal.h"

// k_force flags
#define PD_F_KICK 1   // second half kick + velocity BCs of the current step
#define PD_F_DRIFT 2  // first half kick + drift + position BCs of the next step
// k_book flags
#define PD_B_FLIP_Y 1
#define PD_B_FLIP_EDP 2
#define PD_B_STEP 4

void pd_launch_theta(hipStream_t s, int prec, const Chunk* chunks, int n, EnvDev* envs);
void pd_launch_td(hipStream_t s, int prec, const Chunk* chunks, int n, EnvDev* envs);
// gpow: some contact model of the envs has a non-integer exponent
void pd_launch_force(hipStream_t s, int kind, int prec, const Chunk* chunks, int n, EnvDev* envs, int flags,
                     bool gpow);
void pd_launch_prologue(hipStream_t s, const Chunk* chunks, int n, EnvDev* envs);
void pd_launch_book(hipStream_t s, EnvDev* envs, const int32_t* slots, int n, int flags);
// deformed latch over n (env slot, block) items: the slices of item i (a block
// of n_b particles has pd_latch_slices(n_b)) are the units (i, 0 .. ns - 1),
// listed from ustart[i] on in the nunits units; part holds one double4 per unit
int pd_latch_slices(int64_t n);
void pd_launch_latch(hipStream_t s, EnvDev* envs, const int2* items, int n, const int2* units, int nunits,
                     const int32_t* ustart, double4* part);

// family build (pd_family.hip); all arrays device, env-local particle ids
struct FamilyGrid {
  double lo[3];
  double w;          // cell width (>= horizon)
  int32_t nc[3];     // cells per dimension
};
void pd_launch_cell_of(hipStream_t s, const double4* x, int first, int n, FamilyGrid g,
                       int32_t* cell_of, int32_t* cell_count);
void pd_launch_cell_fill(hipStream_t s, int n, const int32_t* cell_of, const int32_t* cell_start,
                         int32_t* cell_cursor, int32_t* cell_items, int first);
void pd_launch_degree(hipStream_t s, const double4* x, int first, int n, FamilyGrid g, double horizon,
                      const int32_t* cell_of, const int32_t* cell_start, const int32_t* cell_count,
                      const int32_t* cell_items, int32_t* deg);
void pd_launch_fill(hipStream_t s, const double4* x, int first, int n, FamilyGrid g, double horizon,
                    const int32_t* cell_of, const int32_t* cell_start, const int32_t* cell_count,
                    const int32_t* cell_items, const uint32_t* pslot0, int lgG, const int32_t* ref, int32_t* col,
                    uint8_t* state, int maxdeg);
void pd_launch_wvol(hipStream_t s, const double4* x, int first, int n, const uint32_t* pslot0, int lgG,
                    const int32_t* deg, const int32_t* col, double* m);
void pd_launch_rev(hipStream_t s, int first, int n, const uint32_t* pslot0, int lgG, const int32_t* ref,
                   const int32_t* deg, const int32_t* col, uint32_t* rev);
void pd_launch_aux_init(hipStream_t s, const double* m, double4* aux, int n);
// quasi-static minimiser: BCs once, then per iteration the drag update of the
// free particles (maxes: 2 x u64, max |f| and max |dy| as non-negative bits)
void pd_launch_qs_bc(hipStream_t s, const Chunk* chunks, int n, EnvDev* envs);
void pd_launch_qs_update(hipStream_t s, const Chunk* chunks, int n, EnvDev* envs, double4* opt, double dt,
                         double lambda, double a, unsigned long long* maxes);
// printdata columns: mode 0 momentum into rows[8 e + 1..3], mode 1 force and
// damage sums into rows[8 e + 4..7] for the nenv envs whose chunks (in `chunks`,
// n in all) are the ranges (first, count); part: n double4 of scratch
// printdata columns of one env into row[0..7] (device), partial sums in part
// (one double4 per chunk): the kernels of pd_env_observables_noalloc
// one contact model's terms added into acc (n = env particles, internal order)
void pd_launch_contact_model(hipStream_t s, const EnvDev* envs, int slot, int cidx, bool gpow, double4* acc, int n);
// tiled envs with contacts (slots[0 .. n)): number the particles with contact
// terms (cflag, cpart, ccount) after a list rebuild; starts[k] = first of the
// nblocks workgroups of env k (prefix over ceil(N / 256))
void pd_launch_cpart_build(hipStream_t s, EnvDev* envs, const int32_t* slots, const int32_t* starts, int n,
                           int nblocks);
// contact terms + Verlet update of the particles the tiled force kernels left
// unfinished (after all force passes of the step), envs slots[0 .. n), nblocks
// workgroups of 4 wavefronts striding over all their particles: one lane per
// particle when they number lane_min or more, else one wavefront per particle
void pd_launch_contact_finish(hipStream_t s, EnvDev* envs, const int32_t* slots, int n, int nblocks, int lane_min,
                              int flags);
void pd_launch_observables_env(hipStream_t s, const Chunk* chunks, int n, EnvDev* envs, int slot, double4* part,
                               double* row);
void pd_launch_observables(hipStream_t s, int mode, const Chunk* chunks, int n, EnvDev* envs, double4* part,
                           const int2* ranges, int nenv, double* rows);
// device-wide exclusive scan of n int32 (scratch: pd_scan_scratch_ints(n) ints)
int pd_scan_scratch_ints(int n);
void pd_scan_exclusive(hipStream_t s, const int32_t* in, int32_t* out, int n, int32_t* scratch);

// contacts (pd_contact.hip)
void pd_launch_contact_update(hipStream_t s, EnvDev* envs, int slot, const EnvDev& host_env,
                              const ContactDev& c, int cidx);
// self contact (ContactModel11): own lists from the bond states, transposed lists
void pd_launch_self_contact_update(hipStream_t s, EnvDev* envs, int slot, const ContactDev& c, int cidx);
// the same update for many (env slot, contact index) items at once, one launch
// per kernel of the sequence (pair contacts; self contacts use the per-env
// form): blocks of item k in class cls = pd_contact_batch_blocks(cls, ...),
// starts[cls * (nitems + 1) + k] their prefix sums, grid[cls] the totals
enum { PD_CB_BOX = 0, PD_CB_ZERO, PD_CB_N2T, PD_CB_CTILE, PD_CB_N2TILE, PD_CB_N1W, PD_CB_N1T, PD_CB_SRCH, PD_CB_SORT,
       PD_CB_NCLS };
int pd_contact_batch_blocks(int cls, int n1, int n2, int ncell_cap);
void pd_launch_contact_update_batch(hipStream_t s, const EnvDev* envs, const int2* items, const int32_t* starts,
                                    int nitems, const int* grid);

// FP32-compensated form (pd_step32.hip)
void pd_launch_theta32(hipStream_t s, const Chunk* chunks, int n, EnvDev* envs);
void pd_launch_td32(hipStream_t s, const Chunk* chunks, int n, EnvDev* envs);
void pd_launch_force32(hipStream_t s, int kind, const Chunk* chunks, int n, EnvDev* envs, int flags, bool gpow);

// tiled passes (pd_tile.hip): LDS bytes per halo record for pass (0 theta,
// 1 td, 2 force), kind and form; xf = 1: every x and volume of the launched
// envs is exactly an FP32 value (compact records), 2: besides, every tiled
// bond-based block is an integer lattice (compact lattice records of the
// bond-based force pass), 3: besides, every such block has one volume (the
// lattice records without it); hcap = largest halo of the launched tiles
int pd_tile_record_bytes(int pass, int kind, int prec, int xf);
// LDS class of a pass's tiles whose largest halo is hmax: 1..4 for the
// compiled strides that let 6, 5, 4, 3 workgroups share a CU (for the pass's
// record size), 5 for the run-time capacity, 0 for the 7-workgroup stride of
// the compact lattice records (bond-based force, xf = 2) or the 8-workgroup
// stride of the one-volume ones (xf = 3).  Envs are launched
// per class, so that one env with a large halo does not lower the occupancy
// of the others.
constexpr int PD_TILE_CLASSES = 6;
int pd_tile_class(int pass, int kind, int prec, int xf, int hmax);
void pd_tile_allow_lds(size_t bytes);
void pd_launch_tile_theta(hipStream_t s, int lgG, int xf, const Tile* tiles, int n, EnvDev* envs, int hcap);
void pd_launch_tile_td(hipStream_t s, int lgG, int xf, const Tile* tiles, int n, EnvDev* envs, int hcap);
void pd_launch_tile_force(hipStream_t s, int kind, int lgG, int xf, const Tile* tiles, int n, EnvDev* envs,
                          int flags, int hcap);
// pk intact bits := state for the listed chunks of tiled envs
void pd_launch_pk_sync(hipStream_t s, const Chunk* chunks, int n, EnvDev* envs);
// FP32-compensated tiled passes (BB / OSB blocks of PD_PREC_FP32C envs)
void pd_launch_tile_theta32(hipStream_t s, int lgG, const Tile* tiles, int n, EnvDev* envs, int hcap);
void pd_launch_tile_force32(hipStream_t s, int kind, int lgG, const Tile* tiles, int n, EnvDev* envs, int flags,
                            int hcap);
// device build of the tile halos: LDS bytes for a hash set of hc (power of two,
// >= twice a tile's bond slots) entries; counts[t] = halo size of tile t; fill
// writes the sorted halo[hoff..] and pk
int pd_tile_build_lds(int hc);
void pd_launch_halo_count(hipStream_t s, const Tile* tiles, int n, const int32_t* col, int32_t* counts, int hc);
void pd_launch_halo_fill(hipStream_t s, const Tile* tiles, int n, const int32_t* col, int32_t* halo, uint64_t* pk,
                         int hc);
// deformed latch of n (env slot, block) items into out[n]
void pd_launch_latch_flags(hipStream_t s, EnvDev* envs, const int2* items, int n, int32_t* out);
// n device-to-device copies of 4-byte words in one launch
struct PdCopy {
  const uint32_t* src;
  uint32_t* dst;
  uint32_t words, pad;
};
void pd_launch_copy_words(hipStream_t s, const PdCopy* d, int n);
