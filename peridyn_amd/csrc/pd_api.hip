// C ABI of libperidyn_hip.so (declared in include/peridyn.h): contexts, env
// upload + GPU family build, the resident run loop and state transfer.
//
// Each env owns its device arrays (layout in pd_internal.h).  All envs of a
// context share one device table EnvDev[slot]; kernels index it through the
// chunk descriptors, so any set of envs is stepped together by one launch per
// kernel class (the batched / served form) and a single env is a set of one.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <map>
#include <thread>
#include <string>
#include <vector>

#include "pd_device.h"
#include "pd_internal.h"
#include "pd_kernels.h"

// Scratch of pd_env_observables_noalloc: one double4 of partial sums per chunk
// of the env (8 MiB, held by every GPU the library runs on until the process
// ends).  Host code never passes the array's name to a kernel: each context
// resolves this GPU's copy with hipGetSymbolAddress at creation and uses that
// address (Ctx::scratch).
constexpr int kObsScratchChunks = 1 << 18;
static __device__ __align__(256) double4 scratch[kObsScratchChunks];

namespace {

// Device memory of one context's envs.  Released blocks are kept by size class
// and handed to later allocations of the same class: hipMalloc maps pages and
// hipFree waits for the whole device, both slow next to admitting a small env.
// All work of a context is ordered on its one stream, so a block reused by a
// later allocation is only touched by work enqueued after its last use.
struct DevMem {
  static constexpr size_t kMaxCached = 8ull << 30;   // bytes kept at most
  static constexpr size_t kMaxBlock = 1ull << 30;    // larger blocks are freed at once
  std::multimap<size_t, void*> spare;                // size class -> block
  size_t cached = 0;
  // poison: every block handed out is first filled with 0xFF bytes on `stream`
  // (NaN doubles, -1 integers), so that a kernel reading memory it has not
  // written gives results that differ from a run without it
  hipStream_t stream = nullptr;
  bool poison = false;
  // powers of two up to 1 MiB, then steps of 1/8 of the power of two below
  static size_t size_class(size_t bytes) {
    size_t p = 256;
    while (p < bytes && p < (1u << 20)) p <<= 1;
    if (p >= bytes) return p;
    size_t q = 1u << 20;
    while ((q << 1) <= bytes) q <<= 1;
    const size_t step = q >> 3;
    return (bytes + step - 1) / step * step;
  }
  void* get(size_t cls) {
    void* p = take(cls);
    if (poison) PD_HIP(hipMemsetAsync(p, 0xff, cls, stream));
    return p;
  }
  void* take(size_t cls) {
    auto it = spare.find(cls);
    if (it != spare.end()) {
      void* p = it->second;
      spare.erase(it);
      cached -= cls;
      return p;
    }
    void* p = nullptr;
    hipError_t r = hipMalloc(&p, cls);
    if (r == hipErrorOutOfMemory && cached > 0) {
      (void)hipGetLastError();
      trim(0);
      r = hipMalloc(&p, cls);
    }
    if (r != hipSuccess) {
      (void)hipGetLastError();
      throw PdError(r == hipErrorOutOfMemory ? PD_E_NOMEM : PD_E_HIP, std::string("hipMalloc: ") + hipGetErrorString(r));
    }
    return p;
  }
  void put(void* p, size_t cls) {
    if (!p) return;
    if (cls > kMaxBlock) {
      (void)hipFree(p);
      return;
    }
    spare.emplace(cls, p);
    cached += cls;
    if (cached > kMaxCached) trim(kMaxCached / 2);
  }
  // frees spare blocks, largest first, until at most keep bytes are cached
  void trim(size_t keep) {
    while (cached > keep && !spare.empty()) {
      auto it = std::prev(spare.end());
      (void)hipFree(it->second);
      cached -= it->first;
      spare.erase(it);
    }
  }
};

// Pinned host staging of one context, reused by every commit: fresh pageable
// vectors fault their pages in on first touch, and copies from pageable memory
// go through the runtime's own staging buffer.  Commits needing more than
// kMaxPinned stage in pageable memory instead (no gigabytes held pinned).
struct HostStage {
  static constexpr size_t kMaxPinned = 1ull << 30;
  char* base = nullptr;      // the buffer in use (pinned or big.data())
  char* pinned = nullptr;
  size_t pinned_cap = 0;
  std::vector<char> big;     // pageable staging of the largest commits
  size_t cap = 0, used = 0;
  template <typename T>
  struct Span {
    T* p;
    T& operator[](size_t i) const { return p[i]; }
    T* data() const { return p; }
  };
  // starts a commit's staging (every commit ends with the stream idle, so no
  // copy from an earlier one is in flight); grows by half again when short
  void begin(size_t bytes, hipStream_t s) {
    used = 0;
    if (bytes > kMaxPinned) {
      PD_HIP(hipStreamSynchronize(s));
      std::vector<char>().swap(big);
      big.resize(bytes);
      base = big.data();
      cap = bytes;
      return;
    }
    std::vector<char>().swap(big);
    if (bytes > pinned_cap) {
      PD_HIP(hipStreamSynchronize(s));
      if (pinned) (void)hipHostFree(pinned);
      pinned = nullptr;
      pinned_cap = 0;
      const size_t want = std::min(std::max<size_t>(bytes + bytes / 2, (size_t)1 << 20), kMaxPinned);
      PD_HIP(hipHostMalloc((void**)&pinned, want, hipHostMallocDefault));
      pinned_cap = want;
    }
    base = pinned;
    cap = pinned_cap;
  }
  template <typename T>
  Span<T> take(size_t n) {
    const size_t off = (used + 255) & ~(size_t)255;
    const size_t end = off + std::max<size_t>(n, 1) * sizeof(T);
    if (end > cap) throw PdError(PD_E_STATE, "host staging overflow");
    used = end;
    return Span<T>{reinterpret_cast<T*>(base + off)};
  }
};
// staged bytes per particle in a commit (x, y, v, block, BC, type, 1/density,
// degree, row start, reference index, FP32C x32, aux32 and m, minimize!'s held
// flag), and the padding
constexpr size_t kStagePerParticle = 3 * 32 + 4 + 4 + 1 + 8 + 4 + 4 + 4 + 16 + 16 + 8 + 1;
constexpr size_t kStagePad = 16 * 256;

struct HostBlock {
  pd_block_desc d{};
  std::vector<double> x, y, v, vol;
  std::vector<int64_t> type;
  std::vector<double> density, s0, c, K, G, psi;
  int32_t first = 0, chunk0 = 0, nchunks = 0, maxdeg = 0;
  bool xint = false;  // integer lattice with horizon^2 <= 32 (BlockDev::xint)
  bool vol1 = false;  // one volume (BlockDev::vol1)
};

struct HostContact {
  pd_contact_desc d{};
  std::vector<int64_t> t1, t2;
};

struct HostBc {
  pd_bc_desc d{};
  std::vector<uint8_t> mask;
};

struct Env {
  int64_t id = 0;
  int32_t slot = -1;
  double dt = 0.0;
  int32_t precision = PD_PREC_FP64;
  int32_t lgG = -1;  // log2 lanes per particle; -1 = choose at commit
  std::vector<HostBlock> blocks;
  std::vector<HostContact> contacts;
  std::vector<HostBc> bcs;
  bool committed = false;
  bool maybe_undeformed = true;
  std::vector<uint8_t> block_latched;  // blocks read back as latched (latches never reset)
  // Blocks whose every particle one BC holds (not onlyatstart): the BC moves
  // them by one displacement per step, so y - y[1] + x[1] - x stays at
  // rounding level (~1e-16 |y|), far inside the latch's rtol = sqrt(eps); once
  // such a block has been read back unlatched after y was last set from the
  // host (block_rigid_ok), later latch tests would all fail too and are skipped.
  std::vector<uint8_t> block_rigid, block_rigid_ok;
  // the contact lists were built from the current positions (nothing has moved
  // them since), so an update_neighs! now would rebuild the same sorted lists
  bool lists_current = false;
  bool has_ep = false;
  int64_t N = 0, nslots = 0, nbonds = 0, npk = 0;
  int64_t time_step = 0;
  int32_t cur = 0, edp_cur = 0;
  std::vector<int32_t> perm, iperm;  // internal id -> reference index and back (empty: identity)
  // tiled FP64 passes: brick order requested at commit (want_tiles), tiles
  // built (tiled: the halos fit the LDS budget); brick of each internal particle
  bool want_tiles = false, tiled = false, xf32 = false;
  std::vector<int32_t> brick;
  // per brick: its (x, y, z) position in its block's brick grid
  std::vector<std::array<int32_t, 3>> brick_pos;
  std::vector<Tile> tiles;
  Tile* d_tiles = nullptr;                       // device copy (run sets copy from it)
  std::vector<std::array<int, 3>> tile_blocks;   // (block, first tile, count), tiles in block order
  int32_t hmax = 0;  // largest tile halo
  std::vector<Chunk> chunks;
  Chunk* d_chunks = nullptr;
  int32_t* d_self = nullptr;  // this env's slot, for single-env bookkeeping launches
  // tiled envs with contacts: {slot, 0}, the one-env list of pd_launch_cpart_build
  int32_t* d_cself = nullptr;
  int32_t cself_host[2] = {0, 0};
  EnvDev dev{};
  std::vector<ContactDev> cdev;
  DevMem* mem = nullptr;  // the context's device memory
  std::vector<std::pair<void*, size_t>> allocs;  // block, size class
  size_t bytes = 0;
};

struct Ctx {
  int device = 0;
  // the stream every call enqueues on (`stream`): the run stream, except
  // while a run started by pd_run_begin is in flight, when the calls allowed
  // then (env construction, commit, destroy) use the admission stream, so that
  // their waits do not wait for the run
  hipStream_t stream = nullptr;
  hipStream_t run_stream = nullptr;
  hipStream_t adm_stream = nullptr;  // created with the first pd_run_begin
  // the bond-based force passes of a step run here, beside the dilatation
  // passes they do not depend on (fork after the latch, join before the
  // contact finish); created with the first run that forks
  hipStream_t fork_stream = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  struct Inflight {
    bool on = false;
    std::vector<int64_t> ids;
    int64_t nsteps = 0, nuf = 1;
    int nl = 0;
  } inflight;
  double4* scratch = nullptr;  // this GPU's copy of the file-scope scratch array
  DevMem mem;  // env arrays and build temporaries
  HostStage stage;  // commit staging
  std::string err;
  std::map<int64_t, Env*> envs;
  int64_t next_id = 1;
  EnvDev* d_envs = nullptr;
  int32_t slot_cap = 0;
  std::vector<int32_t> free_slots;
  // cached run set
  std::vector<int64_t> set_ids;
  bool set_valid = false;
  bool set_full = false;  // the all-chunk lists per pass (th, td, fk) were built too
  // stepping lists: chunks of untiled envs per pass, tiles of tiled envs per
  // pass and lanes per particle (lgG 1..3)
  int n_theta_u = 0, n_td_u = 0, n_force_u[3] = {0, 0, 0};
  int off_theta_u = 0, off_td_u = 0, off_force_u[3] = {0, 0, 0};
  Tile* d_tset = nullptr;
  size_t tset_cap = 0;
  // tile lists [form class: 0 FP64 kernels, 1 FP32C kernels][lgG][LDS class]
  struct TileList {
    int off = 0, n = 0;
    int hcap = 1;  // largest halo of the listed envs
  };
  TileList tl_theta[2][4][PD_TILE_CLASSES], tl_td[4][PD_TILE_CLASSES], tl_force[3][2][4][PD_TILE_CLASSES];
  bool set_tiled = false;
  int set_xf = 0;  // every tiled env of the set has FP32-exact x and volumes
  // pair contacts of the set, rebuilt together (pd_launch_contact_update_batch)
  int2* d_citems = nullptr;
  size_t citems_cap = 0;
  int32_t* d_cstarts = nullptr;
  size_t cstarts_cap = 0;
  int n_citems = 0;
  int cgrid[PD_CB_NCLS] = {};
  // tiled envs of the set with contacts: slots and first cpart-build workgroups
  int32_t* d_cfin = nullptr;
  size_t cfin_cap = 0;
  int32_t* d_cfin_starts = nullptr;
  size_t cfin_starts_cap = 0;
  int n_cfin = 0, cfin_blocks = 0;
  int cfin_grid = 0;  // workgroups of the contact-finish pass
  PdCopy* d_copies = nullptr;  // build_set's copy descriptors
  size_t copies_cap = 0;
  Chunk* d_set = nullptr;
  size_t set_cap = 0;
  int n_all = 0, n_theta = 0, n_td = 0, n_force[3] = {0, 0, 0};
  int off_theta = 0, off_td = 0, off_force[3] = {0, 0, 0};
  int32_t* d_slots = nullptr;
  int n_slots = 0;
  size_t slots_cap = 0;
  bool set_has_ep = false, set_has_theta = false;
  bool set_gpow = false;  // some contact model of the set has a non-integer exponent
  // observables: chunk range (first, count) of each env of the set in its
  // chunk list, per-chunk partial sums, one 8-column row per env and logged step
  int2* d_obs_ranges = nullptr;
  size_t obs_ranges_cap = 0;
  double4* d_obs_part = nullptr;
  size_t obs_part_cap = 0;
  double* d_obs_rows = nullptr;
  size_t obs_rows_cap = 0;
  int2* d_latch = nullptr;  // (env slot, block) items of the latch test
  size_t latch_cap = 0;
  int2* d_latch_units = nullptr;  // (item, slice): one workgroup each
  size_t latch_units_cap = 0;
  int32_t* d_latch_ustart = nullptr;  // first unit of each item
  size_t latch_ustart_cap = 0;
  int latch_nunits = 0;
  std::vector<std::pair<int64_t, int>> latch_items;  // (env id, block) of the items
  double4* d_latch_part = nullptr;  // per-unit partial sums of the latch test
  size_t latch_part_cap = 0;
  int32_t* d_latch_flags = nullptr;  // per latch item, read back after a run
  size_t latch_flags_cap = 0;
  // per-kernel-class HIP-event timing on the context stream (pd_ctx_profile)
  bool prof = false;
  struct Rec { int cls; hipEvent_t a, b; };
  std::vector<Rec> pending;
  std::vector<hipEvent_t> pool;
  double prof_ms[PD_PROF_NCLASS] = {};
  int64_t prof_n[PD_PROF_NCLASS] = {};
};

hipEvent_t get_event(Ctx* c) {
  if (!c->pool.empty()) {
    hipEvent_t e = c->pool.back();
    c->pool.pop_back();
    return e;
  }
  hipEvent_t e;
  PD_HIP(hipEventCreate(&e));
  return e;
}

// Runs `f` (kernel launches on c->stream) bracketed by events when profiling.
template <typename F>
void timed(Ctx* c, int cls, F&& f) {
  if (!c->prof) {
    f();
    return;
  }
  hipEvent_t a = get_event(c), b = get_event(c);
  PD_HIP(hipEventRecord(a, c->stream));
  f();
  PD_HIP(hipEventRecord(b, c->stream));
  c->pending.push_back({cls, a, b});
}

void drain_profile(Ctx* c) {
  PD_HIP(hipStreamSynchronize(c->stream));
  for (auto& r : c->pending) {
    float ms = 0.f;
    PD_HIP(hipEventElapsedTime(&ms, r.a, r.b));
    c->prof_ms[r.cls] += ms;
    c->prof_n[r.cls] += 1;
    c->pool.push_back(r.a);
    c->pool.push_back(r.b);
  }
  c->pending.clear();
}

template <typename T>
T* dalloc(Env& e, size_t n) {
  const size_t cls = DevMem::size_class(std::max<size_t>(n, 1) * sizeof(T));
  void* p = e.mem->get(cls);
  e.allocs.emplace_back(p, cls);
  e.bytes += cls;
  return static_cast<T*>(p);
}

void free_env_device(Env& e) {
  for (auto& a : e.allocs) e.mem->put(a.first, a.second);
  e.allocs.clear();
  e.bytes = 0;
}

Ctx* C(void* c) { return static_cast<Ctx*>(c); }

Env& get_env(Ctx* c, int64_t id) {
  auto it = c->envs.find(id);
  if (it == c->envs.end()) throw PdError(PD_E_ARG, "unknown env id " + std::to_string(id));
  return *it->second;
}

void check_sym(const std::vector<double>& m, int nt, const char* name) {
  for (int a = 0; a < nt; ++a)
    for (int b = 0; b < nt; ++b)
      if (!(m[a + nt * b] == m[b + nt * a]) && !(std::isnan(m[a + nt * b]) && std::isnan(m[b + nt * a])))
        throw PdError(PD_E_ARG, std::string(name) + " matrix must be symmetric (make_matrix)");
}

std::vector<double> copy_mat(const double* p, int nt, const char* name, bool required) {
  std::vector<double> out((size_t)nt * nt, 0.0);
  if (!p) {
    if (required) throw PdError(PD_E_ARG, std::string(name) + " is required");
    return out;
  }
  std::memcpy(out.data(), p, sizeof(double) * nt * nt);
  check_sym(out, nt, name);
  return out;
}

// ---- family build ---------------------------------------------------------
struct FamTmp {
  FamilyGrid g{};
  int32_t *cell_of = nullptr, *cell_count = nullptr, *cell_start = nullptr, *cursor = nullptr, *items = nullptr;
  DevMem* dm = nullptr;
  std::vector<std::pair<void*, size_t>> mem;
};

void fam_free(FamTmp& t) {
  for (auto& a : t.mem) t.dm->put(a.first, a.second);
  t.mem.clear();
}

template <typename T>
T* talloc(FamTmp& t, size_t n) {
  const size_t cls = DevMem::size_class(std::max<size_t>(n, 1) * sizeof(T));
  void* p = t.dm->get(cls);
  t.mem.emplace_back(p, cls);
  return static_cast<T*>(p);
}

void build_cells(Ctx* c, Env& e, HostBlock& b, FamTmp& t) {
  const int n = (int)b.d.n;
  double lo[3], hi[3];
  for (int d = 0; d < 3; ++d) { lo[d] = b.x[d]; hi[d] = b.x[d]; }
  for (int i = 0; i < n; ++i)
    for (int d = 0; d < 3; ++d) {
      lo[d] = std::min(lo[d], b.x[3 * i + d]);
      hi[d] = std::max(hi[d], b.x[3 * i + d]);
    }
  double w = b.d.horizon * (1.0 + 9.313225746154785e-10);
  if (!(w > 0.0) || !std::isfinite(w)) throw PdError(PD_E_ARG, "horizon must be positive and finite");
  int64_t tot = 0;
  int nc[3];
  for (int it = 0; it < 200; ++it) {
    tot = 1;
    for (int d = 0; d < 3; ++d) {
      nc[d] = (int)std::min(std::floor((hi[d] - lo[d]) / w) + 1.0, 1.0e6);
      tot *= nc[d];
    }
    if (tot <= 8LL * n + 64) break;
    w *= 1.25;
  }
  for (int d = 0; d < 3; ++d) { t.g.lo[d] = lo[d]; t.g.nc[d] = nc[d]; }
  t.g.w = w;
  t.cell_of = talloc<int32_t>(t, n);
  t.cell_count = talloc<int32_t>(t, tot);
  t.cell_start = talloc<int32_t>(t, tot);
  t.cursor = talloc<int32_t>(t, tot);
  t.items = talloc<int32_t>(t, n);
  PD_HIP(hipMemsetAsync(t.cell_count, 0, sizeof(int32_t) * tot, c->stream));
  PD_HIP(hipMemsetAsync(t.cursor, 0, sizeof(int32_t) * tot, c->stream));
  pd_launch_cell_of(c->stream, e.dev.x, b.first, n, t.g, t.cell_of, t.cell_count);
  std::vector<int32_t> cnt(tot), start(tot);
  PD_HIP(hipMemcpyAsync(cnt.data(), t.cell_count, sizeof(int32_t) * tot, hipMemcpyDeviceToHost, c->stream));
  PD_HIP(hipStreamSynchronize(c->stream));
  int64_t run = 0;
  for (int64_t i = 0; i < tot; ++i) { start[i] = (int32_t)run; run += cnt[i]; }
  PD_HIP(hipMemcpyAsync(t.cell_start, start.data(), sizeof(int32_t) * tot, hipMemcpyHostToDevice, c->stream));
  pd_launch_cell_fill(c->stream, n, t.cell_of, t.cell_start, t.cursor, t.items, b.first);
  PD_HIP(hipStreamSynchronize(c->stream));  // `start` goes out of scope
}

void upload_env_table(Ctx* c, Env& e) {
  // the device table's buffer indices and step counter are advanced on the
  // device (k_book); the host tracks the same values in Env
  e.dev.cur = e.cur;
  e.dev.edp_cur = e.edp_cur;
  e.dev.time_step = e.time_step;
  PD_HIP(hipMemcpyAsync(c->d_envs + e.slot, &e.dev, sizeof(EnvDev), hipMemcpyHostToDevice, c->stream));
}

void ensure_slot(Ctx* c, Env& e) {
  if (e.slot >= 0) return;
  if (c->free_slots.empty()) {
    // an in-flight run reads the table: let it finish before the table moves
    if (c->inflight.on) PD_HIP(hipStreamSynchronize(c->run_stream));
    const int32_t ncap = std::max(64, c->slot_cap * 2);
    EnvDev* nt = nullptr;
    PD_HIP(hipMalloc(&nt, sizeof(EnvDev) * ncap));
    if (c->d_envs) {
      PD_HIP(hipMemcpyAsync(nt, c->d_envs, sizeof(EnvDev) * c->slot_cap, hipMemcpyDeviceToDevice, c->stream));
      PD_HIP(hipStreamSynchronize(c->stream));
      PD_HIP(hipFree(c->d_envs));
    }
    for (int32_t s = ncap - 1; s >= c->slot_cap; --s) c->free_slots.push_back(s);
    c->d_envs = nt;
    c->slot_cap = ncap;
  }
  e.slot = c->free_slots.back();
  c->free_slots.pop_back();
}

// ---- internal particle order ---------------------------------------------------
// Device arrays may hold a block's particles in an internal order (perm:
// internal id -> reference index, iperm: back; empty = identity).  All
// host-facing arrays stay in reference order; perm / iperm translate at the
// boundary.  The reference-order form always keeps the reference order (its
// summation order is defined by reference indices).
inline int64_t internal_of(const Env& e, int64_t r) { return e.iperm.empty() ? r : e.iperm[r]; }
inline int64_t reference_of(const Env& e, int64_t p) { return e.perm.empty() ? p : e.perm[p]; }

uint64_t morton_spread(uint64_t v) {
  v &= 0x1fffffULL;
  v = (v | v << 32) & 0x1f00000000ffffULL;
  v = (v | v << 16) & 0x1f0000ff0000ffULL;
  v = (v | v << 8) & 0x100f00f00f00f00fULL;
  v = (v | v << 4) & 0x10c30c30c30c30c3ULL;
  v = (v | v << 2) & 0x1249249249249249ULL;
  return v;
}

// Tiled passes for the FP64 form (PERIDYN_TILES=0 turns them off).
bool tiles_enabled() {
  const char* v = std::getenv("PERIDYN_TILES");
  return !(v && std::string(v) == "0");
}

// Brick of a tile in lattice cells for lgG: 4 chunks of C = 64/G particles,
// each chunk one bx x by layer of the brick.
void brick_dims(int lgG, int& bx, int& by, int& bz) {
  if (lgG == 1) { bx = 8; by = 4; bz = 4; }
  else if (lgG == 2) { bx = 4; by = 4; bz = 4; }
  else { bx = 4; by = 2; bz = 4; }
}

// Brick order of every block: particles binned into lattice cells of the
// block's particle size, cells grouped into bricks (brick_dims), bricks in
// z-y-x order and cells inside a brick likewise; ties keep the input order.
// e.brick numbers each particle's brick (unique over the env).
void build_brick_order(Env& e) {
  e.perm.resize(e.N);
  e.iperm.resize(e.N);
  e.brick.assign(e.N, 0);
  e.brick_pos.clear();
  int bx, by, bz;
  brick_dims(e.lgG, bx, by, bz);
  int32_t next_brick = 0;
  for (auto& b : e.blocks) {
    const int64_t n = b.d.n;
    if (n == 0) continue;
    double lo[3] = {INFINITY, INFINITY, INFINITY};
    for (int64_t i = 0; i < n; ++i)
      for (int d = 0; d < 3; ++d) lo[d] = std::min(lo[d], b.x[3 * i + d]);
    const double ps = b.d.particle_size > 0.0 ? b.d.particle_size : b.d.horizon / 3.0;
    std::vector<int64_t> q(3 * n);
    int64_t nb[3] = {1, 1, 1};
    const int dims[3] = {bx, by, bz};
    for (int64_t i = 0; i < n; ++i)
      for (int d = 0; d < 3; ++d) {
        const double v = std::floor((b.x[3 * i + d] - lo[d]) / ps + 0.5);
        const int64_t qi = (int64_t)std::min(std::max(v, 0.0), 1048575.0);
        q[3 * i + d] = qi;
        nb[d] = std::max(nb[d], qi / dims[d] + 1);
      }
    // brick and cell of every particle (brick dimensions are powers of two),
    // then a stable sort by (brick, cell): ties keep the input order
    const int lx = __builtin_ctz(bx), ly = __builtin_ctz(by), lz = __builtin_ctz(bz);
    const int64_t nbr = nb[0] * nb[1] * nb[2];
    std::vector<int64_t> brk(n);
    std::vector<uint8_t> cel(n);
    for (int64_t i = 0; i < n; ++i) {
      const int64_t* qq = &q[3 * i];
      brk[i] = ((qq[2] >> lz) * nb[1] + (qq[1] >> ly)) * nb[0] + (qq[0] >> lx);
      cel[i] = (uint8_t)((((qq[2] & (bz - 1)) << ly | (qq[1] & (by - 1))) << lx) | (qq[0] & (bx - 1)));
    }
    std::vector<int32_t> order(n);
    if (nbr < (1LL << 30) && bx * by * bz <= 128) {
      // LSD radix sort of the keys brick << 7 | cell, 11 bits per pass, as many
      // passes as the largest key has digits
      std::vector<int32_t> tmp(n);
      for (int64_t i = 0; i < n; ++i) order[i] = (int32_t)i;
      auto key = [&](int32_t i) { return ((uint64_t)brk[i] << 7) | cel[i]; };
      uint64_t kmax = 0;
      for (int64_t i = 0; i < n; ++i) kmax = std::max(kmax, key((int32_t)i));
      std::vector<int64_t> cnt(1 << 11);
      for (int shift = 0; shift < 64 && (kmax >> shift) != 0; shift += 11) {
        std::fill(cnt.begin(), cnt.end(), 0);
        for (int64_t k = 0; k < n; ++k) ++cnt[(key(order[k]) >> shift) & 0x7ff];
        int64_t run = 0;
        for (auto& c : cnt) { const int64_t v = c; c = run; run += v; }
        for (int64_t k = 0; k < n; ++k) tmp[cnt[(key(order[k]) >> shift) & 0x7ff]++] = order[k];
        order.swap(tmp);
      }
    } else {
      for (int64_t i = 0; i < n; ++i) order[i] = (int32_t)i;
      std::sort(order.begin(), order.end(), [&](int32_t a, int32_t c) {
        if (brk[a] != brk[c]) return brk[a] < brk[c];
        if (cel[a] != cel[c]) return cel[a] < cel[c];
        return a < c;
      });
    }
    int64_t prev = -1;
    for (int64_t k = 0; k < n; ++k) {
      const int32_t i = order[k];
      const int64_t br = brk[i];
      if (k > 0 && br != prev) ++next_brick;
      if (k == 0 || br != prev)
        e.brick_pos.push_back({(int32_t)(br % nb[0]), (int32_t)((br / nb[0]) % nb[1]), (int32_t)(br / (nb[0] * nb[1]))});
      prev = br;
      e.perm[b.first + k] = b.first + i;
      e.iperm[b.first + i] = (int32_t)(b.first + k);
      e.brick[b.first + k] = next_brick;
    }
    ++next_brick;
  }
}

void build_order(Env& e) {
  e.perm.clear();
  e.iperm.clear();
  e.brick.clear();
  e.brick_pos.clear();
  if (e.precision == PD_PREC_EXACT) return;
  if (e.want_tiles) {
    build_brick_order(e);
    return;
  }
  // Measured on lattice specimens the input (lattice row) order is faster: the
  // k-th neighbours of consecutive lanes are consecutive particles, so neighbour
  // gathers coalesce; PERIDYN_ORDER=morton selects the Z-order layout.
  const char* ord = std::getenv("PERIDYN_ORDER");
  if (!ord || std::string(ord) != "morton") return;
  e.perm.resize(e.N);
  e.iperm.resize(e.N);
  for (auto& b : e.blocks) {
    const int64_t n = b.d.n;
    double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int64_t i = 0; i < n; ++i)
      for (int d = 0; d < 3; ++d) {
        lo[d] = std::min(lo[d], b.x[3 * i + d]);
        hi[d] = std::max(hi[d], b.x[3 * i + d]);
      }
    double cell = b.d.horizon / 2.0;
    for (int d = 0; d < 3; ++d) cell = std::max(cell, (hi[d] - lo[d]) / 2097151.0);
    std::vector<std::pair<uint64_t, int32_t>> key(n);
    for (int64_t i = 0; i < n; ++i) {
      uint64_t k = 0;
      for (int d = 0; d < 3; ++d) {
        const double q = std::floor((b.x[3 * i + d] - lo[d]) / cell);
        const uint64_t qi = (uint64_t)std::min(std::max(q, 0.0), 2097151.0);
        k |= morton_spread(qi) << d;
      }
      key[i] = {k, (int32_t)i};
    }
    std::stable_sort(key.begin(), key.end(),
                     [](const std::pair<uint64_t, int32_t>& a, const std::pair<uint64_t, int32_t>& c) {
                       return a.first < c.first;
                     });
    for (int64_t q = 0; q < n; ++q) {
      e.perm[b.first + q] = b.first + key[q].second;
      e.iperm[b.first + key[q].second] = (int32_t)(b.first + q);
    }
  }
}

// LDS a tile may use for its largest pass (2 workgroups fit a CU's 160 KiB)
constexpr size_t kTileLdsBudget = 80 * 1024;
// tiles of one launch that keep every CU busy for about two rounds (5
// workgroups per CU on 256 CUs)
constexpr int kMergeTiles = 2 * 5 * 256;

// Tiles of a brick-ordered env: consecutive chunks of one brick, up to
// PD_TILE_CHUNKS per tile, or of two partial bricks next to each other
// (merge_partial_tiles).  Each tile's halo is the sorted set of neighbour ids
// of its particles; pk receives every bond's halo-local index with the intact
// bit (every bond is intact after the family build).  Leaves e.tiled false
// (untiled passes) when a halo exceeds 15-bit indices or the LDS budget.
// LDS record bytes of the env's largest tiled pass
size_t tile_record_bytes(const Env& e) {
  size_t rb = 0;
  for (auto& b : e.blocks) {
    rb = std::max<size_t>(rb, pd_tile_record_bytes(2, b.d.kind, e.precision, 0));
    if (b.d.kind != PD_BB) rb = std::max<size_t>(rb, pd_tile_record_bytes(0, b.d.kind, e.precision, 0));
    if (b.d.kind == PD_EP) rb = std::max<size_t>(rb, pd_tile_record_bytes(1, PD_EP, e.precision, 0));
  }
  return rb;
}

// A tile before its build: the indices (into Env::chunks) of its chunks.
struct TileChunks {
  int n = 0;
  int ck[PD_TILE_CHUNKS] = {};
};

Tile tile_of(const Env& e, const TileChunks& r) {
  Tile T;
  std::memset(&T, 0, sizeof(T));
  T.block = e.chunks[r.ck[0]].block;
  T.nch = r.n;
  for (int k = 0; k < r.n; ++k) T.ch[k] = e.chunks[r.ck[k]];
  return T;
}

// Device build (k_halo_count / k_halo_fill, one workgroup per tile): tiles of
// at most 16384 bond slots (an LDS hash set of twice that).  Returns false
// (nothing allocated) when a halo exceeds the limits.
bool build_tiles_device(Ctx* c, Env& e, const std::vector<TileChunks>& tr, int ncap) {
  hipStream_t s = c->stream;
  const int nt = (int)tr.size();
  std::vector<Tile> tt(nt);
  for (int t = 0; t < nt; ++t) tt[t] = tile_of(e, tr[t]);
  Tile* dt = nullptr;
  int32_t* dcnt = nullptr;
  std::vector<int32_t> cnt(nt);
  const size_t cls_t = DevMem::size_class(sizeof(Tile) * std::max(nt, 1));
  const size_t cls_c = DevMem::size_class(sizeof(int32_t) * std::max(nt, 1));
  auto release = [&] {
    c->mem.put(dt, cls_t);
    c->mem.put(dcnt, cls_c);
    dt = nullptr;
    dcnt = nullptr;
  };
  try {
    dt = static_cast<Tile*>(c->mem.get(cls_t));
    dcnt = static_cast<int32_t*>(c->mem.get(cls_c));
    PD_HIP(hipMemcpyAsync(dt, tt.data(), sizeof(Tile) * nt, hipMemcpyHostToDevice, s));
    const int hc = std::max(2 * ncap, 512);  // LDS hash set, at most half full
    pd_launch_halo_count(s, dt, nt, e.dev.col, dcnt, hc);
    PD_HIP(hipGetLastError());
    PD_HIP(hipMemcpyAsync(cnt.data(), dcnt, sizeof(int32_t) * nt, hipMemcpyDeviceToHost, s));
    PD_HIP(hipStreamSynchronize(s));
    size_t hmax = 1, total = 0;
    for (int t = 0; t < nt; ++t) {
      hmax = std::max<size_t>(hmax, (size_t)cnt[t]);
      tt[t].hoff = (uint32_t)total;
      tt[t].hn = cnt[t];
      total += (size_t)cnt[t];
    }
    if (hmax > 0x7fff || tile_record_bytes(e) * hmax > kTileLdsBudget || total * sizeof(int32_t) >= PD_MAX_ARRAY_BYTES) {
      release();
      return false;
    }
    PD_HIP(hipMemcpyAsync(dt, tt.data(), sizeof(Tile) * nt, hipMemcpyHostToDevice, s));
    e.dev.halo = dalloc<int32_t>(e, std::max<size_t>(total, 1));
    e.dev.pk = dalloc<uint64_t>(e, (size_t)std::max<int64_t>(e.npk, 1));
    PD_HIP(hipMemsetAsync(e.dev.pk, 0, sizeof(uint64_t) * (size_t)std::max<int64_t>(e.npk, 1), s));
    pd_launch_halo_fill(s, dt, nt, e.dev.col, e.dev.halo, e.dev.pk, hc);
    PD_HIP(hipGetLastError());
    PD_HIP(hipStreamSynchronize(s));
    e.hmax = (int32_t)hmax;
  } catch (...) {
    (void)hipStreamSynchronize(s);
    release();
    throw;
  }
  release();
  e.tiles = std::move(tt);
  return true;
}

// Halo sizes (distinct neighbour ids) of the tiles with chunk ranges tr, by
// k_halo_count on the device.
std::vector<int32_t> halo_counts_device(Ctx* c, Env& e, const std::vector<TileChunks>& tr, int ncap) {
  hipStream_t s = c->stream;
  const int nt = (int)tr.size();
  std::vector<Tile> tt(nt);
  for (int t = 0; t < nt; ++t) tt[t] = tile_of(e, tr[t]);
  std::vector<int32_t> cnt(nt);
  const size_t cls_t = DevMem::size_class(sizeof(Tile) * std::max(nt, 1));
  const size_t cls_c = DevMem::size_class(sizeof(int32_t) * std::max(nt, 1));
  Tile* dt = nullptr;
  int32_t* dcnt = nullptr;
  try {
    dt = static_cast<Tile*>(c->mem.get(cls_t));
    dcnt = static_cast<int32_t*>(c->mem.get(cls_c));
    PD_HIP(hipMemcpyAsync(dt, tt.data(), sizeof(Tile) * nt, hipMemcpyHostToDevice, s));
    pd_launch_halo_count(s, dt, nt, e.dev.col, dcnt, std::max(2 * ncap, 512));
    PD_HIP(hipGetLastError());
    PD_HIP(hipMemcpyAsync(cnt.data(), dcnt, sizeof(int32_t) * nt, hipMemcpyDeviceToHost, s));
    PD_HIP(hipStreamSynchronize(s));
  } catch (...) {
    (void)hipStreamSynchronize(s);
    c->mem.put(dt, cls_t);
    c->mem.put(dcnt, cls_c);
    throw;
  }
  c->mem.put(dt, cls_t);
  c->mem.put(dcnt, cls_c);
  return cnt;
}

// The same sizes on the host, from the downloaded neighbour columns.
std::vector<int32_t> halo_counts_host(const Env& e, const std::vector<int32_t>& col,
                                      const std::vector<TileChunks>& tr) {
  std::vector<int32_t> cnt(tr.size(), 0);
  std::vector<int32_t> stamp((size_t)e.N, -1);
  for (size_t t = 0; t < tr.size(); ++t)
    for (int q = 0; q < tr[t].n; ++q) {
      const Chunk& ch = e.chunks[tr[t].ck[q]];
      const size_t nsl = (size_t)PD_WAVE * ch.width;
      for (size_t i = 0; i < nsl; ++i) {
        const int32_t j = col[ch.slot0 + i];
        if (j >= 0 && stamp[j] != (int32_t)t) {
          stamp[j] = (int32_t)t;
          ++cnt[t];
        }
      }
    }
  return cnt;
}

// Partial bricks (fewer than PD_TILE_CHUNKS chunks: the last brick layer of a
// block whose extent is not a multiple of the brick's) leave waves of their
// workgroup idle while the workgroup holds a full LDS class.  Two partial
// bricks next to each other (in x, y or z) become one tile when their chunks
// fit one and the merged halo is no larger than the largest halo of the env's
// unmerged tiles, so that the env's LDS classes are unchanged.  Candidate
// pairs are taken greedily, those that fill a whole tile first, then in brick
// order.  counts(tiles) returns the halo sizes of chunk lists.
template <class F>
void merge_partial_tiles(const Env& e, std::vector<TileChunks>& tr, int64_t maxent, F&& counts) {
  auto ent = [&](const TileChunks& a, const TileChunks& b) {
    int64_t v = 0;
    for (int k = 0; k < a.n; ++k) v += (int64_t)PD_WAVE * e.chunks[a.ck[k]].width;
    for (int k = 0; k < b.n; ++k) v += (int64_t)PD_WAVE * e.chunks[b.ck[k]].width;
    return v;
  };
  // partial tiles by (block, brick position)
  std::map<std::array<int32_t, 4>, int> at;
  for (int t = 0; t < (int)tr.size(); ++t) {
    if (tr[t].n >= PD_TILE_CHUNKS) continue;
    const Chunk& c0 = e.chunks[tr[t].ck[0]];
    const auto& bp = e.brick_pos[e.brick[c0.first]];
    at[{c0.block, bp[0], bp[1], bp[2]}] = t;
  }
  std::vector<std::pair<int, int>> cand;  // (t, u): merge tiles t < u
  for (const auto& kv : at) {
    const int t = kv.second;
    for (int d = 0; d < 3; ++d) {
      std::array<int32_t, 4> k = kv.first;
      ++k[1 + d];
      const auto it = at.find(k);
      if (it == at.end()) continue;
      const int u = it->second;
      if (tr[t].n + tr[u].n > PD_TILE_CHUNKS || ent(tr[t], tr[u]) > maxent) continue;
      cand.emplace_back(std::min(t, u), std::max(t, u));
    }
  }
  if (cand.empty()) return;
  std::sort(cand.begin(), cand.end(), [&](const std::pair<int, int>& a, const std::pair<int, int>& b) {
    const bool fa = tr[a.first].n + tr[a.second].n == PD_TILE_CHUNKS;
    const bool fb = tr[b.first].n + tr[b.second].n == PD_TILE_CHUNKS;
    if (fa != fb) return fa;
    return a < b;
  });
  std::vector<TileChunks> q(tr);
  for (const auto& pr : cand) {
    TileChunks m = tr[pr.first];
    for (int k = 0; k < tr[pr.second].n; ++k) m.ck[m.n++] = tr[pr.second].ck[k];
    q.push_back(m);
  }
  const std::vector<int32_t> cnt = counts(q);
  int32_t h0 = 0;
  for (size_t t = 0; t < tr.size(); ++t) h0 = std::max(h0, cnt[t]);
  std::vector<int> partner(tr.size(), -1);
  for (size_t i = 0; i < cand.size(); ++i) {
    const int t = cand[i].first, u = cand[i].second;
    if (partner[t] >= 0 || partner[u] >= 0 || cnt[tr.size() + i] > h0) continue;
    partner[t] = u;
    partner[u] = t;
  }
  // a merged tile takes the place of its first brick
  std::vector<TileChunks> out;
  out.reserve(tr.size());
  for (int t = 0; t < (int)tr.size(); ++t) {
    if (partner[t] < 0) {
      out.push_back(tr[t]);
    } else if (partner[t] > t) {
      TileChunks m = tr[t];
      const TileChunks& o = tr[partner[t]];
      for (int k = 0; k < o.n; ++k) m.ck[m.n++] = o.ck[k];
      out.push_back(m);
    }
  }
  tr.swap(out);
}

void build_tiles(Ctx* c, Env& e) {
  hipStream_t s = c->stream;
  // tiles: the chunks of one brick each
  std::vector<TileChunks> tr;
  for (int k = 0; k < (int)e.chunks.size();) {
    const int br = e.brick[e.chunks[k].first], blk = e.chunks[k].block;
    TileChunks r;
    r.ck[r.n++] = k++;
    while (r.n < PD_TILE_CHUNKS && k < (int)e.chunks.size() && e.chunks[k].block == blk &&
           e.brick[e.chunks[k].first] == br)
      r.ck[r.n++] = k++;
    tr.push_back(r);
  }
  int64_t maxent = 1;
  for (auto& r : tr) {
    int64_t ent = 0;
    for (int k = 0; k < r.n; ++k) ent += (int64_t)PD_WAVE * e.chunks[r.ck[k]].width;
    maxent = std::max(maxent, ent);
  }
  int ncap = 1;
  while (ncap < maxent) ncap <<= 1;
  // PERIDYN_TILE_BUILD=host selects the host build below (it is the path of
  // tiles beyond the device sort buffer; tests compare the two)
  const char* hb = std::getenv("PERIDYN_TILE_BUILD");
  const bool host_build = (hb && std::string(hb) == "host") || ncap > 16384;
  std::vector<int32_t> col;
  if (host_build) {
    col.resize((size_t)std::max<int64_t>(e.nslots, 1));
    PD_HIP(hipMemcpyAsync(col.data(), e.dev.col, sizeof(int32_t) * (size_t)e.nslots, hipMemcpyDeviceToHost, s));
    PD_HIP(hipStreamSynchronize(s));
  }
  // merged tiles hold at most maxent bond slots, so ncap stands
  merge_partial_tiles(e, tr, maxent, [&](const std::vector<TileChunks>& q) {
    return host_build ? halo_counts_host(e, col, q) : halo_counts_device(c, e, q, ncap);
  });
  const int nt = (int)tr.size();
  if (!host_build) {
    e.tiled = build_tiles_device(c, e, tr, ncap);
    if (!e.tiled) e.tiles.clear();
    return;
  }
  // host build for tiles beyond the device sort buffer
  std::vector<std::vector<int32_t>> halos(nt);
  std::vector<uint64_t> pk((size_t)std::max<int64_t>(e.npk, 1), 0ull);
  const int64_t N = e.N;
  auto work = [&](int t0, int t1) {
    std::vector<int32_t> stamp(N, -1), pos(N, 0);
    for (int t = t0; t < t1; ++t) {
      std::vector<int32_t>& h = halos[t];
      for (int q = 0; q < tr[t].n; ++q) {
        const Chunk& ch = e.chunks[tr[t].ck[q]];
        const size_t nsl = (size_t)PD_WAVE * ch.width;
        for (size_t i = 0; i < nsl; ++i) {
          const int32_t j = col[ch.slot0 + i];
          if (j >= 0 && stamp[j] != t) {
            stamp[j] = t;
            h.push_back(j);
          }
        }
      }
      std::sort(h.begin(), h.end());
      if (h.size() > 0x7fff) continue;  // rejected below
      for (size_t i = 0; i < h.size(); ++i) pos[h[i]] = (int32_t)i;
      for (int q = 0; q < tr[t].n; ++q) {
        const Chunk& ch = e.chunks[tr[t].ck[q]];
        for (int st = 0; st < ch.width; ++st)
          for (int l = 0; l < PD_WAVE; ++l) {
            const int32_t j = col[ch.slot0 + (size_t)PD_WAVE * st + l];
            if (j < 0) continue;
            const uint64_t v = 0x8000u | (uint32_t)pos[j];
            pk[ch.pk0 + (size_t)PD_WAVE * (st >> 2) + l] |= v << (16 * (st & 3));
          }
      }
    }
  };
  const int nth = std::max(1, std::min<int>({16, (int)std::thread::hardware_concurrency(), nt}));
  std::vector<std::thread> pool;
  for (int w = 0; w < nth; ++w) pool.emplace_back(work, (int)((int64_t)nt * w / nth), (int)((int64_t)nt * (w + 1) / nth));
  for (auto& th : pool) th.join();
  size_t hmax = 1, total = 0;
  for (auto& h : halos) {
    hmax = std::max(hmax, h.size());
    total += h.size();
  }
  if (hmax > 0x7fff || tile_record_bytes(e) * hmax > kTileLdsBudget ||
      total * sizeof(int32_t) >= PD_MAX_ARRAY_BYTES) {
    e.tiles.clear();
    e.tiled = false;
    return;
  }
  std::vector<int32_t> hall;
  hall.reserve(total);
  e.tiles.assign(nt, Tile{});
  for (int t = 0; t < nt; ++t) {
    Tile& T = e.tiles[t];
    T = tile_of(e, tr[t]);
    T.env = 0;
    T.hn = (int32_t)halos[t].size();
    T.hoff = (uint32_t)hall.size();
    hall.insert(hall.end(), halos[t].begin(), halos[t].end());
  }
  e.hmax = (int32_t)hmax;
  e.dev.halo = dalloc<int32_t>(e, std::max<size_t>(hall.size(), 1));
  e.dev.pk = dalloc<uint64_t>(e, pk.size());
  if (!hall.empty())
    PD_HIP(hipMemcpyAsync(e.dev.halo, hall.data(), sizeof(int32_t) * hall.size(), hipMemcpyHostToDevice, s));
  PD_HIP(hipMemcpyAsync(e.dev.pk, pk.data(), sizeof(uint64_t) * pk.size(), hipMemcpyHostToDevice, s));
  PD_HIP(hipStreamSynchronize(s));
  e.tiled = true;
}

// Lanes per particle (log2) of a tiled env without an explicit schedule: 4
// (tiles of 4x4x4 particles) unless the horizon makes their halos outgrow the
// smaller compiled LDS stride, then 8 (4x2x4 tiles, ~25% smaller halos, so more
// workgroups fit a CU).  The halo of a brick spanning Lx x Ly x Lz particle
// spacings is estimated by the volume of the brick dilated by the horizon r
// (Steiner's formula, in particle volumes), which tracks lattice counts within
// ~10% for r = 2.5 .. 5 spacings.
int tiled_lanes_log2(const Env& e) {
  double r = 0.0;
  for (auto& b : e.blocks) {
    const double ps = b.d.particle_size > 0.0 ? b.d.particle_size : b.d.horizon / 3.0;
    r = std::max(r, b.d.horizon / ps);
  }
  const double pi = 3.141592653589793;
  auto halo = [&](double lx, double ly, double lz) {
    return (int)(lx * ly * lz + 2.0 * r * (lx * ly + ly * lz + lz * lx) + pi * r * r * (lx + ly + lz) +
                 (4.0 / 3.0) * pi * r * r * r);
  };
  const int h4 = halo(3.0, 3.0, 3.0), h8 = halo(3.0, 1.0, 3.0);
  if (h4 <= 608) return 2;  // 4x4x4 halos this small take the densest LDS classes anyway
  int c4 = 0, c8 = 0;  // LDS classes of the force pass
  for (auto& b : e.blocks) {
    c4 = std::max(c4, pd_tile_class(2, b.d.kind, e.precision, 1, h4));
    c8 = std::max(c8, pd_tile_class(2, b.d.kind, e.precision, 1, h8));
  }
  return c8 < c4 ? 3 : 2;
}

void commit_env(Ctx* c, Env& e) {
  if (e.committed) throw PdError(PD_E_STATE, "env already committed");
  if (e.blocks.empty()) throw PdError(PD_E_ARG, "env has no material blocks");
  int64_t N = 0;
  for (auto& b : e.blocks) { b.first = (int32_t)N; N += b.d.n; }
  if (N >= (1LL << 31) - 64 || (uint64_t)N * sizeof(double4) >= PD_MAX_ARRAY_BYTES)
    throw PdError(PD_E_ARG, "env too large: its particle arrays would exceed 4 GiB; split it into several envs");
  e.N = N;
  // bond layout: the reference-order form walks each row on one lane; the
  // untiled forms spread a particle over G lanes so that small specimens still
  // fill the chip (>= 786,432 lanes = 48 waves per CU over 256 CUs; measured
  // best untiled on the 100k-particle EP plates (G=8) and the 721k-particle BB
  // disk (G=2)), unless a schedule was set explicitly
  // tiled passes: FP64 form, 2..8 lanes per particle, single-type blocks and
  // integer contact exponents (the tiled force kernel carries no general pow);
  // 4 lanes per particle unless a schedule was set (measured best on the EP
  // plates, equal to 2 on the BB disk) or the horizon calls for 8
  // (tiled_lanes_log2)
  bool tileable = (e.precision == PD_PREC_FP64 || e.precision == PD_PREC_FP32C) && tiles_enabled();
  for (auto& b : e.blocks) tileable = tileable && b.d.ntype == 1;
  // the FP32C elastoplastic passes are untiled kernels (k_td32, k_force32<EP>)
  if (e.precision == PD_PREC_FP32C)
    for (auto& b : e.blocks) tileable = tileable && b.d.kind != PD_EP;
  for (auto& hc : e.contacts) tileable = tileable && hc.d.exponent_is_int;
  if (e.precision == PD_PREC_EXACT) {
    e.lgG = 0;
  } else if (e.lgG < 0 && tileable) {
    e.lgG = tiled_lanes_log2(e);
  } else if (e.lgG < 0) {
    e.lgG = 0;
    while (e.lgG < 4 && (N << e.lgG) < 786432) ++e.lgG;
  }
  e.want_tiles = tileable && e.lgG >= 1 && e.lgG <= 3;
  e.tiled = false;
  build_order(e);
  // ---- particle arrays (host staging, internal order)
  HostStage& hs = c->stage;
  hs.begin((size_t)N * kStagePerParticle + kStagePad, c->stream);
  auto x4 = hs.take<double4>(N), y4 = hs.take<double4>(N), v4 = hs.take<double4>(N);
  auto pblock = hs.take<int32_t>(N), bcid = hs.take<int32_t>(N);
  auto ptype = hs.take<uint8_t>(N);
  auto inv_den = hs.take<double>(N);
  std::fill(bcid.p, bcid.p + N, -1);
  for (size_t bi = 0; bi < e.blocks.size(); ++bi) {
    HostBlock& b = e.blocks[bi];
    for (int64_t i = 0; i < b.d.n; ++i) {
      const int64_t p = internal_of(e, b.first + i);
      x4[p] = make_double4(b.x[3 * i], b.x[3 * i + 1], b.x[3 * i + 2], b.vol[i]);
      y4[p] = make_double4(b.y[3 * i], b.y[3 * i + 1], b.y[3 * i + 2], 0.0);
      v4[p] = make_double4(b.v[3 * i], b.v[3 * i + 1], b.v[3 * i + 2], 0.0);
      pblock[p] = (int32_t)bi;
      const int64_t t = b.type[i] - b.d.type_start;
      ptype[p] = (uint8_t)t;
      // Env: mass = volume * density[t]; density = mass ./ volume (environment.jl:173-214)
      const double mass = b.vol[i] * b.density[t];
      const double den = mass / b.vol[i];
      inv_den[p] = 1.0 / den;
    }
  }
  bool any_onlyatstart = false;
  for (size_t k = 0; k < e.bcs.size(); ++k) {
    if ((int64_t)e.bcs[k].mask.size() != N) throw PdError(PD_E_ARG, "BC mask length must equal env particles");
    for (int64_t r = 0; r < N; ++r)
      if (e.bcs[k].mask[r]) bcid[internal_of(e, r)] = (int32_t)k;
    any_onlyatstart = any_onlyatstart || (e.bcs[k].d.options & PD_BC_ONLYATSTART);
  }
  e.block_rigid.assign(e.blocks.size(), 0);
  e.block_rigid_ok.assign(e.blocks.size(), 0);
  for (size_t bi = 0; bi < e.blocks.size(); ++bi) {
    const HostBlock& b = e.blocks[bi];
    if (b.d.n <= 0) continue;
    const int32_t k0 = bcid[internal_of(e, b.first)];
    bool one = k0 >= 0 && !(e.bcs[k0].d.options & PD_BC_ONLYATSTART);
    for (int64_t i = 1; one && i < b.d.n; ++i) one = bcid[internal_of(e, b.first + i)] == k0;
    e.block_rigid[bi] = one ? 1 : 0;
  }
  EnvDev& D = e.dev;
  std::memset(&D, 0, sizeof(D));
  D.N = (int32_t)N;
  D.nblocks = (int32_t)e.blocks.size();
  D.ncontacts = (int32_t)e.contacts.size();
  D.nbc = (int32_t)e.bcs.size();
  D.precision = e.precision;
  D.dt = e.dt;
  D.x = dalloc<double4>(e, N);
  D.y[0] = dalloc<double4>(e, N);
  D.y[1] = dalloc<double4>(e, N);
  D.v = dalloc<double4>(e, N);
  D.f = dalloc<double4>(e, N);
  D.pblock = dalloc<int32_t>(e, N);
  D.ptype = dalloc<uint8_t>(e, N);
  D.inv_den = dalloc<double>(e, N);
  D.bcid = dalloc<int32_t>(e, N);
  D.deg = dalloc<int32_t>(e, N);
  D.pslot0 = dalloc<uint32_t>(e, N);
  D.m = dalloc<double>(e, N);
  D.theta = dalloc<double>(e, N);
  D.td = dalloc<double>(e, N);
  hipStream_t s = c->stream;
  PD_HIP(hipMemcpyAsync(D.x, x4.data(), sizeof(double4) * N, hipMemcpyHostToDevice, s));
  PD_HIP(hipMemcpyAsync(D.y[0], y4.data(), sizeof(double4) * N, hipMemcpyHostToDevice, s));
  PD_HIP(hipMemcpyAsync(D.y[1], y4.data(), sizeof(double4) * N, hipMemcpyHostToDevice, s));
  PD_HIP(hipMemcpyAsync(D.v, v4.data(), sizeof(double4) * N, hipMemcpyHostToDevice, s));
  PD_HIP(hipMemsetAsync(D.f, 0, sizeof(double4) * N, s));
  PD_HIP(hipMemcpyAsync(D.pblock, pblock.data(), sizeof(int32_t) * N, hipMemcpyHostToDevice, s));
  PD_HIP(hipMemcpyAsync(D.ptype, ptype.data(), N, hipMemcpyHostToDevice, s));
  PD_HIP(hipMemcpyAsync(D.inv_den, inv_den.data(), sizeof(double) * N, hipMemcpyHostToDevice, s));
  PD_HIP(hipMemcpyAsync(D.bcid, bcid.data(), sizeof(int32_t) * N, hipMemcpyHostToDevice, s));
  if (any_onlyatstart) {
    // minimize!'s held particles: those of any BC without onlyatstart (minimize.jl:138-143)
    auto held = hs.take<uint8_t>(N);
    std::fill(held.p, held.p + N, (uint8_t)0);
    for (const HostBc& hb : e.bcs)
      if (!(hb.d.options & PD_BC_ONLYATSTART))
        for (int64_t r = 0; r < N; ++r)
          if (hb.mask[r]) held[internal_of(e, r)] = 1;
    D.qheld = dalloc<uint8_t>(e, N);
    PD_HIP(hipMemcpyAsync(D.qheld, held.data(), N, hipMemcpyHostToDevice, s));
  }
  PD_HIP(hipMemsetAsync(D.theta, 0, sizeof(double) * N, s));
  PD_HIP(hipMemsetAsync(D.td, 0, sizeof(double) * N, s));
  PD_HIP(hipMemsetAsync(D.f, 0, sizeof(double4) * N, s));

  // ---- family: degrees
  std::vector<FamTmp> tmp(e.blocks.size());
  for (auto& t : tmp) t.dm = &c->mem;
  auto deg = hs.take<int32_t>(N);
  try {
    for (size_t bi = 0; bi < e.blocks.size(); ++bi) {
      HostBlock& b = e.blocks[bi];
      build_cells(c, e, b, tmp[bi]);
      FamTmp& t = tmp[bi];
      pd_launch_degree(s, D.x, b.first, (int)b.d.n, t.g, b.d.horizon, t.cell_of, t.cell_start, t.cell_count,
                       t.items, D.deg);
    }
    PD_HIP(hipMemcpyAsync(deg.data(), D.deg, sizeof(int32_t) * N, hipMemcpyDeviceToHost, s));
    PD_HIP(hipStreamSynchronize(s));
    // ---- SELL-64 layout
    auto pslot0 = hs.take<uint32_t>(N);
    uint64_t slots = 0;
    e.chunks.clear();
    e.npk = 0;
    e.nbonds = 0;
    for (size_t bi = 0; bi < e.blocks.size(); ++bi) {
      HostBlock& b = e.blocks[bi];
      int32_t maxdeg = 0;
      for (int64_t i = 0; i < b.d.n; ++i) maxdeg = std::max(maxdeg, deg[b.first + i]);
      b.maxdeg = maxdeg;
      // cal_family (peri.jl:571-576) keeps M = min(max_neigh, n) rows; filling
      // beyond M or maxdeg == M (general_material.jl:88-89) is a BoundsError
      const int64_t M = std::min<int64_t>(b.d.max_neigh, b.d.n);
      if (maxdeg >= M && maxdeg > 0)
        throw PdError(PD_E_OVERFLOW, "BoundsError: block " + std::to_string(bi) + " has a particle with " +
                                         std::to_string(maxdeg) + " neighbours; max_neigh=" +
                                         std::to_string(b.d.max_neigh) + " must exceed the maximum degree");
      b.chunk0 = (int32_t)e.chunks.size();
      const int G = 1 << e.lgG, Cp = PD_WAVE >> e.lgG;
      for (int64_t c0 = 0; c0 < b.d.n;) {
        int64_t cnt = std::min<int64_t>(Cp, b.d.n - c0);
        if (e.want_tiles) {  // a chunk never spans two bricks
          const int32_t br = e.brick[b.first + c0];
          int64_t k = 1;
          while (k < cnt && e.brick[b.first + c0 + k] == br) ++k;
          cnt = k;
        }
        Chunk ch{};
        ch.env = 0;
        ch.block = (int32_t)bi;
        ch.first = (int32_t)(b.first + c0);
        ch.count = (int32_t)cnt;
        c0 += cnt;
        int32_t md = 0;
        for (int l = 0; l < ch.count; ++l) md = std::max(md, deg[ch.first + l]);
        const int32_t wdt = (md + G - 1) / G;  // slots per lane
        ch.width = wdt;
        ch.slot0 = (uint32_t)slots;
        for (int l = 0; l < ch.count; ++l) {
          pslot0[ch.first + l] = (uint32_t)(slots + l);
          e.nbonds += deg[ch.first + l];
        }
        slots += (uint64_t)PD_WAVE * wdt;
        if (slots >= 0xfffffff0ULL) throw PdError(PD_E_ARG, "env too large (bond slots exceed 2^32)");
        ch.pk0 = (uint32_t)e.npk;
        if (e.want_tiles) e.npk += (int64_t)PD_WAVE * ((wdt + 3) / 4);
        e.chunks.push_back(ch);
      }
      b.nchunks = (int32_t)e.chunks.size() - b.chunk0;
      if (b.d.kind == PD_EP) e.has_ep = true;
    }
    e.nslots = (int64_t)slots;
    // kernels address every array with 32-bit byte offsets
    if ((uint64_t)slots * (e.has_ep ? sizeof(double) : sizeof(int32_t)) >= PD_MAX_ARRAY_BYTES)
      throw PdError(PD_E_ARG, "env too large: its bond arrays would exceed 4 GiB; split it into several envs");
    if ((uint64_t)e.npk * sizeof(uint64_t) >= PD_MAX_ARRAY_BYTES) e.want_tiles = false;
    PD_HIP(hipMemcpyAsync(D.pslot0, pslot0.data(), sizeof(uint32_t) * N, hipMemcpyHostToDevice, s));
    D.col = dalloc<int32_t>(e, slots);
    D.state = dalloc<uint8_t>(e, slots);
    // reference index of every internal particle: rows are ordered by it
    auto refh = hs.take<int32_t>(N);
    for (int64_t p = 0; p < N; ++p) refh[p] = (int32_t)reference_of(e, p);
    int32_t* d_ref = dalloc<int32_t>(e, N);
    PD_HIP(hipMemcpyAsync(d_ref, refh.data(), sizeof(int32_t) * N, hipMemcpyHostToDevice, s));
    PD_HIP(hipMemsetAsync(D.col, 0xff, sizeof(int32_t) * std::max<uint64_t>(slots, 1), s));
    PD_HIP(hipMemsetAsync(D.state, 0, std::max<uint64_t>(slots, 1), s));
    for (size_t bi = 0; bi < e.blocks.size(); ++bi) {
      HostBlock& b = e.blocks[bi];
      FamTmp& t = tmp[bi];
      pd_launch_fill(s, D.x, b.first, (int)b.d.n, t.g, b.d.horizon, t.cell_of, t.cell_start, t.cell_count,
                     t.items, D.pslot0, e.lgG, d_ref, D.col, D.state, b.maxdeg);
      pd_launch_wvol(s, D.x, b.first, (int)b.d.n, D.pslot0, e.lgG, D.deg, D.col, D.m);
    }
    D.aux = dalloc<double4>(e, N);
    pd_launch_aux_init(s, D.m, D.aux, (int)N);
    if (e.has_ep) {
      D.rev = dalloc<uint32_t>(e, slots);
      D.edp[0] = dalloc<double>(e, slots);
      D.edp[1] = dalloc<double>(e, slots);
      PD_HIP(hipMemsetAsync(D.edp[0], 0, sizeof(double) * std::max<uint64_t>(slots, 1), s));
      PD_HIP(hipMemsetAsync(D.edp[1], 0, sizeof(double) * std::max<uint64_t>(slots, 1), s));
      for (auto& b : e.blocks)
        pd_launch_rev(s, b.first, (int)b.d.n, D.pslot0, e.lgG, d_ref, D.deg, D.col, D.rev);
    }
    if (e.want_tiles) build_tiles(c, e);
    if (e.precision == PD_PREC_FP32C) {
      // x32 relative to each block's minimum corner (bonds never leave a block)
      auto x32 = hs.take<float4>(N), a32 = hs.take<float4>(N);
      auto mh = hs.take<double>(N);
      PD_HIP(hipMemcpyAsync(mh.data(), D.m, sizeof(double) * N, hipMemcpyDeviceToHost, s));
      PD_HIP(hipStreamSynchronize(s));
      for (auto& b : e.blocks) {
        double lo[3] = {INFINITY, INFINITY, INFINITY};
        for (int64_t i = 0; i < b.d.n; ++i)
          for (int d = 0; d < 3; ++d) lo[d] = std::min(lo[d], b.x[3 * i + d]);
        for (int64_t i = 0; i < b.d.n; ++i) {
          const int64_t p = internal_of(e, b.first + i);
          x32[p] = make_float4((float)(b.x[3 * i] - lo[0]), (float)(b.x[3 * i + 1] - lo[1]),
                               (float)(b.x[3 * i + 2] - lo[2]), (float)b.vol[i]);
        }
      }
      for (int64_t p = 0; p < N; ++p) a32[p] = make_float4(0.0f, (float)(1.0 / mh[p]), 0.0f, 0.0f);
      D.x32 = dalloc<float4>(e, N);
      D.aux32 = dalloc<float4>(e, N);
      PD_HIP(hipMemcpyAsync(D.x32, x32.data(), sizeof(float4) * N, hipMemcpyHostToDevice, s));
      PD_HIP(hipMemcpyAsync(D.aux32, a32.data(), sizeof(float4) * N, hipMemcpyHostToDevice, s));
      PD_HIP(hipStreamSynchronize(s));
    }
    PD_HIP(hipStreamSynchronize(s));
    for (auto& t : tmp) fam_free(t);
  } catch (...) {
    (void)hipStreamSynchronize(s);
    for (auto& t : tmp) fam_free(t);
    throw;
  }
  D.has_ep = e.has_ep ? 1 : 0;
  D.lgG = e.lgG;
  D.tiled = e.tiled ? 1 : 0;
  // tiled FP64 passes may stage x and volume as FP32 when that is exact
  e.xf32 = e.tiled;
  for (auto& b : e.blocks)
    for (int64_t i = 0; e.xf32 && i < b.d.n; ++i) {
      for (int d = 0; d < 3; ++d) e.xf32 = e.xf32 && (double)(float)b.x[3 * i + d] == b.x[3 * i + d];
      e.xf32 = e.xf32 && (double)(float)b.vol[i] == b.vol[i];
    }
  D.xf32 = e.xf32 ? 1 : 0;

  // ---- blocks
  std::vector<BlockDev> bd(e.blocks.size());
  bool undeformed = false;
  for (size_t bi = 0; bi < e.blocks.size(); ++bi) {
    HostBlock& b = e.blocks[bi];
    BlockDev& B = bd[bi];
    std::memset(&B, 0, sizeof(B));
    const int nt = b.d.ntype;
    B.kind = b.d.kind;
    B.ntype = nt;
    B.first = b.first;
    B.n = (int32_t)b.d.n;
    B.chunk0 = b.chunk0;
    B.nchunks = b.nchunks;
    B.deformed = b.d.deformed ? 1 : 0;
    if (!B.deformed) undeformed = true;
    B.c11 = (b.d.kind == PD_BB) ? b.c[0] : 0.0;
    B.horizon = b.d.horizon;
    // FP32C: origin of x32 (the block's minimum corner) and s0 rounded up to FP32
    for (int d = 0; d < 3; ++d) B.corner[d] = INFINITY;
    for (int64_t i = 0; i < b.d.n; ++i)
      for (int d = 0; d < 3; ++d) B.corner[d] = std::min(B.corner[d], b.x[3 * i + d]);
    // integer lattice: bond lengths from the tiled kernels' table
    {
      bool xi = b.d.horizon * b.d.horizon <= 32.0;
      for (int64_t i = 0; xi && i < 3 * b.d.n; ++i)
        xi = std::floor(b.x[i]) == b.x[i] && std::fabs(b.x[i]) < 4194304.0;
      B.xint = xi ? 1 : 0;
      b.xint = xi;
    }
    B.vol0 = b.d.n > 0 ? b.vol[0] : 0.0;
    B.vol1 = 1;
    for (int64_t i = 0; i < b.d.n; ++i) B.vol1 = B.vol1 && b.vol[i] == B.vol0;
    b.vol1 = B.vol1 != 0;
    {
      float f = (float)b.s0[0];
      if ((double)f < b.s0[0]) f = std::nextafter(f, INFINITY);
      B.s0f_up = f;
    }
    for (int t = 0; t < nt; ++t) B.density[t] = b.density[t];
    for (int q = 0; q < nt * nt; ++q) {
      B.s0[q] = b.s0[q];
      B.K[q] = b.K[q];
      B.G[q] = b.G[q];
      B.psi[q] = b.psi[q];
    }
  }
  e.maybe_undeformed = undeformed;
  D.blocks = dalloc<BlockDev>(e, bd.size());
  PD_HIP(hipMemcpyAsync(D.blocks, bd.data(), sizeof(BlockDev) * bd.size(), hipMemcpyHostToDevice, s));

  // ---- BCs
  std::vector<BcDev> bcd(e.bcs.size());
  for (size_t k = 0; k < e.bcs.size(); ++k) {
    std::memset(&bcd[k], 0, sizeof(BcDev));
    bcd[k].kind = e.bcs[k].d.kind;
    for (int d = 0; d < 3; ++d) bcd[k].rate[d] = e.bcs[k].d.rate[d];
    bcd[k].freq = e.bcs[k].d.freq;
    bcd[k].applyafter = e.bcs[k].d.applyafter;
  }
  D.bcs = dalloc<BcDev>(e, std::max<size_t>(bcd.size(), 1));
  if (!bcd.empty())
    PD_HIP(hipMemcpyAsync(D.bcs, bcd.data(), sizeof(BcDev) * bcd.size(), hipMemcpyHostToDevice, s));

  // ---- contacts
  e.cdev.assign(e.contacts.size(), ContactDev{});
  std::vector<std::vector<int32_t>> keep;  // staging kept alive until sync
  for (size_t k = 0; k < e.contacts.size(); ++k) {
    HostContact& hc = e.contacts[k];
    ContactDev& cd = e.cdev[k];
    std::memset(&cd, 0, sizeof(cd));
    std::vector<int32_t> idx1, idx2, inv1(N, -1), inv2(N, -1);
    std::vector<int64_t> types(N);
    for (auto& b : e.blocks)
      for (int64_t i = 0; i < b.d.n; ++i) types[b.first + i] = b.type[i];
    // member ids follow the reference order (the type masks of contacts.jl:212-219);
    // idx maps a member to its internal particle id
    const bool self = hc.t2.empty();  // ContactModel11 (contacts.jl:105-119)
    for (int64_t r = 0; r < N; ++r) {
      const bool in1 = std::find(hc.t1.begin(), hc.t1.end(), types[r]) != hc.t1.end();
      const bool in2 = std::find(hc.t2.begin(), hc.t2.end(), types[r]) != hc.t2.end();
      if (in1 && in2) throw PdError(PD_E_ARG, "contact pair type sets must be disjoint");
      const int32_t p = (int32_t)internal_of(e, r);
      if (in1) { inv1[p] = (int32_t)idx1.size(); idx1.push_back(p); }
      if (in2) { inv2[p] = (int32_t)idx2.size(); idx2.push_back(p); }
    }
    if (self) {
      // one material block, members in the block's own order: short_range_contact!
      // and update_contact_neighs! of ContactModel11 index the particles of the
      // block with its family (contacts.jl:361-365, 574-591)
      int64_t nb_match = 0;
      for (auto& b : e.blocks) {
        int64_t in = 0;
        for (int64_t i = 0; i < b.d.n; ++i)
          in += std::find(hc.t1.begin(), hc.t1.end(), b.type[i]) != hc.t1.end();
        if (in == b.d.n && in == (int64_t)idx1.size()) ++nb_match;
      }
      if (nb_match != 1) throw PdError(PD_E_ARG, "a self contact must cover exactly the particles of one block");
      idx2 = idx1;
      inv2 = inv1;
    }
    cd.self_contact = self ? 1 : 0;
    cd.n1 = (int32_t)idx1.size();
    cd.n2 = (int32_t)idx2.size();
    cd.max_neighs = (int32_t)std::max<int64_t>(1, std::min<int64_t>(hc.d.max_neighs, cd.n2));
    cd.expo_is_int = hc.d.exponent_is_int;
    cd.expo = hc.d.exponent;
    cd.K = hc.d.stiffness;
    cd.equi = hc.d.equi_dist;
    cd.distance = hc.d.distance;
    cd.ncell_cap = 2 * cd.n2 + 64;
    cd.rev_cap = cd.n1 * cd.max_neighs;
    int32_t* didx1 = dalloc<int32_t>(e, cd.n1);
    int32_t* didx2 = dalloc<int32_t>(e, cd.n2);
    int32_t* dinv1 = dalloc<int32_t>(e, N);
    int32_t* dinv2 = dalloc<int32_t>(e, N);
    cd.idx1 = didx1; cd.idx2 = didx2; cd.inv1 = dinv1; cd.inv2 = dinv2;
    cd.own_cnt = dalloc<int32_t>(e, cd.n1);
    cd.own = dalloc<int32_t>(e, (size_t)cd.n1 * cd.max_neighs);
    cd.rev_cnt = dalloc<int32_t>(e, cd.n2);
    cd.rev_off = dalloc<int32_t>(e, cd.n2);
    cd.rev = dalloc<int32_t>(e, (size_t)cd.rev_cap);
    cd.rev_cursor = dalloc<int32_t>(e, cd.n2);
    cd.cell_start = dalloc<int32_t>(e, 2 * (size_t)cd.ncell_cap);
    cd.cell_count = dalloc<int32_t>(e, cd.ncell_cap);
    cd.cell_items = dalloc<int32_t>(e, cd.n2);
    cd.cell_of = dalloc<int32_t>(e, cd.n2);
    cd.wlist = dalloc<int32_t>(e, std::max(cd.n1, cd.n2));
    cd.wcount = dalloc<int32_t>(e, 2);
    // reduction partials + scan scratch (pd_contact.hip)
    cd.box = dalloc<double>(e, 16 + 128 * 12 + pd_scan_scratch_ints(std::max(cd.ncell_cap, cd.n2)) / 2 + 8);
    PD_HIP(hipMemcpyAsync(didx1, idx1.data(), sizeof(int32_t) * idx1.size(), hipMemcpyHostToDevice, s));
    PD_HIP(hipMemcpyAsync(didx2, idx2.data(), sizeof(int32_t) * idx2.size(), hipMemcpyHostToDevice, s));
    PD_HIP(hipMemcpyAsync(dinv1, inv1.data(), sizeof(int32_t) * N, hipMemcpyHostToDevice, s));
    PD_HIP(hipMemcpyAsync(dinv2, inv2.data(), sizeof(int32_t) * N, hipMemcpyHostToDevice, s));
    PD_HIP(hipMemsetAsync(cd.own_cnt, 0, sizeof(int32_t) * std::max(cd.n1, 1), s));
    PD_HIP(hipMemsetAsync(cd.rev_cnt, 0, sizeof(int32_t) * std::max(cd.n2, 1), s));
    PD_HIP(hipMemsetAsync(cd.rev_off, 0, sizeof(int32_t) * std::max(cd.n2, 1), s));
    PD_HIP(hipStreamSynchronize(s));
  }
  if (!e.cdev.empty()) {
    D.contacts = dalloc<ContactDev>(e, e.cdev.size());
    PD_HIP(hipMemcpyAsync(D.contacts, e.cdev.data(), sizeof(ContactDev) * e.cdev.size(), hipMemcpyHostToDevice, s));
  }
  ensure_slot(c, e);
  for (auto& ch : e.chunks) ch.env = e.slot;
  for (auto& t : e.tiles) {
    t.env = e.slot;
    for (int k = 0; k < t.nch; ++k) t.ch[k].env = e.slot;
  }
  e.d_self = dalloc<int32_t>(e, 1);
  PD_HIP(hipMemcpyAsync(e.d_self, &e.slot, sizeof(int32_t), hipMemcpyHostToDevice, s));
  if (e.tiled && !e.contacts.empty()) {
    // particles with contact terms, numbered at every list rebuild
    D.cflag = dalloc<uint8_t>(e, N);
    D.cpart = dalloc<int32_t>(e, N);
    D.ccount = dalloc<int32_t>(e, 1);
    if (e.cdev.size() == 1 && !e.cdev[0].self_contact) D.crec = dalloc<int32_t>(e, 8 * (size_t)N);
    PD_HIP(hipMemsetAsync(D.cflag, 0, (size_t)N, s));
    PD_HIP(hipMemsetAsync(D.ccount, 0, sizeof(int32_t), s));
    e.cself_host[0] = e.slot;
    e.cself_host[1] = 0;
    e.d_cself = dalloc<int32_t>(e, 2);
    PD_HIP(hipMemcpyAsync(e.d_cself, e.cself_host, 2 * sizeof(int32_t), hipMemcpyHostToDevice, s));
  }
  e.d_chunks = dalloc<Chunk>(e, e.chunks.size());
  PD_HIP(hipMemcpyAsync(e.d_chunks, e.chunks.data(), sizeof(Chunk) * e.chunks.size(), hipMemcpyHostToDevice, s));
  e.tile_blocks.clear();
  if (e.tiled && !e.tiles.empty()) {
    e.d_tiles = dalloc<Tile>(e, e.tiles.size());
    PD_HIP(hipMemcpyAsync(e.d_tiles, e.tiles.data(), sizeof(Tile) * e.tiles.size(), hipMemcpyHostToDevice, s));
    for (int t = 0; t < (int)e.tiles.size(); ++t) {
      if (e.tile_blocks.empty() || e.tile_blocks.back()[0] != e.tiles[t].block)
        e.tile_blocks.push_back({e.tiles[t].block, t, 0});
      ++e.tile_blocks.back()[2];
    }
  }
  upload_env_table(c, e);
  PD_HIP(hipStreamSynchronize(s));
  e.committed = true;
  c->set_valid = false;
}

// ---- run sets ----------------------------------------------------------------
template <typename T>
void ensure_cap(T*& p, size_t& cap, size_t n, const DevMem& dm) {
  if (n <= cap && p) return;
  if (p) PD_HIP(hipFree(p));
  p = nullptr;
  cap = std::max<size_t>(n, 16) * 2;
  PD_HIP(hipMalloc(&p, sizeof(T) * cap));
  if (dm.poison) PD_HIP(hipMemsetAsync(p, 0xff, sizeof(T) * cap, dm.stream));
}

bool env_gpow(const Env& e) {
  for (const ContactDev& cd : e.cdev)
    if (!cd.expo_is_int) return true;
  return false;
}

// A run of an env's tiles (first, n) placed at dst in the set's tile array.
struct TileRun {
  int dst;
  const Env* e;
  int first, n;
};

// Relative cost of a tile's workgroup: the waves of its chunks run side by
// side, each through its chunk's bond steps, after the halo staging.
int tile_cost(const Tile& T) {
  int w = 0;
  for (int k = 0; k < T.nch; ++k) w = std::max(w, T.ch[k].width);
  return 4 * w + T.hn / 16;
}

// PERIDYN_TILE_ORDER=0 keeps every launch's tiles in env and brick order,
// =1 orders every launch by cost; unset, a launch of fewer than kOrderTiles
// tiles is ordered by cost where that shortens its predicted span by
// kOrderGain percent or more (below)
int tile_order_mode() {
  const char* v = std::getenv("PERIDYN_TILE_ORDER");
  if (!v) return -1;
  return std::string(v) == "0" ? 0 : 1;
}
constexpr int kOrderTiles = 16384, kOrderGain = 4, kOrderRun = 16;

// Predicted span of a launch whose tiles (costs in list order) are dispatched
// as pd_xcd_block does: XCD x runs its contiguous range of the list in order
// on its 32 CUs' wg workgroup slots, taken here as rounds of that many tiles,
// each round lasting as long as its costliest tile (linear in the tiles, as
// the set build runs before every served chunk).
int64_t tile_span(const std::vector<int>& cost, int wg) {
  const int n = (int)cost.size(), q = n / 8, r = n % 8;
  const int slots = 32 * std::max(wg, 1);
  int64_t span = 0;
  int at = 0;
  for (int x = 0; x < 8; ++x) {
    const int m = q + (x < r ? 1 : 0);
    int64_t t = 0;
    for (int i0 = at; i0 < at + m; i0 += slots) {
      int c = 0;
      for (int i = i0; i < std::min(at + m, i0 + slots); ++i) c = std::max(c, cost[i]);
      t += c;
    }
    span = std::max(span, t);
    at += m;
  }
  return span;
}

// Copies of the tile runs into the set's tile array, one launch range of the
// lists at a time.  In brick order, a launch's workgroups go to the chip's 8
// XCDs in 8 contiguous ranges of the list (pd_xcd_block), so neighbouring
// bricks share their XCD's L2.  Ordered by cost, the tiles are sorted by
// decreasing cost (stable) and dealt round-robin over the 8 ranges, so that
// each XCD receives the same mix of costly and cheap tiles and runs its own
// in decreasing cost: the last workgroups of a launch are the cheapest ones.
// That pays where the tiles' costs are unequal and the launch runs few
// rounds of workgroups (the disk's partial bricks: force pass 129 -> 117 us;
// specimens of different horizons stepped together: 22 served requests, 3%
// faster per step).  Dealt one tile at a time, neighbouring bricks' shared
// halo records miss the XCD's L2 (the plates' near-equal tiles: force pass
// 10% slower), so the deal hands out runs of kOrderRun consecutive ranks
// (brick order within equal costs: the plates then 1% faster, runs of 2: 5%
// slower).  A launch is ordered by cost where tile_span predicts that order
// to be kOrderGain percent shorter, and only below kOrderTiles tiles: a
// longer launch runs many rounds of workgroups and has little tail to gain
// (the 256 tensile specimens, 33k tiles per launch: 6% slower dealt one tile
// at a time).
void place_tile_runs(Ctx* c, std::vector<TileRun>& runs, std::vector<PdCopy>& copies,
                     Ctx::TileList (&l_th)[2][4][PD_TILE_CLASSES], Ctx::TileList (&l_td)[4][PD_TILE_CLASSES],
                     Ctx::TileList (&l_fk)[3][2][4][PD_TILE_CLASSES], int xfl) {
  const int mode = tile_order_mode();
  const char* lg_env = std::getenv("PERIDYN_TILE_ORDER_LOG");
  const bool log = lg_env && std::string(lg_env) == "1";
  auto run_copy = [&](const Env* e, int t, int n, int dst) {
    copies.push_back(PdCopy{reinterpret_cast<const uint32_t*>(e->d_tiles + t),
                            reinterpret_cast<uint32_t*>(c->d_tset + dst), (uint32_t)(sizeof(Tile) * n / 4), 0u});
  };
  if (mode == 0) {
    for (const TileRun& r : runs) run_copy(r.e, r.first, r.n, r.dst);
    return;
  }
  std::sort(runs.begin(), runs.end(), [](const TileRun& a, const TileRun& b) { return a.dst < b.dst; });
  std::vector<std::pair<const Env*, int>> tl;
  std::vector<int> cost, ord, arranged;
  size_t ri = 0;
  // rec: LDS bytes per halo record of the launch's kernel (its workgroups per CU)
  auto launch = [&](const Ctx::TileList& l, int rec) {
    if (l.n <= 0) return;
    const size_t r0 = ri;
    auto in_order = [&] {
      for (ri = r0; ri < runs.size() && runs[ri].dst < l.off + l.n; ++ri)
        if (runs[ri].dst >= l.off) run_copy(runs[ri].e, runs[ri].first, runs[ri].n, runs[ri].dst);
    };
    if (mode < 0 && l.n >= kOrderTiles) return in_order();
    tl.clear();
    while (ri < runs.size() && runs[ri].dst < l.off) ++ri;
    for (; ri < runs.size() && runs[ri].dst < l.off + l.n; ++ri)
      for (int t = 0; t < runs[ri].n; ++t) tl.push_back({runs[ri].e, runs[ri].first + t});
    const int n = (int)tl.size();
    if (n != l.n) throw PdError(PD_E_STATE, "tile runs do not cover their launch");
    cost.resize(n);
    ord.resize(n);
    arranged.resize(n);
    for (int i = 0; i < n; ++i) {
      cost[i] = tile_cost(tl[i].first->tiles[tl[i].second]);
      ord[i] = i;
    }
    {  // stable counting sort by decreasing cost (costs are small integers)
      int cmax = 0;
      for (int i = 0; i < n; ++i) cmax = std::max(cmax, cost[i]);
      std::vector<int> at(cmax + 2, 0);
      for (int i = 0; i < n; ++i) ++at[cmax - cost[i] + 1];
      for (int k = 1; k <= cmax + 1; ++k) at[k] += at[k - 1];
      for (int i = 0; i < n; ++i) ord[at[cmax - cost[i]]++] = i;
    }
    // rank b -> list position: runs of kOrderRun consecutive ranks (often
    // neighbouring bricks of equal cost, which share halo records) are dealt
    // round-robin over the XCD ranges, the ranks past the last whole round of
    // runs one at a time, so that XCD x still receives q (+1 if x < r) tiles
    const int q = n / 8, r = n % 8, whole = n / (8 * kOrderRun) * (8 * kOrderRun);
    auto pos_of = [&](int b) {
      int x, i;
      if (b < whole) {
        const int run = b / kOrderRun;
        x = run % 8;
        i = run / 8 * kOrderRun + b % kOrderRun;
      } else {
        x = (b - whole) % 8;
        i = whole / 8 + (b - whole) / 8;
      }
      return x * q + std::min(x, r) + i;
    };
    for (int b = 0; b < n; ++b) arranged[pos_of(b)] = cost[ord[b]];
    const int wg = (int)std::min<int64_t>(8, std::max<int64_t>(1, 163840 / ((int64_t)rec * std::max(l.hcap, 1))));
    const int64_t s_brick = tile_span(cost, wg), s_cost = tile_span(arranged, wg);
    const bool by_cost = mode == 1 || (mode < 0 && s_cost * 100 <= s_brick * (100 - kOrderGain));
    if (log)
      std::fprintf(stderr, "tile launch at %d: %d tiles, %d per CU, span %lld in brick order, %lld by cost: %s\n",
                   l.off, n, wg, (long long)s_brick, (long long)s_cost, by_cost ? "by cost" : "brick order");
    if (!by_cost) return in_order();
    // one copy per run of ranks whose tiles follow each other both in their
    // env's array and in the list
    for (int b = 0; b < n;) {
      const Env* e = tl[ord[b]].first;
      const int t0 = tl[ord[b]].second, p0 = pos_of(b);
      int k = 1;
      while (b + k < n && tl[ord[b + k]].first == e && tl[ord[b + k]].second == t0 + k && pos_of(b + k) == p0 + k)
        ++k;
      run_copy(e, t0, k, l.off + p0);
      b += k;
    }
  };
  // launch ranges in ascending set offset (the order place() gave them)
  std::vector<std::pair<const Ctx::TileList*, int>> ls;
  for (int lg = 1; lg <= 3; ++lg) {
    for (int pc = 0; pc < 2; ++pc) {
      const int prec = pc ? PD_PREC_FP32C : PD_PREC_FP64;
      for (int h = 0; h < PD_TILE_CLASSES; ++h)
        ls.push_back({&l_th[pc][lg][h], pd_tile_record_bytes(0, PD_OSB, prec, xfl)});
      for (int k = 0; k < 3; ++k)
        for (int h = 0; h < PD_TILE_CLASSES; ++h)
          ls.push_back({&l_fk[k][pc][lg][h], pd_tile_record_bytes(2, k, prec, xfl)});
    }
    for (int h = 0; h < PD_TILE_CLASSES; ++h)
      ls.push_back({&l_td[lg][h], pd_tile_record_bytes(1, PD_EP, PD_PREC_FP64, xfl)});
  }
  std::sort(ls.begin(), ls.end(), [](const auto& a, const auto& b) { return a.first->off < b.first->off; });
  for (const auto& l : ls) launch(*l.first, l.second);
}

// Run set of the listed envs.  The all-chunk list and the tile lists are
// gathered on the device from each env's own arrays (no host staging); the
// untiled envs' stepping lists and, with full (QSDrag's passes), the all-chunk
// lists per pass are built on the host.
void build_set(Ctx* c, const std::vector<int64_t>& ids, bool full = false) {
  if (c->set_valid && ids == c->set_ids && (c->set_full || !full)) return;
  static_assert(sizeof(Chunk) % 4 == 0 && sizeof(Tile) % 4 == 0, "copied as 4-byte words");
  std::vector<PdCopy> copies;  // per-env chunk and tile lists into the set's arrays
  std::vector<TileRun> truns;  // per-env tile runs at their list positions
  std::vector<int2> citems;    // pair-contact items
  std::vector<int32_t> cfin, cfin_starts;  // tiled envs with contacts
  std::vector<int32_t> cst;    // their per-class workgroup offsets
  std::vector<Chunk> th, td, fk[3], thu, tdu, fku[3];
  std::vector<int32_t> slots;
  std::vector<int2> ranges;
  bool has_ep = false, has_th = false, gpow = false, tiled = false, xf = true;
  int n_all = 0, n_th = 0, n_td = 0, n_fk[3] = {0, 0, 0};
  // tile lists: theta [pc][lg][cls], td [lg][cls], force [kind][pc][lg][cls]
  using TileList = Ctx::TileList;
  TileList l_th[2][4][PD_TILE_CLASSES], l_td[4][PD_TILE_CLASSES], l_fk[3][2][4][PD_TILE_CLASSES];
  bool lat = true;   // every tiled FP64 bond-based block of the set is an integer lattice
  bool onev = true;  // ... and has one volume
  for (int64_t id : ids) {
    const Env& e = get_env(c, id);
    if (e.tiled) xf = xf && e.xf32;  // the FP64 records' size, hence the LDS classes
    if (e.tiled && e.precision == PD_PREC_FP64)
      for (auto& b : e.blocks)
        if (b.d.kind == PD_BB) {
          lat = lat && b.xint;
          onev = onev && b.vol1;
        }
  }
  const int xfl = xf ? (lat ? (onev ? 3 : 2) : 1) : 0;
  auto add = [](TileList& l, int n, int hmax) {
    l.n += n;
    l.hcap = std::max(l.hcap, hmax);
  };
  for (int64_t id : ids) {
    Env& e = get_env(c, id);
    if (!e.committed) throw PdError(PD_E_STATE, "env " + std::to_string(id) + " is not committed");
    slots.push_back(e.slot);
    gpow = gpow || env_gpow(e);
    ranges.push_back(make_int2(n_all, (int)e.chunks.size()));
    n_all += (int)e.chunks.size();
    for (auto& b : e.blocks) {
      const int kind = b.d.kind;
      n_fk[kind] += b.nchunks;
      if (kind != PD_BB) { n_th += b.nchunks; has_th = true; }
      if (kind == PD_EP) { n_td += b.nchunks; has_ep = true; }
    }
    if (full || !e.tiled)
      for (const Chunk& ch : e.chunks) {
        const int kind = e.blocks[ch.block].d.kind;
        if (full) {
          fk[kind].push_back(ch);
          if (kind != PD_BB) th.push_back(ch);
          if (kind == PD_EP) td.push_back(ch);
        }
        if (!e.tiled) {
          fku[kind].push_back(ch);
          if (kind != PD_BB) thu.push_back(ch);
          if (kind == PD_EP) tdu.push_back(ch);
        }
      }
    if (e.tiled) {
      tiled = true;
      const int hm = (int)e.hmax, lg = e.lgG, pr = e.precision;
      const int pc = (pr == PD_PREC_FP32C) ? 1 : 0;  // FP32C envs tile BB / OSB blocks only
      for (auto& tb : e.tile_blocks) {
        const int kind = e.blocks[tb[0]].d.kind;
        add(l_fk[kind][pc][lg][pd_tile_class(2, kind, pr, xfl, hm)], tb[2], hm);
        if (kind != PD_BB) add(l_th[pc][lg][pd_tile_class(0, kind, pr, xfl, hm)], tb[2], hm);
        if (kind == PD_EP) add(l_td[lg][pd_tile_class(1, kind, pr, xfl, hm)], tb[2], hm);
      }
    }
  }
  // chunk set: [all | th td fk (full) | thu tdu fku]
  std::vector<Chunk> rest;
  c->off_theta = n_all + (int)rest.size(); rest.insert(rest.end(), th.begin(), th.end());
  c->off_td = n_all + (int)rest.size(); rest.insert(rest.end(), td.begin(), td.end());
  for (int k = 0; k < 3; ++k) { c->off_force[k] = n_all + (int)rest.size(); rest.insert(rest.end(), fk[k].begin(), fk[k].end()); }
  c->off_theta_u = n_all + (int)rest.size(); rest.insert(rest.end(), thu.begin(), thu.end());
  c->off_td_u = n_all + (int)rest.size(); rest.insert(rest.end(), tdu.begin(), tdu.end());
  for (int k = 0; k < 3; ++k) {
    c->off_force_u[k] = n_all + (int)rest.size();
    rest.insert(rest.end(), fku[k].begin(), fku[k].end());
  }
  c->n_all = n_all;
  c->n_theta = n_th;
  c->n_td = n_td;
  for (int k = 0; k < 3; ++k) c->n_force[k] = n_fk[k];
  c->n_theta_u = (int)thu.size();
  c->n_td_u = (int)tdu.size();
  for (int k = 0; k < 3; ++k) c->n_force_u[k] = (int)fku[k].size();
  ensure_cap(c->d_set, c->set_cap, (size_t)n_all + rest.size(), c->mem);
  {
    int off = 0;
    for (int64_t id : ids) {
      Env& e = get_env(c, id);
      if (!e.chunks.empty())
        copies.push_back(PdCopy{reinterpret_cast<const uint32_t*>(e.d_chunks),
                                reinterpret_cast<uint32_t*>(c->d_set + off),
                                (uint32_t)(sizeof(Chunk) * e.chunks.size() / 4), 0u});
      off += (int)e.chunks.size();
    }
  }
  if (!rest.empty())
    PD_HIP(hipMemcpyAsync(c->d_set + n_all, rest.data(), sizeof(Chunk) * rest.size(), hipMemcpyHostToDevice,
                          c->stream));
  // tile set: per lgG, form and pass the lists of the LDS classes in class
  // order, so that neighbouring classes can share a launch
  int tot = 0;
  auto place = [&](TileList* l) {
    for (int h = 0; h < PD_TILE_CLASSES; ++h) {
      l[h].off = tot;
      tot += l[h].n;
    }
  };
  for (int lg = 1; lg <= 3; ++lg) {
    for (int pc = 0; pc < 2; ++pc) {
      place(l_th[pc][lg]);
      for (int k = 0; k < 3; ++k) place(l_fk[k][pc][lg]);
    }
    place(l_td[lg]);
  }
  if (tot > 0) {
    ensure_cap(c->d_tset, c->tset_cap, (size_t)tot, c->mem);
    // fill positions
    int f_th[2][4][PD_TILE_CLASSES] = {}, f_td[4][PD_TILE_CLASSES] = {}, f_fk[3][2][4][PD_TILE_CLASSES] = {};
    auto put = [&](int dst, const Env& e, const std::array<int, 3>& tb) {
      truns.push_back(TileRun{dst, &e, tb[1], tb[2]});
    };
    for (int64_t id : ids) {
      Env& e = get_env(c, id);
      if (!e.tiled) continue;
      const int lg = e.lgG, hm = (int)e.hmax, pr = e.precision;
      const int pc = (pr == PD_PREC_FP32C) ? 1 : 0;
      for (auto& tb : e.tile_blocks) {
        const int kind = e.blocks[tb[0]].d.kind;
        const int hf = pd_tile_class(2, kind, pr, xfl, hm);
        put(l_fk[kind][pc][lg][hf].off + f_fk[kind][pc][lg][hf], e, tb);
        f_fk[kind][pc][lg][hf] += tb[2];
        if (kind != PD_BB) {
          const int ht = pd_tile_class(0, kind, pr, xfl, hm);
          put(l_th[pc][lg][ht].off + f_th[pc][lg][ht], e, tb);
          f_th[pc][lg][ht] += tb[2];
        }
        if (kind == PD_EP) {
          const int hd = pd_tile_class(1, kind, pr, xfl, hm);
          put(l_td[lg][hd].off + f_td[lg][hd], e, tb);
          f_td[lg][hd] += tb[2];
        }
      }
    }
  }
  // Launch plan: one launch per class only where the class fills the chip
  // (>= kMergeTiles tiles); smaller lists join the next non-empty class's
  // launch, which runs them with its larger stride.  A launch lasts at least
  // one tile's time, so a small set split over several classes would pay that
  // once per class and pass.
  auto merge = [](TileList* l) {
    int start = -1, n = 0, hc = 1;
    for (int h = 0; h < PD_TILE_CLASSES; ++h) {
      if (!l[h].n) continue;
      if (start < 0) start = l[h].off;
      n += l[h].n;
      hc = std::max(hc, l[h].hcap);
      l[h].n = 0;
      bool later = false;
      for (int g = h + 1; g < PD_TILE_CLASSES; ++g) later = later || l[g].n > 0;
      if (n >= kMergeTiles || !later) {
        l[h].off = start;
        l[h].n = n;
        l[h].hcap = hc;
        start = -1;
        n = 0;
        hc = 1;
      }
    }
  };
  for (int lg = 1; lg <= 3; ++lg) {
    for (int pc = 0; pc < 2; ++pc) {
      merge(l_th[pc][lg]);
      for (int k = 0; k < 3; ++k) merge(l_fk[k][pc][lg]);
    }
    merge(l_td[lg]);
  }
  place_tile_runs(c, truns, copies, l_th, l_td, l_fk, xfl);
  std::copy(&l_th[0][0][0], &l_th[0][0][0] + 2 * 4 * PD_TILE_CLASSES, &c->tl_theta[0][0][0]);
  std::copy(&l_td[0][0], &l_td[0][0] + 4 * PD_TILE_CLASSES, &c->tl_td[0][0]);
  std::copy(&l_fk[0][0][0][0], &l_fk[0][0][0][0] + 3 * 2 * 4 * PD_TILE_CLASSES, &c->tl_force[0][0][0][0]);
  c->set_tiled = tiled;
  c->set_xf = xfl;
  // pair-contact items: (slot, contact index) and per-class workgroup offsets
  {
    std::vector<int2>& items = citems;
    std::vector<std::array<int, 3>> sz;  // n1, n2, ncell_cap
    for (int64_t id : ids) {
      Env& e = get_env(c, id);
      for (size_t k = 0; k < e.cdev.size(); ++k)
        if (!e.cdev[k].self_contact) {
          items.push_back(make_int2(e.slot, (int)k));
          sz.push_back({e.cdev[k].n1, e.cdev[k].n2, e.cdev[k].ncell_cap});
        }
    }
    const int ni = (int)items.size();
    c->n_citems = ni;
    if (ni > 0) {
      std::vector<int32_t>& st = cst;
      st.assign((size_t)PD_CB_NCLS * (ni + 1), 0);
      for (int cls = 0; cls < PD_CB_NCLS; ++cls) {
        int64_t run = 0;
        for (int k = 0; k < ni; ++k) {
          st[(size_t)cls * (ni + 1) + k] = (int32_t)run;
          run += pd_contact_batch_blocks(cls, sz[k][0], sz[k][1], sz[k][2]);
        }
        if (run >= (1LL << 31)) throw PdError(PD_E_ARG, "contact batch too large");
        st[(size_t)cls * (ni + 1) + ni] = (int32_t)run;
        c->cgrid[cls] = (int)run;
      }
      ensure_cap(c->d_citems, c->citems_cap, (size_t)ni, c->mem);
      ensure_cap(c->d_cstarts, c->cstarts_cap, st.size(), c->mem);
      PD_HIP(hipMemcpyAsync(c->d_citems, items.data(), sizeof(int2) * ni, hipMemcpyHostToDevice, c->stream));
      PD_HIP(hipMemcpyAsync(c->d_cstarts, st.data(), sizeof(int32_t) * st.size(), hipMemcpyHostToDevice,
                            c->stream));
    }
  }
  {
    std::vector<int32_t>& cf = cfin;
    std::vector<int32_t>& cfs = cfin_starts;
    int blocks = 0;
    for (int64_t id : ids) {
      const Env& e = get_env(c, id);
      if (!e.dev.cflag) continue;
      cf.push_back(e.slot);
      cfs.push_back(blocks);
      blocks += (int)((e.N + 255) / 256);
    }
    c->n_cfin = (int)cf.size();
    c->cfin_blocks = blocks;
    // one wavefront per 16 particles of these envs (a bound on those with
    // contact terms), at most the 4 per SIMD (of the chip's 1024) that the
    // kernel's registers let reside at once
    c->cfin_grid = std::min(1024, std::max(1, blocks * 4));
    if (!cf.empty()) {
      ensure_cap(c->d_cfin, c->cfin_cap, cf.size(), c->mem);
      ensure_cap(c->d_cfin_starts, c->cfin_starts_cap, cfs.size(), c->mem);
      PD_HIP(hipMemcpyAsync(c->d_cfin, cf.data(), sizeof(int32_t) * cf.size(), hipMemcpyHostToDevice, c->stream));
      PD_HIP(hipMemcpyAsync(c->d_cfin_starts, cfs.data(), sizeof(int32_t) * cfs.size(), hipMemcpyHostToDevice,
                            c->stream));
    }
  }
  ensure_cap(c->d_slots, c->slots_cap, slots.size(), c->mem);
  PD_HIP(hipMemcpyAsync(c->d_slots, slots.data(), sizeof(int32_t) * slots.size(), hipMemcpyHostToDevice, c->stream));
  ensure_cap(c->d_obs_ranges, c->obs_ranges_cap, ranges.size(), c->mem);
  PD_HIP(hipMemcpyAsync(c->d_obs_ranges, ranges.data(), sizeof(int2) * ranges.size(), hipMemcpyHostToDevice,
                        c->stream));
  ensure_cap(c->d_obs_part, c->obs_part_cap, std::max<size_t>((size_t)n_all, 1), c->mem);
  if (!copies.empty()) {
    ensure_cap(c->d_copies, c->copies_cap, copies.size(), c->mem);
    PD_HIP(hipMemcpyAsync(c->d_copies, copies.data(), sizeof(PdCopy) * copies.size(), hipMemcpyHostToDevice,
                          c->stream));
    pd_launch_copy_words(c->stream, c->d_copies, (int)copies.size());
    PD_HIP(hipGetLastError());
  }
  PD_HIP(hipStreamSynchronize(c->stream));  // the host vectors above go out of scope
  c->n_slots = (int)slots.size();
  c->set_has_ep = has_ep;
  c->set_has_theta = has_th;
  c->set_gpow = gpow;
  c->set_ids = ids;
  c->set_full = full;
  c->set_valid = true;
}

int precision_of(Ctx* c, const std::vector<int64_t>& ids) {
  int prec = -1;
  for (int64_t id : ids) {
    Env& e = get_env(c, id);
    if (prec < 0) prec = e.precision;
    else if (prec != e.precision) throw PdError(PD_E_ARG, "envs stepped together must share one precision");
  }
  return prec < 0 ? PD_PREC_FP64 : prec;
}

// deformed-latch work list: the blocks not yet read back as latched of the
// envs that may still hold undeformed blocks, and their slices
int prepare_latch(Ctx* c, const std::vector<int64_t>& ids) {
  std::vector<int2> items, units;
  std::vector<int32_t> ustart;
  c->latch_items.clear();
  for (int64_t id : ids) {
    Env& e = get_env(c, id);
    if (!e.maybe_undeformed) continue;
    e.block_latched.resize(e.blocks.size(), 0);
    e.block_rigid_ok.resize(e.blocks.size(), 0);
    for (size_t b = 0; b < e.blocks.size(); ++b) {
      if (e.block_latched[b] || e.block_rigid_ok[b]) continue;
      const int item = (int)items.size();
      items.push_back(make_int2(e.slot, (int)b));
      c->latch_items.push_back({id, (int)b});
      ustart.push_back((int32_t)units.size());
      const int ns = pd_latch_slices(e.blocks[b].d.n);
      for (int q = 0; q < ns; ++q) units.push_back(make_int2(item, q));
    }
  }
  c->latch_nunits = (int)units.size();
  if (items.empty()) return 0;
  ensure_cap(c->d_latch, c->latch_cap, items.size(), c->mem);
  ensure_cap(c->d_latch_flags, c->latch_flags_cap, items.size(), c->mem);
  ensure_cap(c->d_latch_units, c->latch_units_cap, units.size(), c->mem);
  ensure_cap(c->d_latch_ustart, c->latch_ustart_cap, ustart.size(), c->mem);
  ensure_cap(c->d_latch_part, c->latch_part_cap, units.size(), c->mem);
  PD_HIP(hipMemcpyAsync(c->d_latch, items.data(), sizeof(int2) * items.size(), hipMemcpyHostToDevice, c->stream));
  PD_HIP(hipMemcpyAsync(c->d_latch_units, units.data(), sizeof(int2) * units.size(), hipMemcpyHostToDevice,
                        c->stream));
  PD_HIP(hipMemcpyAsync(c->d_latch_ustart, ustart.data(), sizeof(int32_t) * ustart.size(), hipMemcpyHostToDevice,
                        c->stream));
  PD_HIP(hipStreamSynchronize(c->stream));
  return (int)items.size();
}

void launch_latch(Ctx* c, hipStream_t s, int nl) {
  pd_launch_latch(s, c->d_envs, c->d_latch, nl, c->d_latch_units, c->latch_nunits, c->d_latch_ustart,
                  c->d_latch_part);
}

// After a run with nl latch items (the last prepare_latch): the latches of
// every listed block in one readback.
void refresh_latch(Ctx* c, const std::vector<int64_t>& ids, int nl) {
  (void)ids;
  if (nl <= 0) return;
  std::vector<int32_t> fl(nl);
  pd_launch_latch_flags(c->stream, c->d_envs, c->d_latch, nl, c->d_latch_flags);
  PD_HIP(hipMemcpyAsync(fl.data(), c->d_latch_flags, sizeof(int32_t) * nl, hipMemcpyDeviceToHost, c->stream));
  PD_HIP(hipStreamSynchronize(c->stream));
  for (int k = 0; k < nl; ++k) {
    Env& e = get_env(c, c->latch_items[k].first);
    const int b = c->latch_items[k].second;
    if (fl[k]) e.block_latched[b] = 1;
    else if (!e.block_rigid.empty() && e.block_rigid[b]) e.block_rigid_ok[b] = 1;
  }
  for (int k = 0; k < nl; ++k) {
    Env& e = get_env(c, c->latch_items[k].first);
    bool any = false;
    for (uint8_t l : e.block_latched) any = any || !l;
    e.maybe_undeformed = any;
  }
}

// numbers the particles with contact terms of a tiled env after its lists changed
void cpart_update(Ctx* c, Env& e) {
  if (e.dev.cflag)
    pd_launch_cpart_build(c->stream, c->d_envs, e.d_cself, e.d_cself + 1, 1, (int)((e.N + 255) / 256));
}

void contact_update(Ctx* c, Env& e) {
  for (size_t k = 0; k < e.cdev.size(); ++k)
    if (e.cdev[k].self_contact) pd_launch_self_contact_update(c->stream, c->d_envs, e.slot, e.cdev[k], (int)k);
    else pd_launch_contact_update(c->stream, c->d_envs, e.slot, e.dev, e.cdev[k], (int)k);
  cpart_update(c, e);
  e.lists_current = true;
}

// contact_update of every env of the current run set: the pair contacts of all
// of them in one batched sequence, self contacts per env
void contact_update_set(Ctx* c, const std::vector<int64_t>& ids) {
  pd_launch_contact_update_batch(c->stream, c->d_envs, c->d_citems, c->d_cstarts, c->n_citems, c->cgrid);
  for (int64_t id : ids) {
    Env& e = get_env(c, id);
    for (size_t k = 0; k < e.cdev.size(); ++k)
      if (e.cdev[k].self_contact) pd_launch_self_contact_update(c->stream, c->d_envs, e.slot, e.cdev[k], (int)k);
    if (!e.contacts.empty()) e.lists_current = true;
  }
  pd_launch_cpart_build(c->stream, c->d_envs, c->d_cfin, c->d_cfin_starts, c->n_cfin, c->cfin_blocks);
}

// one force evaluation (update_acc!) for a chunk subset of one env
void force_pass(Ctx* c, Env& e, const std::vector<Chunk>& sub, int flags_extra) {
  std::vector<Chunk> th, td, fk[3];
  for (auto& ch : sub) {
    const int kind = e.blocks[ch.block].d.kind;
    fk[kind].push_back(ch);
    if (kind != PD_BB) th.push_back(ch);
    if (kind == PD_EP) td.push_back(ch);
  }
  std::vector<Chunk> cat;
  cat.insert(cat.end(), th.begin(), th.end());
  const int o_td = (int)cat.size();
  cat.insert(cat.end(), td.begin(), td.end());
  int o_f[3];
  for (int k = 0; k < 3; ++k) { o_f[k] = (int)cat.size(); cat.insert(cat.end(), fk[k].begin(), fk[k].end()); }
  Chunk* d = nullptr;
  PD_HIP(hipMalloc(&d, sizeof(Chunk) * std::max<size_t>(cat.size(), 1)));
  int nl = 0;
  try {
    PD_HIP(hipMemcpyAsync(d, cat.data(), sizeof(Chunk) * cat.size(), hipMemcpyHostToDevice, c->stream));
    std::vector<int64_t> one{e.id};
    nl = prepare_latch(c, one);
    if (nl) launch_latch(c, c->stream, nl);
    pd_launch_theta(c->stream, e.precision, d, (int)th.size(), c->d_envs);
    pd_launch_td(c->stream, e.precision, d + o_td, (int)td.size(), c->d_envs);
    for (int k = 0; k < 3; ++k)
      pd_launch_force(c->stream, k, e.precision, d + o_f[k], (int)fk[k].size(), c->d_envs, flags_extra,
                      env_gpow(e));
    if (e.has_ep && !td.empty()) {
      // the force pass wrote the updated plastic strain into the other buffer
      pd_launch_book(c->stream, c->d_envs, e.d_self, 1, PD_B_FLIP_EDP);
      e.edp_cur ^= 1;
    }
    // bonds broken by these (untiled) passes: refresh the tiles' intact bits
    if (e.tiled) pd_launch_pk_sync(c->stream, e.d_chunks, (int)e.chunks.size(), c->d_envs);
    PD_HIP(hipGetLastError());
    PD_HIP(hipStreamSynchronize(c->stream));
  } catch (...) {
    (void)hipStreamSynchronize(c->stream);
    (void)hipFree(d);
    throw;
  }
  PD_HIP(hipFree(d));
  refresh_latch(c, {e.id}, nl);
}

// 3 x N reference-order host array <-> internal-order double4 device array
void download_xyz(Ctx* c, const Env& e, const double4* src, double* out) {
  const int64_t N = e.N;
  std::vector<double4> h(N);
  PD_HIP(hipMemcpyAsync(h.data(), src, sizeof(double4) * N, hipMemcpyDeviceToHost, c->stream));
  PD_HIP(hipStreamSynchronize(c->stream));
  for (int64_t p = 0; p < N; ++p) {
    const int64_t r = reference_of(e, p);
    out[3 * r] = h[p].x; out[3 * r + 1] = h[p].y; out[3 * r + 2] = h[p].z;
  }
}

void upload_xyz(Ctx* c, const Env& e, double4* dst, const double* in, bool keep_w) {
  const int64_t N = e.N;
  std::vector<double4> h(N);
  if (keep_w) {
    PD_HIP(hipMemcpyAsync(h.data(), dst, sizeof(double4) * N, hipMemcpyDeviceToHost, c->stream));
    PD_HIP(hipStreamSynchronize(c->stream));
  }
  for (int64_t p = 0; p < N; ++p) {
    const int64_t r = reference_of(e, p);
    h[p].x = in[3 * r]; h[p].y = in[3 * r + 1]; h[p].z = in[3 * r + 2];
    if (!keep_w) h[p].w = 0.0;
  }
  PD_HIP(hipMemcpyAsync(dst, h.data(), sizeof(double4) * N, hipMemcpyHostToDevice, c->stream));
  PD_HIP(hipStreamSynchronize(c->stream));
}

// Positions from the host into the current y buffer; rigidly held blocks are
// tested again (block_rigid_ok).
void upload_y(Ctx* c, Env& e, const double* y) {
  upload_xyz(c, e, e.dev.y[e.cur], y, false);
  std::fill(e.block_rigid_ok.begin(), e.block_rigid_ok.end(), (uint8_t)0);
}

struct BondsHost {
  std::vector<int32_t> col, deg;
  std::vector<uint8_t> state;
  std::vector<uint32_t> pslot0;
};

void download_bonds(Ctx* c, Env& e, BondsHost& h) {
  h.col.resize(std::max<int64_t>(e.nslots, 1));
  h.state.resize(std::max<int64_t>(e.nslots, 1));
  h.deg.resize(e.N);
  h.pslot0.resize(e.N);
  PD_HIP(hipMemcpyAsync(h.col.data(), e.dev.col, sizeof(int32_t) * e.nslots, hipMemcpyDeviceToHost, c->stream));
  PD_HIP(hipMemcpyAsync(h.state.data(), e.dev.state, e.nslots, hipMemcpyDeviceToHost, c->stream));
  PD_HIP(hipMemcpyAsync(h.deg.data(), e.dev.deg, sizeof(int32_t) * e.N, hipMemcpyDeviceToHost, c->stream));
  PD_HIP(hipMemcpyAsync(h.pslot0.data(), e.dev.pslot0, sizeof(uint32_t) * e.N, hipMemcpyDeviceToHost, c->stream));
  PD_HIP(hipStreamSynchronize(c->stream));
}

HostBlock& get_block(Env& e, int64_t b) {
  if (b < 0 || b >= (int64_t)e.blocks.size()) throw PdError(PD_E_ARG, "block index out of range");
  return e.blocks[b];
}

// Bonds of reference particle r as (reference neighbour index, slot), in the
// family's descending order (general_material.jl:88-90).
void row_of(const Env& e, const BondsHost& h, int64_t r, std::vector<std::pair<int64_t, uint32_t>>& out) {
  const int64_t p = internal_of(e, r);
  const int d = h.deg[p];
  out.clear();
  for (int k = 0; k < d; ++k) {
    const uint32_t slot = pd_slot(h.pslot0[p], k, e.lgG);
    out.push_back({reference_of(e, h.col[slot]), slot});
  }
  std::sort(out.begin(), out.end(),
            [](const std::pair<int64_t, uint32_t>& a, const std::pair<int64_t, uint32_t>& b) { return a.first > b.first; });
}

thread_local std::string g_err;

// One captured step sequence, replayed with hipGraphLaunch; destroyed with the
// pd_run call that made it.
struct StepGraph {
  hipGraphExec_t exec = nullptr;
  void reset() {  // the stream must be idle
    if (exec) (void)hipGraphExecDestroy(exec);
    exec = nullptr;
  }
  template <typename F>
  void capture(hipStream_t s, F&& enqueue) {
    hipGraph_t g = nullptr;
    PD_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    try {
      enqueue();
    } catch (...) {
      (void)hipStreamEndCapture(s, &g);
      if (g) (void)hipGraphDestroy(g);
      throw;
    }
    PD_HIP(hipStreamEndCapture(s, &g));
    const hipError_t e = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) {
      exec = nullptr;
      throw PdError(PD_E_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
    }
  }
  ~StepGraph() {
    if (exec) (void)hipGraphExecDestroy(exec);
  }
};

constexpr int64_t kLatchPoll = 16;
constexpr int64_t kGraphMinSteps = 32;  // shorter pd_run calls launch every step eagerly

// Contact-finish shape: one lane per particle once the set's particles with
// contact terms number at least this (PERIDYN_CFIN_LANE_MIN overrides), one
// wavefront per particle below it.
int cfin_lane_min() {
  static const int v = [] {
    const char* e = std::getenv("PERIDYN_CFIN_LANE_MIN");
    return e ? std::atoi(e) : 32768;
  }();
  return v;
}

constexpr int kForkTiles = 16384, kForkMinTiles = 1024;

// PERIDYN_FORK=0 keeps every pass of a step on one stream, =1 forks whenever
// the step has both kinds of pass (tests); unset, the sizes decide
int fork_mode() {
  const char* e = std::getenv("PERIDYN_FORK");
  if (!e) return -1;
  return std::string(e) == "0" ? 0 : 1;
}

// PERIDYN_GRAPH=0 launches every step eagerly (for comparison)
bool graphs_enabled() {
  const char* v = std::getenv("PERIDYN_GRAPH");
  return !(v && std::string(v) == "0");
}

// Host bookkeeping after the steps of a run have completed: every env's step
// counter, buffer indices and list state, and the deformed latches read back.
void end_run(Ctx* c, const std::vector<int64_t>& ids, int64_t nsteps, int64_t nuf, int nl) {
  for (int64_t id : ids) {
    Env& e = get_env(c, id);
    e.time_step += nsteps;
    e.cur ^= (int32_t)(nsteps & 1);
    if (e.has_ep) e.edp_cur ^= (int32_t)(nsteps & 1);
    e.lists_current = (nsteps % nuf) == 0;  // rebuilt after the last step
  }
  refresh_latch(c, ids, nl);
}

// run! for the listed envs (pd_run); with every > 0 the printdata columns of
// step 1 and of every every-th step (solvers.jl:158) go to log_host (rows x
// nenv x 8, rows in step order).  async (pd_run_begin): every step is launched
// eagerly, the deformed latches are read back only at the end, and the call
// returns once the steps are enqueued; pd_run_end completes it.
void run_impl(Ctx* c, const std::vector<int64_t>& ids, int64_t nsteps, int64_t nuf, int64_t every,
              double* log_host, bool async = false) {
  PD_HIP(hipSetDevice(c->device));
  const int prec = precision_of(c, ids);
  build_set(c, ids);
  hipStream_t s = c->stream;
  const int nenv = (int)ids.size();
  const int64_t nrows = every > 0 ? pd_log_rows(nsteps, every) : 0;
  if (nrows > 0) ensure_cap(c->d_obs_rows, c->obs_rows_cap, (size_t)(nrows * nenv * 8), c->mem);
  // run! start: update_neighs! (solvers.jl:114); apply_bc_at0 records
  // BC.last = y, which the resident state already holds.  Contact lists are
  // a deterministic function of the positions (sorted), so lists built from
  // the current positions are kept rather than rebuilt identically.
  for (int64_t id : ids) {
    Env& e = get_env(c, id);
    if (!e.contacts.empty() && !e.lists_current) contact_update(c, e);
  }
  if (nsteps == 0) return;
  int nl = prepare_latch(c, ids);
  EnvDev* T = c->d_envs;
  Chunk* d = c->d_set;
  timed(c, PD_PROF_PROLOGUE, [&] { pd_launch_prologue(s, d, c->n_all, T); });
  pd_launch_book(s, T, c->d_slots, c->n_slots, PD_B_FLIP_Y);
  int64_t row = 0;
  // the kernels of one velocity_verlet_step! (all envs of the set); a logged
  // step takes the momentum from the half-step velocities just before the
  // force pass (update_misc! runs inside update_acc!) and the force and damage
  // sums after it
  // the bond-based force passes read neither theta nor td and write only their
  // own blocks' particles and bonds, so with dilatation passes in the step
  // they run on the fork stream meanwhile (not when profiling: every class
  // then keeps its own events on the main stream).  That pays where the
  // launches are short: a launch of kForkTiles tiles or more fills the chip's
  // workgroup slots ten times over, its tail and launch gap are a few percent,
  // and the concurrent passes only contend for the same busy VALUs (measured:
  // 256 specimens, 33k tiles per launch, 0.5% slower forked; 22-64 served
  // requests, 2k-10k tiles per launch, 7-10% faster).  Both sides must hold
  // at least a round of the chip's workgroups (kForkMinTiles), or the events
  // between the streams cost more than the overlap gains (one indentation
  // request alone: its rigid indenter is the only bond-based block)
  int max_launch = 0, bb_tiles = c->n_force_u[PD_BB] / PD_TILE_CHUNKS, other_tiles = 0;
  for (int k = 1; k < 3; ++k) other_tiles += c->n_force_u[k] / PD_TILE_CHUNKS;
  for (int lg = 1; lg <= 3; ++lg)
    for (int h = 0; h < PD_TILE_CLASSES; ++h)
      for (int pc = 0; pc < 2; ++pc) {
        max_launch = std::max(max_launch, c->tl_theta[pc][lg][h].n);
        other_tiles += c->tl_theta[pc][lg][h].n;
        for (int k = 0; k < 3; ++k) {
          max_launch = std::max(max_launch, c->tl_force[k][pc][lg][h].n);
          (k == PD_BB ? bb_tiles : other_tiles) += c->tl_force[k][pc][lg][h].n;
        }
      }
  const int fm = fork_mode();
  const bool sized = max_launch < kForkTiles && bb_tiles >= kForkMinTiles && other_tiles >= kForkMinTiles;
  const bool fork = fm != 0 && !c->prof && c->n_force[PD_BB] > 0 && (c->set_has_theta || c->set_has_ep) &&
                    (fm == 1 || sized);
  if (fork && !c->fork_stream) {
    PD_HIP(hipStreamCreateWithFlags(&c->fork_stream, hipStreamNonBlocking));
    PD_HIP(hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming));
    PD_HIP(hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming));
  }
  auto force_kind = [&](hipStream_t fs, int k, int fl) {
    Tile* dt = c->d_tset;
    const int xf = c->set_xf;
    if (c->n_force_u[k]) pd_launch_force(fs, k, prec, d + c->off_force_u[k], c->n_force_u[k], T, fl, c->set_gpow);
    for (int lg = 1; lg <= 3; ++lg)
      for (int h = 0; h < PD_TILE_CLASSES; ++h) {
        const Ctx::TileList& l0 = c->tl_force[k][0][lg][h];
        const Ctx::TileList& l1 = c->tl_force[k][1][lg][h];
        if (l0.n) pd_launch_tile_force(fs, k, lg, xf, dt + l0.off, l0.n, T, fl, l0.hcap);
        if (l1.n) pd_launch_tile_force32(fs, k, lg, dt + l1.off, l1.n, T, fl, l1.hcap);
      }
  };
  auto step_kernels = [&](bool last, bool logged) {
    if (nl) timed(c, PD_PROF_LATCH, [&] { launch_latch(c, s, nl); });
    Tile* dt = c->d_tset;
    const int xf = c->set_xf;
    const int fl = PD_F_KICK | (last ? 0 : PD_F_DRIFT);
    const bool fk = fork && !logged;  // logged steps read the half-step velocities in between
    if (fk) {
      PD_HIP(hipEventRecord(c->ev_fork, s));
      PD_HIP(hipStreamWaitEvent(c->fork_stream, c->ev_fork, 0));
      force_kind(c->fork_stream, PD_BB, fl);
      PD_HIP(hipEventRecord(c->ev_join, c->fork_stream));
    }
    if (c->set_has_theta)
      timed(c, PD_PROF_THETA, [&] {
        if (c->n_theta_u) pd_launch_theta(s, prec, d + c->off_theta_u, c->n_theta_u, T);
        for (int lg = 1; lg <= 3; ++lg)
          for (int h = 0; h < PD_TILE_CLASSES; ++h) {
            const Ctx::TileList& l0 = c->tl_theta[0][lg][h];
            const Ctx::TileList& l1 = c->tl_theta[1][lg][h];
            if (l0.n) pd_launch_tile_theta(s, lg, xf, dt + l0.off, l0.n, T, l0.hcap);
            if (l1.n) pd_launch_tile_theta32(s, lg, dt + l1.off, l1.n, T, l1.hcap);
          }
      });
    if (c->set_has_ep)
      timed(c, PD_PROF_TD, [&] {
        if (c->n_td_u) pd_launch_td(s, prec, d + c->off_td_u, c->n_td_u, T);
        for (int lg = 1; lg <= 3; ++lg)
          for (int h = 0; h < PD_TILE_CLASSES; ++h) {
            const Ctx::TileList& l = c->tl_td[lg][h];
            if (l.n) pd_launch_tile_td(s, lg, xf, dt + l.off, l.n, T, l.hcap);
          }
      });
    double* rows = logged ? c->d_obs_rows + (size_t)row * nenv * 8 : nullptr;
    if (logged) pd_launch_observables(s, 0, d, c->n_all, T, c->d_obs_part, c->d_obs_ranges, nenv, rows);
    for (int k = 0; k < 3; ++k)
      if (c->n_force[k] && !(fk && k == PD_BB)) timed(c, PD_PROF_FORCE_BB + k, [&] { force_kind(s, k, fl); });
    if (fk) PD_HIP(hipStreamWaitEvent(s, c->ev_join, 0));
    if (c->n_cfin) timed(c, PD_PROF_CONTACT, [&] { pd_launch_contact_finish(s, T, c->d_cfin, c->n_cfin, c->cfin_grid, cfin_lane_min(), fl); });
    if (logged) pd_launch_observables(s, 1, d, c->n_all, T, c->d_obs_part, c->d_obs_ranges, nenv, rows);
  };
  auto book = [&](bool last) {
    timed(c, PD_PROF_BOOK, [&] {
      pd_launch_book(s, T, c->d_slots, c->n_slots,
                     PD_B_STEP | (last ? 0 : PD_B_FLIP_Y) | (c->set_has_ep ? PD_B_FLIP_EDP : 0));
    });
  };
  // Steps that are neither the last nor logged are identical launch sequences
  // (all step-dependent state lives on the device), so they replay captured
  // HIP graphs: one for plain steps, one for steps followed by the contact
  // rebuild (whose kernels read the positions on the device; no host sync).
  // Profiled runs launch eagerly so that every kernel class keeps its own events.
  StepGraph graph, rgraph;
  // capturing and instantiating the two step graphs costs ~2-3 ms; measured on
  // the 256-specimen batch (4 launches per step) a 20-step call is faster
  // eager (1.48 vs 1.52 ms/step), a 100-step call faster replayed (1.38 vs 1.39)
  const bool use_graph = graphs_enabled() && !c->prof && nsteps >= kGraphMinSteps && !async;
  for (int64_t st = 1; st <= nsteps; ++st) {
    const bool last = (st == nsteps);
    const bool logged = every > 0 && (st == 1 || st % every == 0);
    // run! (solvers.jl:133,177) updates neighbours after loop step i when
    // i % neigh_update_freq == 0, i counted from 1 in each call (start_at = 0)
    bool rebuild = false;
    if (st % nuf == 0)
      for (int64_t id : ids) rebuild = rebuild || !get_env(c, id).contacts.empty();
    // with latch items pending, step 1 runs eagerly and the latches are read
    // back after it (below), so that the captured steps carry no latch kernels
    // once every block has latched (the latch never resets, general_material.jl:240-256)
    if (use_graph && !last && !logged && !(st == 1 && nl)) {
      StepGraph& g = rebuild ? rgraph : graph;
      if (!g.exec) {
        g.capture(s, [&] {
          step_kernels(false, false);
          if (rebuild) contact_update_set(c, ids);
          book(false);
        });
      }
      PD_HIP(hipGraphLaunch(g.exec, s));
    } else {
      step_kernels(last, logged);
      if (logged) ++row;
      if (rebuild) timed(c, PD_PROF_CONTACT, [&] { contact_update_set(c, ids); });
      book(last);
      if (st == 1 && nl && !last && !async) {
        refresh_latch(c, ids, nl);
        nl = prepare_latch(c, ids);
      }
    }
    // while latches are pending, they are read back every kLatchPoll steps;
    // when envs have latched, the graphs are captured again without their
    // latch kernels (a no-op once latched)
    if (nl && use_graph && !last && st % kLatchPoll == 0) {
      refresh_latch(c, ids, nl);  // synchronizes the stream
      const int nl2 = prepare_latch(c, ids);
      if (nl2 != nl) {
        graph.reset();
        rgraph.reset();
      }
      nl = nl2;
    }
  }
  PD_HIP(hipGetLastError());
  if (async) {
    c->inflight.on = true;
    c->inflight.ids = ids;
    c->inflight.nsteps = nsteps;
    c->inflight.nuf = nuf;
    c->inflight.nl = nl;
    return;
  }
  PD_HIP(hipStreamSynchronize(s));  // also before the step graph is destroyed
  if (nrows > 0) {
    PD_HIP(hipMemcpyAsync(log_host, c->d_obs_rows, sizeof(double) * (size_t)(nrows * nenv * 8),
                          hipMemcpyDeviceToHost, s));
    PD_HIP(hipStreamSynchronize(s));
    // time column (time_step * dt after the logged step) and mean damage
    int64_t r = 0;
    for (int64_t st = 1; st <= nsteps; ++st) {
      if (!(st == 1 || st % every == 0)) continue;
      for (int k = 0; k < nenv; ++k) {
        const Env& e = get_env(c, ids[k]);
        double* out = log_host + ((size_t)r * nenv + k) * 8;
        out[0] = (double)(e.time_step + st) * e.dt;
        out[7] = out[7] / (double)e.N;
      }
      ++r;
    }
  }
  end_run(c, ids, nsteps, nuf, nl);
}

// ---- resident-state snapshot ----------------------------------------------
// Blob: 16 x int64 header (magic, version, N, nslots, nblocks, precision, lgG,
// has_ep, time_step, tiled, 6 reserved), nblocks x int32 deformed latches, then y, v,
// f, aux (N double4 each), theta, td (N double each), bond state (nslots
// bytes) and, for elastoplastic envs, edp (nslots double); all in the env's
// internal particle / slot order.
constexpr int64_t kStateMagic = 0x5453445045524550LL;  // "PEREPDST"
constexpr int64_t kStateVersion = 1;

int64_t state_bytes(const Env& e) {
  int64_t b = 16 * 8 + 4 * (int64_t)e.blocks.size();
  b = (b + 7) / 8 * 8;
  b += 4 * 32 * e.N + 2 * 8 * e.N;
  b += (e.nslots + 7) / 8 * 8;
  if (e.has_ep) b += 8 * e.nslots;
  return b;
}

}  // namespace

// =============================================================================
extern "C" {

static int run_guard(void* ctx, const std::function<void()>& f) {
  try {
    f();
    return PD_OK;
  } catch (const PdError& e) {
    if (ctx) C(ctx)->err = e.msg; else g_err = e.msg;
    return e.code;
  } catch (const std::exception& e) {
    if (ctx) C(ctx)->err = e.what(); else g_err = e.what();
    return PD_E_ARG;
  }
}

// Entry points that read or change state an in-flight run uses take their
// context through idle(): they fail while a pd_run_begin run is in flight.
static Ctx* idle(void* ctx) {
  Ctx* c = C(ctx);
  if (c->inflight.on)
    throw PdError(PD_E_STATE, "a run started by pd_run_begin is in flight: call pd_run_end first");
  return c;
}

// The calls allowed during an in-flight run (env construction, commit,
// destroy) enqueue on the admission stream meanwhile.
struct AdmissionScope {
  Ctx* c;
  bool swapped = false;
  explicit AdmissionScope(Ctx* cc) : c(cc) {
    if (c->inflight.on) {
      c->stream = c->adm_stream;
      c->mem.stream = c->adm_stream;
      swapped = true;
    }
  }
  ~AdmissionScope() {
    if (swapped) {
      c->stream = c->run_stream;
      c->mem.stream = c->run_stream;
    }
  }
};

int pd_ctx_create(int device, void** out) {
  return run_guard(nullptr, [&] {
    if (!out) throw PdError(PD_E_ARG, "null output pointer");
    int n = 0;
    PD_HIP(hipGetDeviceCount(&n));
    if (device < 0 || device >= n) throw PdError(PD_E_ARG, "no HIP device " + std::to_string(device));
    PD_HIP(hipSetDevice(device));
    Ctx* c = new Ctx();
    c->device = device;
    hipError_t r = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (r != hipSuccess) { delete c; PD_HIP(r); }
    c->run_stream = c->stream;
    // the address of this GPU's copy of `scratch` (device is current)
    void* sp = nullptr;
    r = hipGetSymbolAddress(&sp, HIP_SYMBOL(scratch));
    if (r != hipSuccess || !sp) {
      (void)hipStreamDestroy(c->stream);
      delete c;
      if (r == hipSuccess) throw PdError(PD_E_HIP, "hipGetSymbolAddress returned no address for the scratch array");
      PD_HIP(r);
    }
    c->scratch = static_cast<double4*>(sp);
    // PERIDYN_POISON=1: fill device memory with 0xFF when handed out (tests)
    const char* pz = std::getenv("PERIDYN_POISON");
    c->mem.stream = c->stream;
    c->mem.poison = pz && std::string(pz) == "1";
    pd_tile_allow_lds(kTileLdsBudget);
    *out = c;
  });
}

int pd_ctx_destroy(void* ctx) {
  if (!ctx) return PD_E_ARG;
  Ctx* c = C(ctx);
  (void)hipSetDevice(c->device);
  (void)hipStreamSynchronize(c->run_stream);
  if (c->adm_stream) (void)hipStreamSynchronize(c->adm_stream);
  if (c->fork_stream) (void)hipStreamSynchronize(c->fork_stream);
  c->inflight.on = false;
  c->stream = c->run_stream;
  for (auto& kv : c->envs) { free_env_device(*kv.second); delete kv.second; }
  c->mem.trim(0);
  if (c->stage.pinned) (void)hipHostFree(c->stage.pinned);
  if (c->d_envs) (void)hipFree(c->d_envs);
  if (c->d_set) (void)hipFree(c->d_set);
  if (c->d_tset) (void)hipFree(c->d_tset);
  if (c->d_slots) (void)hipFree(c->d_slots);
  if (c->d_latch) (void)hipFree(c->d_latch);
  if (c->d_latch_part) (void)hipFree(c->d_latch_part);
  if (c->d_latch_flags) (void)hipFree(c->d_latch_flags);
  if (c->d_latch_units) (void)hipFree(c->d_latch_units);
  if (c->d_latch_ustart) (void)hipFree(c->d_latch_ustart);
  if (c->d_cfin) (void)hipFree(c->d_cfin);
  if (c->d_cfin_starts) (void)hipFree(c->d_cfin_starts);
  if (c->d_obs_ranges) (void)hipFree(c->d_obs_ranges);
  if (c->d_obs_part) (void)hipFree(c->d_obs_part);
  if (c->d_obs_rows) (void)hipFree(c->d_obs_rows);
  if (c->d_citems) (void)hipFree(c->d_citems);
  if (c->d_copies) (void)hipFree(c->d_copies);
  if (c->d_cstarts) (void)hipFree(c->d_cstarts);
  for (auto& r : c->pending) { c->pool.push_back(r.a); c->pool.push_back(r.b); }
  for (hipEvent_t ev : c->pool) (void)hipEventDestroy(ev);
  (void)hipStreamDestroy(c->run_stream);
  if (c->adm_stream) (void)hipStreamDestroy(c->adm_stream);
  if (c->fork_stream) (void)hipStreamDestroy(c->fork_stream);
  if (c->ev_fork) (void)hipEventDestroy(c->ev_fork);
  if (c->ev_join) (void)hipEventDestroy(c->ev_join);
  delete c;
  return PD_OK;
}

const char* pd_last_error(void* ctx) { return ctx ? C(ctx)->err.c_str() : g_err.c_str(); }

int pd_sync(void* ctx) {
  return run_guard(ctx, [&] { PD_HIP(hipStreamSynchronize(idle(ctx)->stream)); });
}

int pd_ctx_mem_info(void* ctx, int64_t* free_bytes, int64_t* total_bytes) {
  return run_guard(ctx, [&] {
    PD_HIP(hipSetDevice(C(ctx)->device));
    size_t fr = 0, tot = 0;
    PD_HIP(hipMemGetInfo(&fr, &tot));
    fr += C(ctx)->mem.cached;  // kept for this context's next envs
    if (free_bytes) *free_bytes = (int64_t)fr;
    if (total_bytes) *total_bytes = (int64_t)tot;
  });
}

int pd_env_create(void* ctx, double dt, int32_t precision, int64_t* env) {
  return run_guard(ctx, [&] {
    if (!env) throw PdError(PD_E_ARG, "null output pointer");
    if (!(std::isfinite(dt))) throw PdError(PD_E_ARG, "dt must be finite");
    if (precision != PD_PREC_EXACT && precision != PD_PREC_FP64 && precision != PD_PREC_FP32C)
      throw PdError(PD_E_ARG, "unsupported precision form " + std::to_string(precision));
    Ctx* c = C(ctx);
    Env* e = new Env();
    e->mem = &c->mem;
    e->id = c->next_id++;
    e->dt = dt;
    e->precision = precision;
    c->envs[e->id] = e;
    *env = e->id;
  });
}

int pd_env_add_block(void* ctx, int64_t env, const pd_block_desc* d, int64_t* block) {
  return run_guard(ctx, [&] {
    Env& e = get_env(C(ctx), env);
    if (e.committed) throw PdError(PD_E_STATE, "env already committed");
    if (!d || d->n <= 0 || !d->x || !d->volume || !d->type || !d->density || !d->critical_stretch)
      throw PdError(PD_E_ARG, "block descriptor is missing required arrays");
    if (d->kind < PD_BB || d->kind > PD_EP) throw PdError(PD_E_ARG, "unknown material kind");
    if (d->ntype < 1 || d->ntype > PD_MAX_TYPES) throw PdError(PD_E_ARG, "ntype must be in 1..8");
    HostBlock b;
    b.d = *d;
    const int64_t n = d->n;
    const int nt = d->ntype;
    b.x.assign(d->x, d->x + 3 * n);
    b.y.assign(d->y ? d->y : d->x, (d->y ? d->y : d->x) + 3 * n);
    if (d->v) b.v.assign(d->v, d->v + 3 * n); else b.v.assign(3 * n, 0.0);
    b.vol.assign(d->volume, d->volume + n);
    b.type.assign(d->type, d->type + n);
    for (int64_t i = 0; i < n; ++i)
      if (b.type[i] < d->type_start || b.type[i] >= d->type_start + nt)
        throw PdError(PD_E_ARG, "particle type outside the block's type range");
    b.density.assign(d->density, d->density + nt);
    b.s0 = copy_mat(d->critical_stretch, nt, "critical_stretch", true);
    if (d->kind == PD_BB) {
      b.c = copy_mat(d->bond_stiffness, nt, "bond_stiffness", true);
    } else {
      b.K = copy_mat(d->bulk_modulus, nt, "bulk_modulus", true);
      b.G = copy_mat(d->shear_modulus, nt, "shear_modulus", true);
      if (d->kind == PD_EP) b.psi = copy_mat(d->psi, nt, "psi", true);
    }
    if (b.K.empty()) b.K.assign(nt * nt, 0.0);
    if (b.G.empty()) b.G.assign(nt * nt, 0.0);
    if (b.psi.empty()) b.psi.assign(nt * nt, 0.0);
    if (b.c.empty()) b.c.assign(nt * nt, 0.0);
    e.blocks.push_back(std::move(b));
    if (block) *block = (int64_t)e.blocks.size() - 1;
  });
}

int pd_env_add_contact(void* ctx, int64_t env, const pd_contact_desc* d) {
  return run_guard(ctx, [&] {
    Env& e = get_env(C(ctx), env);
    if (e.committed) throw PdError(PD_E_STATE, "env already committed");
    if (!d || !d->types1 || d->ntypes1 <= 0 || d->ntypes2 < 0 || (d->ntypes2 > 0 && !d->types2))
      throw PdError(PD_E_ARG, "contact descriptor needs a first type list (and a second one unless self contact)");
    if ((int)e.contacts.size() >= PD_MAX_CONTACTS) throw PdError(PD_E_ARG, "too many contact models");
    if (d->exponent_is_int && d->exponent != std::floor(d->exponent))
      throw PdError(PD_E_ARG, "integer exponent must be integer-valued");
    HostContact hc;
    hc.d = *d;
    hc.t1.assign(d->types1, d->types1 + d->ntypes1);
    if (d->ntypes2 > 0) hc.t2.assign(d->types2, d->types2 + d->ntypes2);
    e.contacts.push_back(std::move(hc));
  });
}

int pd_env_add_bc(void* ctx, int64_t env, const pd_bc_desc* d) {
  return run_guard(ctx, [&] {
    Env& e = get_env(C(ctx), env);
    if (e.committed) throw PdError(PD_E_STATE, "env already committed");
    if (!d || !d->mask) throw PdError(PD_E_ARG, "BC descriptor needs a mask");
    if (d->kind != PD_BC_FIX && d->kind != PD_BC_TOFRO) throw PdError(PD_E_ARG, "unknown BC kind");
    if (d->kind == PD_BC_TOFRO && !std::isinf(d->freq) && (d->freq != std::floor(d->freq) || d->freq <= 0))
      throw PdError(PD_E_ARG, "ToFroBC freq must be a positive integer or Inf");
    if (d->options & ~(PD_BC_ONLYATSTART | PD_BC_DIMS(7))) throw PdError(PD_E_ARG, "unknown BC options");
    const int dims = (d->options >> 8) & 7;
    if (dims != 0 && dims != 7)
      throw PdError(PD_E_ARG,
                    "BC dims must select all three dimensions: the reference assigns its 3 x n update into "
                    "y[dims, bool] (boundary_conditions.jl:108-110), a DimensionMismatch for a subset");
    int64_t N = 0;
    for (auto& b : e.blocks) N += b.d.n;
    HostBc hb;
    hb.d = *d;
    hb.mask.assign(d->mask, d->mask + N);
    e.bcs.push_back(std::move(hb));
  });
}

int pd_env_set_schedule(void* ctx, int64_t env, int32_t lanes_per_particle) {
  return run_guard(ctx, [&] {
    Env& e = get_env(C(ctx), env);
    if (e.committed) throw PdError(PD_E_STATE, "env already committed");
    if (lanes_per_particle == 0) { e.lgG = -1; return; }
    int lg = 0;
    while ((1 << lg) < lanes_per_particle) ++lg;
    if ((1 << lg) != lanes_per_particle || lg > 6)
      throw PdError(PD_E_ARG, "lanes_per_particle must be a power of two in 1..64");
    if (e.precision == PD_PREC_EXACT && lg != 0)
      throw PdError(PD_E_ARG, "the reference-order form needs one lane per particle");
    e.lgG = lg;
  });
}

int pd_env_schedule(void* ctx, int64_t env, int32_t* lanes_per_particle) {
  return run_guard(ctx, [&] {
    Env& e = get_env(C(ctx), env);
    if (!e.committed) throw PdError(PD_E_STATE, "env not committed");
    *lanes_per_particle = 1 << e.lgG;
  });
}

int pd_env_layout(void* ctx, int64_t env, int32_t* lanes_per_particle, int32_t* tiled, int32_t* max_halo) {
  return run_guard(ctx, [&] {
    Env& e = get_env(C(ctx), env);
    if (!e.committed) throw PdError(PD_E_STATE, "env not committed");
    if (lanes_per_particle) *lanes_per_particle = 1 << e.lgG;
    if (tiled) *tiled = e.tiled ? (e.xf32 ? 2 : 1) : 0;
    if (max_halo) *max_halo = e.tiled ? e.hmax : 0;
  });
}

int pd_env_commit(void* ctx, int64_t env) {
  return run_guard(ctx, [&] {
    Ctx* c = C(ctx);
    PD_HIP(hipSetDevice(c->device));
    Env& e = get_env(c, env);
    AdmissionScope adm(c);
    try {
      commit_env(c, e);
    } catch (...) {
      free_env_device(e);
      throw;
    }
  });
}

int pd_env_destroy(void* ctx, int64_t env) {
  return run_guard(ctx, [&] {
    Ctx* c = C(ctx);
    Env& e = get_env(c, env);
    if (c->inflight.on && std::find(c->inflight.ids.begin(), c->inflight.ids.end(), env) != c->inflight.ids.end())
      throw PdError(PD_E_STATE, "env " + std::to_string(env) + " is stepped by the in-flight run: call pd_run_end first");
    AdmissionScope adm(c);
    PD_HIP(hipStreamSynchronize(c->stream));
    free_env_device(e);
    if (e.slot >= 0) c->free_slots.push_back(e.slot);
    c->envs.erase(env);
    delete &e;
    c->set_valid = false;
  });
}

int pd_env_info(void* ctx, int64_t env, int64_t* np, int64_t* nb, int64_t* bytes) {
  return run_guard(ctx, [&] {
    Env& e = get_env(C(ctx), env);
    if (np) *np = e.N;
    if (nb) *nb = e.nbonds;
    if (bytes) *bytes = (int64_t)e.bytes;
  });
}

int pd_block_family_rows(void* ctx, int64_t env, int64_t block, int64_t* rows) {
  return run_guard(ctx, [&] {
    Env& e = get_env(idle(ctx), env);
    if (!e.committed) throw PdError(PD_E_STATE, "env not committed");
    *rows = get_block(e, block).maxdeg + 1;
  });
}

int pd_block_family_export(void* ctx, int64_t env, int64_t block, int64_t* family, uint8_t* intact, int64_t rows) {
  return run_guard(ctx, [&] {
    Ctx* c = idle(ctx);
    Env& e = get_env(c, env);
    if (!e.committed) throw PdError(PD_E_STATE, "env not committed");
    HostBlock& b = get_block(e, block);
    if (rows < b.maxdeg + 1) throw PdError(PD_E_ARG, "rows must be >= max degree + 1");
    BondsHost h;
    download_bonds(c, e, h);
    std::vector<std::pair<int64_t, uint32_t>> row;
    for (int64_t i = 0; i < b.d.n; ++i) {
      row_of(e, h, b.first + i, row);
      const int64_t d = (int64_t)row.size();
      for (int64_t r = 0; r < rows; ++r) {
        int64_t v = 0;
        uint8_t s = 0;
        if (r < d) {
          v = row[r].first - b.first + 1;
          s = h.state[row[r].second];
        }
        if (family) family[i * rows + r] = v;
        if (intact) intact[i * rows + r] = s;
      }
    }
  });
}

static void download_block_vec(Ctx* c, Env& e, HostBlock& b, const double* src, double* out) {
  std::vector<double> h(e.N);
  PD_HIP(hipMemcpyAsync(h.data(), src, sizeof(double) * e.N, hipMemcpyDeviceToHost, c->stream));
  PD_HIP(hipStreamSynchronize(c->stream));
  for (int64_t i = 0; i < b.d.n; ++i) out[i] = h[internal_of(e, b.first + i)];
}

int pd_block_weighted_volume(void* ctx, int64_t env, int64_t block, double* m) {
  return run_guard(ctx, [&] {
    Ctx* c = idle(ctx);
    Env& e = get_env(c, env);
    if (!e.committed) throw PdError(PD_E_STATE, "env not committed");
    download_block_vec(c, e, get_block(e, block), e.dev.m, m);
  });
}

int pd_block_theta(void* ctx, int64_t env, int64_t block, double* theta) {
  return run_guard(ctx, [&] {
    Ctx* c = idle(ctx);
    Env& e = get_env(c, env);
    if (!e.committed) throw PdError(PD_E_STATE, "env not committed");
    download_block_vec(c, e, get_block(e, block), e.dev.theta, theta);
  });
}

int pd_block_edp(void* ctx, int64_t env, int64_t block, double* edp, int64_t rows) {
  return run_guard(ctx, [&] {
    Ctx* c = idle(ctx);
    Env& e = get_env(c, env);
    if (!e.committed) throw PdError(PD_E_STATE, "env not committed");
    HostBlock& b = get_block(e, block);
    if (!e.has_ep) throw PdError(PD_E_STATE, "env has no elastoplastic block");
    if (rows < b.maxdeg + 1) throw PdError(PD_E_ARG, "rows must be >= max degree + 1");
    BondsHost h;
    download_bonds(c, e, h);
    std::vector<double> ed(std::max<int64_t>(e.nslots, 1));
    PD_HIP(hipMemcpyAsync(ed.data(), e.dev.edp[e.edp_cur], sizeof(double) * e.nslots, hipMemcpyDeviceToHost,
                          c->stream));
    PD_HIP(hipStreamSynchronize(c->stream));
    std::vector<std::pair<int64_t, uint32_t>> row;
    for (int64_t i = 0; i < b.d.n; ++i) {
      row_of(e, h, b.first + i, row);
      const int64_t d = (int64_t)row.size();
      for (int64_t r = 0; r < rows; ++r) edp[i * rows + r] = (r < d) ? ed[row[r].second] : 0.0;
    }
  });
}

int64_t pd_log_rows(int64_t nsteps, int64_t every) {
  if (nsteps < 1 || every < 1) return 0;
  return nsteps / every + (every > 1 ? 1 : 0);
}

int pd_run(void* ctx, const int64_t* envs, int64_t nenv, int64_t nsteps, int64_t nuf) {
  return run_guard(ctx, [&] {
    if (nenv <= 0 || !envs) throw PdError(PD_E_ARG, "empty env list");
    if (nsteps < 0) throw PdError(PD_E_ARG, "nsteps must be >= 0");
    if (nuf < 1) throw PdError(PD_E_ARG, "neigh_update_freq must be >= 1");
    run_impl(idle(ctx), std::vector<int64_t>(envs, envs + nenv), nsteps, nuf, 0, nullptr);
  });
}

int pd_run_begin(void* ctx, const int64_t* envs, int64_t nenv, int64_t nsteps, int64_t nuf) {
  return run_guard(ctx, [&] {
    Ctx* c = idle(ctx);
    if (nenv <= 0 || !envs) throw PdError(PD_E_ARG, "empty env list");
    if (nsteps < 1) throw PdError(PD_E_ARG, "pd_run_begin needs nsteps >= 1");
    if (nuf < 1) throw PdError(PD_E_ARG, "neigh_update_freq must be >= 1");
    if (c->prof) throw PdError(PD_E_STATE, "pd_run_begin does not run under pd_ctx_profile");
    PD_HIP(hipSetDevice(c->device));
    if (!c->adm_stream) PD_HIP(hipStreamCreateWithFlags(&c->adm_stream, hipStreamNonBlocking));
    run_impl(c, std::vector<int64_t>(envs, envs + nenv), nsteps, nuf, 0, nullptr, true);
  });
}

int pd_run_query(void* ctx, int32_t* done) {
  return run_guard(ctx, [&] {
    Ctx* c = C(ctx);
    if (!done) throw PdError(PD_E_ARG, "null output pointer");
    if (!c->inflight.on) {
      *done = 1;
      return;
    }
    const hipError_t r = hipStreamQuery(c->run_stream);
    if (r == hipErrorNotReady) {
      *done = 0;
      return;
    }
    PD_HIP(r);
    *done = 1;
  });
}

int pd_run_end(void* ctx) {
  return run_guard(ctx, [&] {
    Ctx* c = C(ctx);
    if (!c->inflight.on) throw PdError(PD_E_STATE, "no run started by pd_run_begin is in flight");
    PD_HIP(hipSetDevice(c->device));
    Ctx::Inflight r = c->inflight;
    c->inflight = Ctx::Inflight{};
    c->stream = c->run_stream;
    c->mem.stream = c->run_stream;
    PD_HIP(hipStreamSynchronize(c->run_stream));
    end_run(c, r.ids, r.nsteps, r.nuf, r.nl);
  });
}

int pd_run_logged(void* ctx, const int64_t* envs, int64_t nenv, int64_t nsteps, int64_t nuf, int64_t every,
                  double* log) {
  return run_guard(ctx, [&] {
    if (nenv <= 0 || !envs) throw PdError(PD_E_ARG, "empty env list");
    if (nsteps < 0) throw PdError(PD_E_ARG, "nsteps must be >= 0");
    if (nuf < 1) throw PdError(PD_E_ARG, "neigh_update_freq must be >= 1");
    if (every < 1) throw PdError(PD_E_ARG, "average_prop_freq must be >= 1");
    if (!log && pd_log_rows(nsteps, every) > 0) throw PdError(PD_E_ARG, "null log buffer");
    run_impl(idle(ctx), std::vector<int64_t>(envs, envs + nenv), nsteps, nuf, every, log);
  });
}

int pd_env_observables(void* ctx, int64_t env, double* out) {
  return run_guard(ctx, [&] {
    Ctx* c = idle(ctx);
    if (!out) throw PdError(PD_E_ARG, "null output");
    PD_HIP(hipSetDevice(c->device));
    Env& e = get_env(c, env);
    if (!e.committed) throw PdError(PD_E_STATE, "env not committed");
    build_set(c, {env});
    ensure_cap(c->d_obs_rows, c->obs_rows_cap, 8, c->mem);
    hipStream_t s = c->stream;
    pd_launch_observables(s, 0, c->d_set, c->n_all, c->d_envs, c->d_obs_part, c->d_obs_ranges, 1, c->d_obs_rows);
    pd_launch_observables(s, 1, c->d_set, c->n_all, c->d_envs, c->d_obs_part, c->d_obs_ranges, 1, c->d_obs_rows);
    PD_HIP(hipGetLastError());
    PD_HIP(hipMemcpyAsync(out, c->d_obs_rows, sizeof(double) * 8, hipMemcpyDeviceToHost, s));
    PD_HIP(hipStreamSynchronize(s));
    out[0] = (double)e.time_step * e.dt;
    out[7] = out[7] / (double)e.N;
  });
}

int pd_env_observables_noalloc(void* ctx, int64_t env, double* out_dev) {
  return run_guard(ctx, [&] {
    Ctx* c = idle(ctx);
    if (!out_dev) throw PdError(PD_E_ARG, "null output");
    PD_HIP(hipSetDevice(c->device));
    Env& e = get_env(c, env);
    if (!e.committed) throw PdError(PD_E_STATE, "env not committed");
    // out_dev must be device memory a kernel of this GPU may write
    hipPointerAttribute_t pa{};
    const hipError_t ar = hipPointerGetAttributes(&pa, out_dev);
    if (ar != hipSuccess) (void)hipGetLastError();
    if (ar != hipSuccess || (pa.type != hipMemoryTypeDevice && pa.type != hipMemoryTypeManaged) ||
        pa.device != c->device)
      throw PdError(PD_E_ARG, "pd_env_observables_noalloc: out_dev must be device memory of the context's GPU");
    const size_t n = e.chunks.size();
    if (n > (size_t)kObsScratchChunks)
      throw PdError(PD_E_ARG, "pd_env_observables_noalloc: the env has " + std::to_string(n) +
                                  " chunks, more than the " + std::to_string(kObsScratchChunks) +
                                  " its fixed scratch holds; use pd_env_observables");
    pd_launch_observables_env(c->stream, e.d_chunks, (int)n, c->d_envs, e.slot, c->scratch, out_dev);
    PD_HIP(hipGetLastError());
  });
}

int pd_env_update_acc(void* ctx, int64_t env, const double* y, double* f) {
  return run_guard(ctx, [&] {
    Ctx* c = idle(ctx);
    PD_HIP(hipSetDevice(c->device));
    Env& e = get_env(c, env);
    if (!e.committed) throw PdError(PD_E_STATE, "env not committed");
    if (y) {
      upload_y(c, e, y);
      e.lists_current = false;
    }
    force_pass(c, e, e.chunks, 0);
    if (f) download_xyz(c, e, e.dev.f, f);
  });
}

int pd_force_density(void* ctx, int64_t env, int64_t block, const double* y, double* f, int64_t lo, int64_t hi) {
  return run_guard(ctx, [&] {
    Ctx* c = idle(ctx);
    PD_HIP(hipSetDevice(c->device));
    Env& e = get_env(c, env);
    if (!e.committed) throw PdError(PD_E_STATE, "env not committed");
    HostBlock& b = get_block(e, block);
    if (lo != b.first + 1 || hi != b.first + b.d.n) throw PdError(PD_E_ARG, "limits do not match the block");
    if (y) {
      upload_y(c, e, y);
      e.lists_current = false;
    }
    std::vector<Chunk> sub(e.chunks.begin() + b.chunk0, e.chunks.begin() + b.chunk0 + b.nchunks);
    // contacts are not part of force_density_T!: zero them out of this pass
    EnvDev saved = e.dev;
    e.dev.ncontacts = 0;
    upload_env_table(c, e);
    try {
      force_pass(c, e, sub, 0);
    } catch (...) {
      e.dev = saved;
      upload_env_table(c, e);
      throw;
    }
    e.dev = saved;
    upload_env_table(c, e);
    PD_HIP(hipStreamSynchronize(c->stream));
    std::vector<double> fb(3 * e.N);
    download_xyz(c, e, e.dev.f, fb.data());
    for (int64_t i = 0; i < b.d.n; ++i)
      for (int d = 0; d < 3; ++d) f[3 * (b.first + i) + d] += fb[3 * (b.first + i) + d];
  });
}

int pd_env_update_neighs(void* ctx, int64_t env) {
  return run_guard(ctx, [&] {
    Ctx* c = idle(ctx);
    Env& e = get_env(c, env);
    if (!e.committed) throw PdError(PD_E_STATE, "env not committed");
    contact_update(c, e);
    PD_HIP(hipStreamSynchronize(c->stream));
  });
}

static ContactDev& get_contact(Env& e, int64_t cidx) {
  if (cidx < 0 || cidx >= (int64_t)e.cdev.size()) throw PdError(PD_E_ARG, "contact model index out of range");
  return e.cdev[cidx];
}

int pd_contact_update(void* ctx, int64_t env, int64_t cidx, const double* y) {
  return run_guard(ctx, [&] {
    Ctx* c = idle(ctx);
    PD_HIP(hipSetDevice(c->device));
    Env& e = get_env(c, env);
    if (!e.committed) throw PdError(PD_E_STATE, "env not committed");
    ContactDev& cd = get_contact(e, cidx);
    if (y) upload_y(c, e, y);
    if (cd.self_contact) pd_launch_self_contact_update(c->stream, c->d_envs, e.slot, cd, (int)cidx);
    else pd_launch_contact_update(c->stream, c->d_envs, e.slot, e.dev, cd, (int)cidx);
    cpart_update(c, e);
    PD_HIP(hipGetLastError());
    PD_HIP(hipStreamSynchronize(c->stream));
    // the lists of the other models were built from earlier positions
    e.lists_current = (e.cdev.size() == 1) && (y != nullptr || e.lists_current);
  });
}

int pd_contact_neighs(void* ctx, int64_t env, int64_t cidx, int64_t* neighs, int64_t max_neighs) {
  return run_guard(ctx, [&] {
    Ctx* c = idle(ctx);
    PD_HIP(hipSetDevice(c->device));
    Env& e = get_env(c, env);
    if (!e.committed) throw PdError(PD_E_STATE, "env not committed");
    const ContactDev& cd = get_contact(e, cidx);
    const int64_t M = cd.max_neighs, n1 = cd.n1;
    if (!neighs || max_neighs != M)
      throw PdError(PD_E_ARG, "neighs must hold max_neighs x n1 entries with max_neighs = " + std::to_string(M));
    std::vector<int32_t> cnt((size_t)std::max<int64_t>(n1, 1)), own((size_t)std::max<int64_t>(n1 * M, 1));
    if (n1 > 0) {
      PD_HIP(hipMemcpyAsync(cnt.data(), cd.own_cnt, sizeof(int32_t) * n1, hipMemcpyDeviceToHost, c->stream));
      PD_HIP(hipMemcpyAsync(own.data(), cd.own, sizeof(int32_t) * n1 * M, hipMemcpyDeviceToHost, c->stream));
    }
    PD_HIP(hipStreamSynchronize(c->stream));
    for (int64_t i = 0; i < n1; ++i)
      for (int64_t k = 0; k < M; ++k) neighs[i * M + k] = (k < cnt[i]) ? (int64_t)own[i * M + k] + 1 : 0;
  });
}

int pd_contact_force(void* ctx, int64_t env, int64_t cidx, const double* y, double* f) {
  return run_guard(ctx, [&] {
    Ctx* c = idle(ctx);
    PD_HIP(hipSetDevice(c->device));
    Env& e = get_env(c, env);
    if (!e.committed) throw PdError(PD_E_STATE, "env not committed");
    const ContactDev& cd = get_contact(e, cidx);
    if (!f) throw PdError(PD_E_ARG, "null f");
    if (y) {
      upload_y(c, e, y);
      e.lists_current = false;
    }
    const size_t cls = DevMem::size_class(sizeof(double4) * (size_t)std::max<int64_t>(e.N, 1));
    double4* acc = static_cast<double4*>(c->mem.get(cls));
    try {
      upload_xyz(c, e, acc, f, false);
      pd_launch_contact_model(c->stream, c->d_envs, e.slot, (int)cidx, !cd.expo_is_int, acc, (int)e.N);
      PD_HIP(hipGetLastError());
      download_xyz(c, e, acc, f);
    } catch (...) {
      (void)hipStreamSynchronize(c->stream);
      c->mem.put(acc, cls);
      throw;
    }
    c->mem.put(acc, cls);
  });
}

int pd_env_download(void* ctx, int64_t env, double* y, double* v, double* f) {
  return run_guard(ctx, [&] {
    Ctx* c = idle(ctx);
    Env& e = get_env(c, env);
    if (!e.committed) throw PdError(PD_E_STATE, "env not committed");
    PD_HIP(hipStreamSynchronize(c->stream));
    if (y) download_xyz(c, e, e.dev.y[e.cur], y);
    if (v) download_xyz(c, e, e.dev.v, v);
    if (f) download_xyz(c, e, e.dev.f, f);
  });
}

int pd_env_upload(void* ctx, int64_t env, const double* y, const double* v, const double* f) {
  return run_guard(ctx, [&] {
    Ctx* c = idle(ctx);
    Env& e = get_env(c, env);
    if (!e.committed) throw PdError(PD_E_STATE, "env not committed");
    PD_HIP(hipStreamSynchronize(c->stream));
    if (y) {
      upload_y(c, e, y);
      e.lists_current = false;
    }
    if (v) upload_xyz(c, e, e.dev.v, v, false);
    if (f) upload_xyz(c, e, e.dev.f, f, false);
  });
}

int pd_env_damage(void* ctx, int64_t env, double* damage) {
  return run_guard(ctx, [&] {
    Ctx* c = idle(ctx);
    Env& e = get_env(c, env);
    if (!e.committed) throw PdError(PD_E_STATE, "env not committed");
    BondsHost h;
    download_bonds(c, e, h);
    for (int64_t p = 0; p < e.N; ++p) {
      int64_t cnt = 0;
      for (int k = 0; k < h.deg[p]; ++k) cnt += h.state[pd_slot(h.pslot0[p], k, e.lgG)];
      // cal_damage: 1 .- intact ./ intact0 (intact0 = initial family size)
      damage[reference_of(e, p)] = 1.0 - (double)cnt / (double)h.deg[p];
    }
  });
}

int pd_ctx_profile(void* ctx, int32_t enable) {
  return run_guard(ctx, [&] {
    Ctx* c = idle(ctx);
    drain_profile(c);
    for (int k = 0; k < PD_PROF_NCLASS; ++k) { c->prof_ms[k] = 0.0; c->prof_n[k] = 0; }
    c->prof = enable != 0;
  });
}

int pd_ctx_profile_read(void* ctx, int32_t cls, double* total_ms, int64_t* launches) {
  return run_guard(ctx, [&] {
    Ctx* c = idle(ctx);
    if (cls < 0 || cls >= PD_PROF_NCLASS) throw PdError(PD_E_ARG, "unknown kernel class");
    drain_profile(c);
    if (total_ms) *total_ms = c->prof_ms[cls];
    if (launches) *launches = c->prof_n[cls];
  });
}

int pd_env_time_step(void* ctx, int64_t env, int64_t* ts) {
  return run_guard(ctx, [&] { *ts = get_env(C(ctx), env).time_step; });
}

int pd_env_state_size(void* ctx, int64_t env, int64_t* bytes) {
  return run_guard(ctx, [&] {
    Env& e = get_env(C(ctx), env);
    if (!e.committed) throw PdError(PD_E_STATE, "env not committed");
    if (!bytes) throw PdError(PD_E_ARG, "null output");
    *bytes = state_bytes(e);
  });
}

int pd_env_state_save(void* ctx, int64_t env, void* buf, int64_t bytes) {
  return run_guard(ctx, [&] {
    Ctx* c = idle(ctx);
    PD_HIP(hipSetDevice(c->device));
    Env& e = get_env(c, env);
    if (!e.committed) throw PdError(PD_E_STATE, "env not committed");
    if (!buf || bytes < state_bytes(e)) throw PdError(PD_E_ARG, "state buffer too small");
    char* p = static_cast<char*>(buf);
    int64_t* hdr = reinterpret_cast<int64_t*>(p);
    std::memset(hdr, 0, 16 * 8);
    hdr[0] = kStateMagic; hdr[1] = kStateVersion; hdr[2] = e.N; hdr[3] = e.nslots;
    hdr[4] = (int64_t)e.blocks.size(); hdr[5] = e.precision; hdr[6] = e.lgG; hdr[7] = e.has_ep ? 1 : 0;
    hdr[8] = e.time_step;
    hdr[9] = e.tiled ? 1 : 0;
    p += 16 * 8;
    std::vector<BlockDev> bd(e.blocks.size());
    hipStream_t s = c->stream;
    PD_HIP(hipMemcpyAsync(bd.data(), e.dev.blocks, sizeof(BlockDev) * bd.size(), hipMemcpyDeviceToHost, s));
    PD_HIP(hipStreamSynchronize(s));
    for (size_t b = 0; b < bd.size(); ++b) reinterpret_cast<int32_t*>(p)[b] = bd[b].deformed;
    p += ((4 * (int64_t)bd.size() + 16 * 8 + 7) / 8 * 8) - 16 * 8;
    const size_t n4 = sizeof(double4) * (size_t)e.N, n1 = sizeof(double) * (size_t)e.N;
    PD_HIP(hipMemcpyAsync(p, e.dev.y[e.cur], n4, hipMemcpyDeviceToHost, s)); p += n4;
    PD_HIP(hipMemcpyAsync(p, e.dev.v, n4, hipMemcpyDeviceToHost, s)); p += n4;
    PD_HIP(hipMemcpyAsync(p, e.dev.f, n4, hipMemcpyDeviceToHost, s)); p += n4;
    PD_HIP(hipMemcpyAsync(p, e.dev.aux, n4, hipMemcpyDeviceToHost, s)); p += n4;
    PD_HIP(hipMemcpyAsync(p, e.dev.theta, n1, hipMemcpyDeviceToHost, s)); p += n1;
    PD_HIP(hipMemcpyAsync(p, e.dev.td, n1, hipMemcpyDeviceToHost, s)); p += n1;
    PD_HIP(hipMemcpyAsync(p, e.dev.state, (size_t)e.nslots, hipMemcpyDeviceToHost, s));
    p += (e.nslots + 7) / 8 * 8;
    if (e.has_ep) PD_HIP(hipMemcpyAsync(p, e.dev.edp[e.edp_cur], 8 * (size_t)e.nslots, hipMemcpyDeviceToHost, s));
    PD_HIP(hipStreamSynchronize(s));
  });
}

int pd_env_state_load(void* ctx, int64_t env, const void* buf, int64_t bytes) {
  return run_guard(ctx, [&] {
    Ctx* c = idle(ctx);
    PD_HIP(hipSetDevice(c->device));
    Env& e = get_env(c, env);
    if (!e.committed) throw PdError(PD_E_STATE, "env not committed");
    if (!buf || bytes < 16 * 8) throw PdError(PD_E_ARG, "state buffer too small");
    const char* p = static_cast<const char*>(buf);
    const int64_t* hdr = reinterpret_cast<const int64_t*>(p);
    if (hdr[0] != kStateMagic || hdr[1] != kStateVersion) throw PdError(PD_E_ARG, "not a state snapshot of this library");
    if (hdr[2] != e.N || hdr[3] != e.nslots || hdr[4] != (int64_t)e.blocks.size() || hdr[5] != e.precision ||
        hdr[6] != e.lgG || hdr[7] != (e.has_ep ? 1 : 0) || hdr[9] != (e.tiled ? 1 : 0))
      throw PdError(PD_E_ARG, "snapshot was taken from an env of another shape, form or schedule");
    if (bytes < state_bytes(e)) throw PdError(PD_E_ARG, "state buffer too small");
    p += 16 * 8;
    hipStream_t s = c->stream;
    std::vector<BlockDev> bd(e.blocks.size());
    PD_HIP(hipMemcpyAsync(bd.data(), e.dev.blocks, sizeof(BlockDev) * bd.size(), hipMemcpyDeviceToHost, s));
    PD_HIP(hipStreamSynchronize(s));
    bool any_undeformed = false;
    for (size_t b = 0; b < bd.size(); ++b) {
      bd[b].deformed = reinterpret_cast<const int32_t*>(p)[b];
      any_undeformed = any_undeformed || !bd[b].deformed;
    }
    PD_HIP(hipMemcpyAsync(e.dev.blocks, bd.data(), sizeof(BlockDev) * bd.size(), hipMemcpyHostToDevice, s));
    p += ((4 * (int64_t)bd.size() + 16 * 8 + 7) / 8 * 8) - 16 * 8;
    const size_t n4 = sizeof(double4) * (size_t)e.N, n1 = sizeof(double) * (size_t)e.N;
    PD_HIP(hipMemcpyAsync(e.dev.y[e.cur], p, n4, hipMemcpyHostToDevice, s)); p += n4;
    PD_HIP(hipMemcpyAsync(e.dev.v, p, n4, hipMemcpyHostToDevice, s)); p += n4;
    PD_HIP(hipMemcpyAsync(e.dev.f, p, n4, hipMemcpyHostToDevice, s)); p += n4;
    PD_HIP(hipMemcpyAsync(e.dev.aux, p, n4, hipMemcpyHostToDevice, s)); p += n4;
    PD_HIP(hipMemcpyAsync(e.dev.theta, p, n1, hipMemcpyHostToDevice, s)); p += n1;
    PD_HIP(hipMemcpyAsync(e.dev.td, p, n1, hipMemcpyHostToDevice, s)); p += n1;
    PD_HIP(hipMemcpyAsync(e.dev.state, p, (size_t)e.nslots, hipMemcpyHostToDevice, s));
    p += (e.nslots + 7) / 8 * 8;
    if (e.has_ep) {
      // both buffers: the fast forms skip writing plastic extensions that stay 0
      PD_HIP(hipMemcpyAsync(e.dev.edp[0], p, 8 * (size_t)e.nslots, hipMemcpyHostToDevice, s));
      PD_HIP(hipMemcpyAsync(e.dev.edp[1], p, 8 * (size_t)e.nslots, hipMemcpyHostToDevice, s));
    }
    e.time_step = hdr[8];
    e.maybe_undeformed = any_undeformed;
    e.block_latched.assign(e.blocks.size(), 0);  // the loaded latches are read back when next tested
    e.block_rigid_ok.assign(e.blocks.size(), 0);
    e.lists_current = false;
    upload_env_table(c, e);
    if (e.tiled) pd_launch_pk_sync(s, e.d_chunks, (int)e.chunks.size(), c->d_envs);
    PD_HIP(hipStreamSynchronize(s));
  });
}

int pd_env_minimize(void* ctx, int64_t env, double step_size, double drag, int64_t max_iter, double x_tol,
                    double f_tol, int64_t* iters, double* f_max, double* x_max) {
  return run_guard(ctx, [&] {
    Ctx* c = idle(ctx);
    if (max_iter < 1) throw PdError(PD_E_ARG, "max_iter must be >= 1");
    if (!(step_size > 0.0) || !std::isfinite(step_size)) throw PdError(PD_E_ARG, "step_size must be > 0");
    PD_HIP(hipSetDevice(c->device));
    Env& e = get_env(c, env);
    if (!e.committed) throw PdError(PD_E_STATE, "env not committed");
    {
      std::vector<uint8_t> held(e.N, 0);
      for (const HostBc& b : e.bcs)
        for (int64_t r = 0; r < e.N; ++r) held[r] |= b.mask[r];
      bool any_free = false;
      for (uint8_t h : held) any_free = any_free || !h;
      if (!any_free) throw PdError(PD_E_ARG, "no particle is free of boundary conditions");
    }
    const std::vector<int64_t> ids{env};
    const int prec = e.precision;
    build_set(c, ids, true);
    hipStream_t s = c->stream;
    const int nl = prepare_latch(c, ids);
    EnvDev* T = c->d_envs;
    Chunk* d = c->d_set;
    const double a = e.blocks[0].d.particle_size / 10.0;  // ps of material block 1 (minimize.jl:127-136)
    char* buf = nullptr;
    PD_HIP(hipMalloc(&buf, sizeof(double4) * (size_t)e.N + 16));
    double4* opt = reinterpret_cast<double4*>(buf);
    auto* maxes = reinterpret_cast<unsigned long long*>(buf + sizeof(double4) * (size_t)e.N);
    try {
      PD_HIP(hipMemsetAsync(opt, 0, sizeof(double4) * (size_t)e.N, s));
      pd_launch_qs_bc(s, d, c->n_all, T);
      int64_t it = 0;
      unsigned long long hm[2] = {0, 0};
      for (int64_t i = 1; i <= max_iter; ++i) {
        // update_acc! (solvers.jl:320-336): the env's own force passes, no Verlet
        if (nl) launch_latch(c, s, nl);
        if (c->set_has_theta) pd_launch_theta(s, prec, d + c->off_theta, c->n_theta, T);
        if (c->set_has_ep) pd_launch_td(s, prec, d + c->off_td, c->n_td, T);
        for (int k = 0; k < 3; ++k)
          if (c->n_force[k]) pd_launch_force(s, k, prec, d + c->off_force[k], c->n_force[k], T, 0, c->set_gpow);
        if (c->set_has_ep) {
          pd_launch_book(s, T, e.d_self, 1, PD_B_FLIP_EDP);
          e.edp_cur ^= 1;
        }
        PD_HIP(hipMemsetAsync(maxes, 0, 16, s));
        pd_launch_qs_update(s, d, c->n_all, T, opt, step_size, drag, a, maxes);
        PD_HIP(hipMemcpyAsync(hm, maxes, 16, hipMemcpyDeviceToHost, s));
        PD_HIP(hipStreamSynchronize(s));
        it = i;
        double fm, xm;
        std::memcpy(&fm, &hm[0], 8);
        std::memcpy(&xm, &hm[1], 8);
        if (f_max) *f_max = fm;
        if (x_max) *x_max = xm;
        if ((fm < f_tol) && (xm < x_tol)) break;
      }
      if (iters) *iters = it;
      // check! of every BC, then time_step += 1 (minimize.jl:225-229)
      pd_launch_book(s, T, e.d_self, 1, PD_B_STEP);
      if (e.tiled) pd_launch_pk_sync(s, e.d_chunks, (int)e.chunks.size(), T);
      PD_HIP(hipGetLastError());
      PD_HIP(hipStreamSynchronize(s));
    } catch (...) {
      (void)hipStreamSynchronize(s);
      (void)hipFree(buf);
      throw;
    }
    PD_HIP(hipFree(buf));
    e.time_step += 1;
    e.lists_current = false;
    refresh_latch(c, ids, nl);
  });
}

int pd_env_check_finite(void* ctx, int64_t env) {
  int rc = run_guard(ctx, [&] {
    Ctx* c = idle(ctx);
    Env& e = get_env(c, env);
    if (!e.committed) throw PdError(PD_E_STATE, "env not committed");
    std::vector<double> a(3 * e.N);
    const double4* srcs[3] = {e.dev.y[e.cur], e.dev.v, e.dev.f};
    const char* names[3] = {"y", "v", "f"};
    for (int q = 0; q < 3; ++q) {
      download_xyz(c, e, srcs[q], a.data());
      for (double x : a)
        if (!std::isfinite(x)) throw PdError(PD_E_NONFINITE, std::string("nan/inf in env.") + names[q]);
    }
  });
  return rc;
}

}  // extern "C"
