// Contact 1-2 neighbour search on the GPU (update_contact_neighs! RM12,
// src/contacts/contacts.jl:434-476, with collision_box :403-416).
//
// Semantics kept from the reference: the AABB intersection grown by `distance`
// gates both sides with strict inequalities; a pair (i, j) is a contact when
//   (a1-a2)*(a1-a2) + (b1-b2)*(b1-b2) + (c1-c2)*(c1-c2) < distance^2
// (block-1 particle first); each block-1 particle keeps the `max_neighs`
// LARGEST block-2 member indices (contacts.jl:468-469).  The brute-force
// O(n1 n2) scan is replaced by a cell list (cells >= distance wide) over the
// in-box block-2 members.  For the gather-form force kernel the lists are also
// transposed: each block-2 member gets its block-1 partners.  The own lists are
// sorted descending and the transposed lists ascending, the order in which the
// reference's loop over i scatters into f1[:, i] and f2[:, j]; the cell-list
// fill uses atomics, so without the sorts the summation order, and the last bits
// of a run, would vary between rebuilds.
#include "pd_device.h"
#include "pd_kernels.h"

namespace {

constexpr int kRedBlocks = 128;

__device__ __forceinline__ const double4* cur_y(const EnvDev& E) { return E.y[E.cur]; }

__device__ __forceinline__ double red_op(int d, double a, double b) {
  return ((d / 3) % 2 == 0) ? fmin(a, b) : fmax(a, b);
}

// stage 1: per-block partial min/max over both member sets (block b of nb)
__device__ void cbox1_body(const EnvDev& E, const ContactDev& C, double* __restrict__ part, int b, int nb) {
  const double4* y = cur_y(E);
  double v[12];
  for (int d = 0; d < 3; ++d) {
    v[d] = INFINITY; v[3 + d] = -INFINITY; v[6 + d] = INFINITY; v[9 + d] = -INFINITY;
  }
  const int stride = nb * blockDim.x;
  for (int i = b * blockDim.x + threadIdx.x; i < C.n1; i += stride) {
    const double4 p = y[C.idx1[i]];
    v[0] = fmin(v[0], p.x); v[1] = fmin(v[1], p.y); v[2] = fmin(v[2], p.z);
    v[3] = fmax(v[3], p.x); v[4] = fmax(v[4], p.y); v[5] = fmax(v[5], p.z);
  }
  for (int i = b * blockDim.x + threadIdx.x; i < C.n2; i += stride) {
    const double4 p = y[C.idx2[i]];
    v[6] = fmin(v[6], p.x); v[7] = fmin(v[7], p.y); v[8] = fmin(v[8], p.z);
    v[9] = fmax(v[9], p.x); v[10] = fmax(v[10], p.y); v[11] = fmax(v[11], p.z);
  }
  __shared__ double sh[12][4];
  for (int o = 32; o > 0; o >>= 1)
    for (int d = 0; d < 12; ++d) v[d] = red_op(d, v[d], __shfl_down(v[d], o));
  const int wv = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0)
    for (int d = 0; d < 12; ++d) sh[d][wv] = v[d];
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int d = 0; d < 12; ++d) {
      double r = sh[d][0];
      for (int q = 1; q < 4; ++q) r = red_op(d, r, sh[d][q]);
      part[b * 12 + d] = r;
    }
  }
}

__global__ __launch_bounds__(256) void k_cbox1(const EnvDev* __restrict__ envs, int slot, int cidx,
                                               double* __restrict__ part) {
  const EnvDev& E = envs[slot];
  cbox1_body(E, E.contacts[cidx], part, blockIdx.x, gridDim.x);
}

// stage 2 (one wavefront): box, overlap flag and cell grid (<= ncell_cap cells)
__device__ void cbox2_body(const ContactDev& C, const double* __restrict__ part, int nparts) {
  double v[12];
  for (int d = 0; d < 12; ++d) v[d] = ((d / 3) % 2 == 0) ? INFINITY : -INFINITY;
  for (int b = threadIdx.x; b < nparts; b += 64)
    for (int d = 0; d < 12; ++d) v[d] = red_op(d, v[d], part[b * 12 + d]);
  for (int o = 32; o > 0; o >>= 1)
    for (int d = 0; d < 12; ++d) v[d] = red_op(d, v[d], __shfl_down(v[d], o));
  if (threadIdx.x != 0) return;
  const double dist = C.distance;
  double bmin[3], bmax[3];
  for (int d = 0; d < 3; ++d) {
    bmin[d] = fmax(v[d] - dist, v[6 + d] - dist);
    bmax[d] = fmin(v[3 + d] + dist, v[9 + d] + dist);
  }
  const double vol = ((bmax[0] - bmin[0]) * (bmax[1] - bmin[1])) * (bmax[2] - bmin[2]);
  double* box = C.box;
  for (int d = 0; d < 3; ++d) { box[d] = bmin[d]; box[3 + d] = bmax[d]; }
  const bool ok = !(vol < 0.0) && C.n1 > 0 && C.n2 > 0;
  box[6] = ok ? 1.0 : 0.0;
  double cw = dist * (1.0 + 9.313225746154785e-10);
  int nc[3] = {1, 1, 1};
  for (int it = 0; it < 64; ++it) {
    long long tot = 1;
    for (int d = 0; d < 3; ++d) {
      const double ext = ok ? (bmax[d] - bmin[d]) : 0.0;
      nc[d] = (int)fmin(floor(ext / cw) + 1.0, 1.0e6);
      if (nc[d] < 1) nc[d] = 1;
      tot *= nc[d];
    }
    if (tot <= C.ncell_cap) break;
    cw *= 1.5;
  }
  box[7] = cw;
  box[8] = nc[0]; box[9] = nc[1]; box[10] = nc[2];
}

__global__ __launch_bounds__(64) void k_cbox2(const EnvDev* __restrict__ envs, int slot, int cidx,
                                              const double* __restrict__ part, int nparts) {
  cbox2_body(envs[slot].contacts[cidx], part, nparts);
}

__global__ void k_izero(int32_t* __restrict__ a, int n) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) a[i] = 0;
}

__device__ __forceinline__ bool in_box(const double4& p, const double* box) {
  return (p.x > box[0]) && (p.y > box[1]) && (p.z > box[2]) && (p.x < box[3]) && (p.y < box[4]) && (p.z < box[5]);
}

__device__ __forceinline__ int cell_id(const double4& p, const double* box, int* cc) {
  const double cw = box[7];
  const int nc[3] = {(int)box[8], (int)box[9], (int)box[10]};
  const double pv[3] = {p.x, p.y, p.z};
  for (int d = 0; d < 3; ++d) {
    int c = (int)floor((pv[d] - box[d]) / cw);
    cc[d] = c < 0 ? 0 : (c >= nc[d] ? nc[d] - 1 : c);
  }
  return cc[0] + nc[0] * (cc[1] + nc[1] * cc[2]);
}

// bin in-box block-2 members
__device__ void cbin_body(const EnvDev& E, const ContactDev& C, int j) {
  if (j >= C.n2) return;
  const double* box = C.box;
  const double4 p = cur_y(E)[C.idx2[j]];
  int c = -1;
  if (box[6] != 0.0 && in_box(p, box)) {
    int cc[3];
    c = cell_id(p, box, cc);
    atomicAdd(&C.cell_count[c], 1);
  }
  C.cell_of[j] = c;
}

__global__ void k_cbin(const EnvDev* __restrict__ envs, int slot, int cidx) {
  const EnvDev& E = envs[slot];
  cbin_body(E, E.contacts[cidx], blockIdx.x * blockDim.x + threadIdx.x);
}

// cursor = the second half of cell_start (ncell_cap entries)
__device__ void cfill_body(const ContactDev& C, int j) {
  if (j >= C.n2) return;
  const int c = C.cell_of[j];
  if (c < 0) return;
  int32_t* cursor = C.cell_start + C.ncell_cap;
  C.cell_items[C.cell_start[c] + atomicAdd(&cursor[c], 1)] = j;
}

__global__ void k_cfill(const EnvDev* __restrict__ envs, int slot, int cidx) {
  cfill_body(envs[slot].contacts[cidx], blockIdx.x * blockDim.x + threadIdx.x);
}

__device__ __forceinline__ int wave_add(int v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}
__device__ __forceinline__ int wave_min(int v) {
  for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o));
  return v;
}
__device__ __forceinline__ int wave_max(int v) {
  for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o));
  return v;
}

// Visits every block-2 member within `distance` of p (27-cell stencil) with the
// 64 lanes of the calling wavefront: lanes 0..26 read the ranges of the 27
// cells at once, and the cells' members, concatenated in stencil order, are
// dealt to the lanes 64 at a time.  f(hit, j) is called by all lanes together
// (wave-uniform trip count) with hit = the lane's member j is a contact.
template <typename F>
__device__ __forceinline__ void wave_candidates(const ContactDev& C, const double4* y, const double4& p,
                                                const double* box, int lane, F&& f) {
  int cc[3];
  cell_id(p, box, cc);
  const int nc[3] = {(int)box[8], (int)box[9], (int)box[10]};
  const double d2max = C.distance * C.distance;
  int my_s = 0, my_n = 0;
  if (lane < 27) {
    const int xx = cc[0] + lane % 3 - 1, yy = cc[1] + (lane / 3) % 3 - 1, zz = cc[2] + lane / 9 - 1;
    if (xx >= 0 && xx < nc[0] && yy >= 0 && yy < nc[1] && zz >= 0 && zz < nc[2]) {
      const int cell = xx + nc[0] * (yy + nc[1] * zz);
      my_s = C.cell_start[cell];
      my_n = C.cell_count[cell];
    }
  }
  int incl = my_n;  // inclusive scan of the cell sizes over the lanes
  for (int o = 1; o < 64; o <<= 1) {
    const int v = __shfl_up(incl, o);
    if (lane >= o) incl += v;
  }
  const int total = __shfl(incl, 63);
  const int excl = incl - my_n;
  for (int t0 = 0; t0 < total; t0 += 64) {
    const int t = t0 + lane;
    // the cell holding concatenated index t: the last of lanes 0..26 whose
    // first index is <= t (empty cells share their successor's first index)
    int c = 0;
    for (int b = 16; b > 0; b >>= 1) {
      const int cand = c + b;
      const int ex = __shfl(excl, cand < 27 ? cand : 26);
      if (cand < 27 && ex <= t) c = cand;
    }
    const int cs = __shfl(my_s, c), ce = __shfl(excl, c);
    bool hit = false;
    int j = -1;
    if (t < total) {
      j = C.cell_items[cs + (t - ce)];
      const double4 q = y[C.idx2[j]];
      const double a = p.x - q.x, b = p.y - q.y, d = p.z - q.z;
      hit = (a * a + b * b + d * d < d2max);
    }
    f(hit, j);
  }
}

// Rank of this lane's value among the n (<= 64) distinct values held by lanes
// 0..n-1: the number of them that sort before it (descending if DESC).
template <bool DESC>
__device__ __forceinline__ int wave_rank(int32_t v, int n) {
  int r = 0;
  for (int m = 0; m < n; ++m) {
    const int32_t w = __shfl(v, m);
    r += DESC ? (w > v) : (w < v);
  }
  return r;
}

// Descending bitonic sort of buf[0, n) (n <= kCandCap) by one wavefront; the
// tail up to the next power of two (>= 64) is padded with -1.
constexpr int kCandCap = 2048;
__device__ void wave_sort_desc(int32_t* buf, int n, int lane) {
  int P = 64;
  while (P < n) P <<= 1;
  for (int t = n + lane; t < P; t += 64) buf[t] = -1;
  wave_sync();
  for (int k = 2; k <= P; k <<= 1) {
    for (int jj = k >> 1; jj > 0; jj >>= 1) {
      for (int t = lane; t < P; t += 64) {
        const int l = t ^ jj;
        if (l > t) {
          const int32_t a = buf[t], b = buf[l];
          const bool lead = (t & k) == 0;
          if (lead ? (a < b) : (a > b)) { buf[t] = b; buf[l] = a; }
        }
      }
      wave_sync();
    }
  }
}

// One wavefront per block-1 member i: the contact list of i is the max_neighs
// largest block-2 member ids within `distance`, stored descending.  Candidates
// are compacted in scan order into the wavefront's LDS buffer and sorted there;
// with more than kCandCap candidates the max_neighs-th largest id is found by
// bisection over the candidates' id range first.
__device__ void csearch_body(const EnvDev& E, const ContactDev& C, int i, int lane, int32_t* buf) {
  if (i >= C.n1) return;  // whole wavefront; the kernels use no workgroup barrier
  const double* box = C.box;
  const double4* y = cur_y(E);
  const double4 p = y[C.idx1[i]];
  const int M = C.max_neighs;
  int32_t* own = C.own + (size_t)i * M;
  int cnt = 0;
  if (M > 0 && box[6] != 0.0 && in_box(p, box)) {
    int total = 0, lo_id = INT_MAX, hi_id = -1;
    wave_candidates(C, y, p, box, lane, [&](bool hit, int j) {
      const uint64_t m = __ballot(hit);
      if (hit) {
        const int pos = total + lanes_below(m);
        if (pos < kCandCap) buf[pos] = j;
        lo_id = min(lo_id, j);
        hi_id = max(hi_id, j);
      }
      total += __popcll(m);
    });
    if (total <= 64) {
      // few candidates: descending ranks in registers, the M largest kept
      wave_sync();  // buf[] written by other lanes above
      const int32_t v = lane < total ? buf[lane] : -1;
      const int r = wave_rank<true>(v, total);
      cnt = min(total, M);
      if (lane < total && r < M) own[r] = v;
    } else if (total <= kCandCap) {
      wave_sort_desc(buf, total, lane);
      cnt = min(total, M);
      for (int k = lane; k < cnt; k += 64) own[k] = buf[k];
    } else {
      // count(>= lo) >= M > count(>= hi); ids are distinct, so lo converges to
      // the M-th largest id and exactly M candidates are >= lo
      int lo = wave_min(lo_id), hi = wave_max(hi_id) + 1;
      if (total > M) {
        while (hi - lo > 1) {
          const int mid = lo + ((hi - lo) >> 1);
          int c = 0;
          wave_candidates(C, y, p, box, lane, [&](bool hit, int j) { c += (hit && j >= mid); });
          if (wave_add(c) >= M) lo = mid; else hi = mid;
        }
      }
      wave_candidates(C, y, p, box, lane, [&](bool hit, int j) {
        const bool keep = hit && j >= lo;
        const uint64_t m = __ballot(keep);
        if (keep) own[cnt + lanes_below(m)] = j;
        cnt += __popcll(m);
      });
      if (cnt <= kCandCap) {
        for (int k = lane; k < cnt; k += 64) buf[k] = own[k];
        wave_sort_desc(buf, cnt, lane);
        for (int k = lane; k < cnt; k += 64) own[k] = buf[k];
      } else if (lane == 0) {
        for (int a = 1; a < cnt; ++a) {  // descending insertion sort (max_neighs > kCandCap)
          const int v = own[a];
          int b = a - 1;
          while (b >= 0 && own[b] < v) { own[b + 1] = own[b]; --b; }
          own[b + 1] = v;
        }
      }
    }
  }
  if (lane == 0) C.own_cnt[i] = cnt;
  wave_sync();  // own[] written by other lanes (or lane 0 alone) above
  for (int k = lane; k < cnt; k += 64) atomicAdd(&C.rev_cnt[own[k]], 1);
}


// One wavefront per block-1 member: append i to the transposed list of each of
// its contacts (order fixed afterwards by the sort below).
__device__ void crev_fill_body(const ContactDev& C, int i, int lane) {
  if (i >= C.n1) return;
  const int cnt = C.own_cnt[i];
  const int32_t* own = C.own + (size_t)i * C.max_neighs;
  for (int k = lane; k < cnt; k += 64) {
    const int j = own[k];
    C.rev[C.rev_off[j] + atomicAdd(&C.rev_cursor[j], 1)] = i;
  }
}

__global__ __launch_bounds__(256) void k_crev_fill(const EnvDev* __restrict__ envs, int slot, int cidx) {
  crev_fill_body(envs[slot].contacts[cidx], blockIdx.x * 4 + (threadIdx.x >> 6), threadIdx.x & 63);
}

// One wavefront per block-2 member: its transposed list sorted ascending.
__device__ void crev_sort_body(const ContactDev& C, int j, int lane, int32_t* buf) {
  if (j >= C.n2) return;  // whole wavefront; no workgroup barrier below
  int32_t* r = C.rev + C.rev_off[j];
  const int n = C.rev_cnt[j];
  if (n <= 1) return;
  if (n <= 64) {  // ascending ranks in registers
    const int32_t v = lane < n ? r[lane] : INT_MAX;
    const int rk = wave_rank<false>(v, n);
    wave_sync();  // every lane has read r[] before any lane overwrites it
    if (lane < n) r[rk] = v;
  } else if (n <= kCandCap) {
    for (int k = lane; k < n; k += 64) buf[k] = r[k];
    wave_sort_desc(buf, n, lane);
    for (int k = lane; k < n; k += 64) r[k] = buf[n - 1 - k];
  } else if (lane == 0) {
    for (int a = 1; a < n; ++a) {
      const int v = r[a];
      int b = a - 1;
      while (b >= 0 && r[b] > v) { r[b + 1] = r[b]; --b; }
      r[b + 1] = v;
    }
  }
}

// Work lists: most members have nothing to search (outside the box) or to
// sort (fewer than two partners), so one lane per member appends those that
// do to wlist (one atomic per wavefront), and a grid sized from the member
// count walks the list, one member per wavefront at a time (search: a
// workgroup per 16 block-1 members, at most 4096; sort: per 64 block-2
// members, at most 1024, as few members have two partners).  Each member is
// processed exactly as before; the order of the list does not enter any result.

__device__ __forceinline__ void wlist_append(const ContactDev& C, int which, bool need, int m) {
  const uint64_t b = __ballot(need);
  if (b == 0) return;
  const int lane = threadIdx.x & 63;
  const int leader = __builtin_ctzll(b);
  int base = 0;
  if (lane == leader) base = atomicAdd(&C.wcount[which], __popcll(b));
  base = __shfl(base, leader);
  if (need) C.wlist[base + lanes_below(b)] = m;
}

// block-1 member i: an empty list at once, or queued for the search
__device__ void csearch_mark(const EnvDev& E, const ContactDev& C, int i) {
  bool need = false;
  if (i < C.n1) {
    need = C.max_neighs > 0 && C.box[6] != 0.0 && in_box(cur_y(E)[C.idx1[i]], C.box);
    if (!need) C.own_cnt[i] = 0;
  }
  wlist_append(C, 0, need, i);
}

// block-2 member j with two or more partners: queued for the sort
__device__ void crev_mark(const ContactDev& C, int j) {
  wlist_append(C, 1, j < C.n2 && C.rev_cnt[j] > 1, j);
}

// wavefront gw of nw walks list `which`
template <bool SORT>
__device__ void walk_list(const EnvDev& E, const ContactDev& C, int gw, int nw, int lane, int32_t* buf) {
  const int n = C.wcount[SORT ? 1 : 0];
  for (int k = gw; k < n; k += nw) {
    const int m = C.wlist[k];
    if (SORT) crev_sort_body(C, m, lane, buf);
    else csearch_body(E, C, m, lane, buf);
  }
}

int list_blocks(int n, bool sort) {
  if (n <= 0) return 0;
  const int per = sort ? 64 : 16, cap = sort ? 1024 : 4096;
  const int b = (n + per - 1) / per;
  return b < cap ? b : cap;
}

__global__ __launch_bounds__(256) void k_csearch_mark(const EnvDev* __restrict__ envs, int slot, int cidx) {
  const EnvDev& E = envs[slot];
  csearch_mark(E, E.contacts[cidx], blockIdx.x * 256 + threadIdx.x);
}

__global__ __launch_bounds__(256) void k_crev_mark(const EnvDev* __restrict__ envs, int slot, int cidx) {
  crev_mark(envs[slot].contacts[cidx], blockIdx.x * 256 + threadIdx.x);
}

template <bool SORT>
__global__ __launch_bounds__(256) void k_walk(const EnvDev* __restrict__ envs, int slot, int cidx) {
  __shared__ int32_t cand[4][kCandCap];
  const int wv = threadIdx.x >> 6;
  const EnvDev& E = envs[slot];
  walk_list<SORT>(E, E.contacts[cidx], blockIdx.x * 4 + wv, gridDim.x * 4, threadIdx.x & 63, cand[wv]);
}

// Self contact (update_contact_neighs! for ContactModel11, contacts.jl:539-591).
// Its merge walk advances through the family column while the column entry is
// smaller than the candidate, expecting the column ascending; GeneralMaterial
// stores it descending with zero padding (general_material.jl:88-90), so the
// walk can only ever match the column's first entry, the largest-index family
// member.  A member therefore gets at most one contact: that particle, when
// the bond to it is broken and the pair is closer than equi_dist (such a pair
// always lies in neighbouring cells of width >= distance >= equi_dist).
// Bond rows are sorted by reference index, so the largest member is bond
// deg - 1.  One thread per member.
__global__ void k_cself(const EnvDev* __restrict__ envs, int slot, int cidx) {
  const EnvDev& E = envs[slot];
  const ContactDev& C = E.contacts[cidx];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= C.n1) return;
  const int p = C.idx1[i];
  const int d = E.deg[p];
  int cnt = 0, jm = -1;
  if (d > 0) {
    const uint32_t sl = pd_slot(E.pslot0[p], d - 1, E.lgG);
    if (!E.state[sl]) {
      const int q = E.col[sl];
      const double4* y = cur_y(E);
      const double4 a = y[p], b = y[q];
      const double dr2 = (a.x - b.x) * (a.x - b.x) + (a.y - b.y) * (a.y - b.y) + (a.z - b.z) * (a.z - b.z);
      if (dr2 < C.equi * C.equi) {
        jm = C.inv1[q];
        cnt = 1;
      }
    }
  }
  if (cnt) C.own[(size_t)i * C.max_neighs] = jm;
  C.own_cnt[i] = cnt;
  if (cnt) atomicAdd(&C.rev_cnt[jm], 1);
}

// ---- device-wide exclusive scan (tiles of 1024 = 256 threads x 4) ----------
constexpr int kTile = 1024;

__device__ __forceinline__ int32_t block_excl_scan(int32_t v, int32_t* sh, int32_t& total) {
  // 256-thread exclusive scan through LDS (wave scans + wave totals)
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
  int32_t x = v;
  for (int o = 1; o < 64; o <<= 1) {
    const int32_t y = __shfl_up(x, o);
    if (lane >= o) x += y;
  }
  if (lane == 63) sh[wv] = x;
  __syncthreads();
  int32_t wofs = 0;
  for (int q = 0; q < wv; ++q) wofs += sh[q];
  total = sh[0] + sh[1] + sh[2] + sh[3];
  __syncthreads();
  return wofs + x - v;
}

__device__ void scan_tiles_body(const int32_t* __restrict__ in, int n, int32_t* __restrict__ tile_sum, int tile) {
  const int base = tile * kTile + threadIdx.x * 4;
  int32_t s = 0;
  for (int k = 0; k < 4; ++k) s += (base + k < n) ? in[base + k] : 0;
  __shared__ int32_t sh[4];
  int32_t tot;
  block_excl_scan(s, sh, tot);
  if (threadIdx.x == 0) tile_sum[tile] = tot;
}

__global__ __launch_bounds__(256) void k_scan_tiles(const int32_t* __restrict__ in, int n,
                                                    int32_t* __restrict__ tile_sum) {
  scan_tiles_body(in, n, tile_sum, blockIdx.x);
}

__device__ void scan_sums_body(int32_t* __restrict__ tile_sum, int ntiles) {
  __shared__ int32_t sh[4];
  int32_t carry = 0;
  for (int b = 0; b < ntiles; b += 256) {
    const int i = b + threadIdx.x;
    const int32_t v = (i < ntiles) ? tile_sum[i] : 0;
    int32_t tot;
    const int32_t ex = block_excl_scan(v, sh, tot);
    if (i < ntiles) tile_sum[i] = carry + ex;
    carry += tot;
  }
}

__global__ __launch_bounds__(256) void k_scan_sums(int32_t* __restrict__ tile_sum, int ntiles) {
  scan_sums_body(tile_sum, ntiles);
}

__device__ void scan_apply_body(const int32_t* __restrict__ in, int32_t* __restrict__ out, int n,
                                const int32_t* __restrict__ tile_off, int tile) {
  const int base = tile * kTile + threadIdx.x * 4;
  int32_t v[4];
  int32_t s = 0;
  for (int k = 0; k < 4; ++k) { v[k] = (base + k < n) ? in[base + k] : 0; s += v[k]; }
  __shared__ int32_t sh[4];
  int32_t tot;
  int32_t run = block_excl_scan(s, sh, tot) + tile_off[tile];
  for (int k = 0; k < 4; ++k) {
    if (base + k < n) out[base + k] = run;
    run += v[k];
  }
}

__global__ __launch_bounds__(256) void k_scan_apply(const int32_t* __restrict__ in, int32_t* __restrict__ out,
                                                    int n, const int32_t* __restrict__ tile_off) {
  scan_apply_body(in, out, n, tile_off, blockIdx.x);
}

// ---- contact updates of many envs at once ----------------------------------
// items[k] = (env slot, contact index); starts[cls * (nitems + 1) + k] = first
// workgroup of item k in the launches of block class cls (no workgroup spans
// two items), starts[cls * (nitems + 1) + nitems] = that class's grid.  Every
// item runs the per-env kernels' bodies, so the lists equal per-env updates.
__device__ __forceinline__ int item_of(const int32_t* __restrict__ st, int nitems, int b) {
  int lo = 0, hi = nitems - 1;  // last k with st[k] <= b
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (st[mid] <= b) lo = mid; else hi = mid - 1;
  }
  return lo;
}

struct BItem {
  const EnvDev* E;
  const ContactDev* C;
  int local;  // workgroup index inside the item
};

__device__ __forceinline__ BItem bitem(const EnvDev* envs, const int2* items, const int32_t* starts, int nitems,
                                       int cls) {
  const int32_t* st = starts + cls * (nitems + 1);
  const int k = item_of(st, nitems, blockIdx.x);
  const int2 it = items[k];
  const EnvDev* E = &envs[it.x];
  return BItem{E, &E->contacts[it.y], (int)blockIdx.x - st[k]};
}

__device__ __forceinline__ int32_t* scan_scratch(const ContactDev& C) {
  return reinterpret_cast<int32_t*>(C.box + 16 + 12 * kRedBlocks);
}

__global__ __launch_bounds__(256) void kb_box1(const EnvDev* envs, const int2* items, const int32_t* starts,
                                               int nitems) {
  const BItem b = bitem(envs, items, starts, nitems, PD_CB_BOX);
  cbox1_body(*b.E, *b.C, b.C->box + 16, b.local, kRedBlocks);
}

__global__ __launch_bounds__(64) void kb_box2(const EnvDev* envs, const int2* items, int nitems) {
  const int2 it = items[blockIdx.x];
  const ContactDev& C = envs[it.x].contacts[it.y];
  cbox2_body(C, C.box + 16, kRedBlocks);
}

// zeroes cell_count, the cell cursor, rev_cursor, rev_cnt and the list lengths
__global__ __launch_bounds__(256) void kb_zero(const EnvDev* envs, const int2* items, const int32_t* starts,
                                               int nitems) {
  const BItem b = bitem(envs, items, starts, nitems, PD_CB_ZERO);
  const ContactDev& C = *b.C;
  int l = b.local * 256 + threadIdx.x;
  if (l < C.ncell_cap) { C.cell_count[l] = 0; return; }
  l -= C.ncell_cap;
  if (l < C.ncell_cap) { C.cell_start[C.ncell_cap + l] = 0; return; }
  l -= C.ncell_cap;
  if (l < C.n2) { C.rev_cursor[l] = 0; return; }
  l -= C.n2;
  if (l < C.n2) { C.rev_cnt[l] = 0; return; }
  l -= C.n2;
  if (l < 2) C.wcount[l] = 0;
}

__global__ __launch_bounds__(256) void kb_bin(const EnvDev* envs, const int2* items, const int32_t* starts,
                                              int nitems) {
  const BItem b = bitem(envs, items, starts, nitems, PD_CB_N2T);
  cbin_body(*b.E, *b.C, b.local * 256 + threadIdx.x);
}

__global__ __launch_bounds__(256) void kb_fill(const EnvDev* envs, const int2* items, const int32_t* starts,
                                               int nitems) {
  const BItem b = bitem(envs, items, starts, nitems, PD_CB_N2T);
  cfill_body(*b.C, b.local * 256 + threadIdx.x);
}

// REV: scan rev_cnt -> rev_off (n2) instead of cell_count -> cell_start (ncell_cap)
template <bool REV>
__global__ __launch_bounds__(256) void kb_scan_tiles(const EnvDev* envs, const int2* items, const int32_t* starts,
                                                     int nitems) {
  const BItem b = bitem(envs, items, starts, nitems, REV ? PD_CB_N2TILE : PD_CB_CTILE);
  const ContactDev& C = *b.C;
  scan_tiles_body(REV ? C.rev_cnt : C.cell_count, REV ? C.n2 : C.ncell_cap, scan_scratch(C), b.local);
}

template <bool REV>
__global__ __launch_bounds__(256) void kb_scan_sums(const EnvDev* envs, const int2* items, int nitems) {
  const int2 it = items[blockIdx.x];
  const ContactDev& C = envs[it.x].contacts[it.y];
  const int n = REV ? C.n2 : C.ncell_cap;
  if (n <= 0) return;
  scan_sums_body(scan_scratch(C), (n + kTile - 1) / kTile);
}

template <bool REV>
__global__ __launch_bounds__(256) void kb_scan_apply(const EnvDev* envs, const int2* items, const int32_t* starts,
                                                     int nitems) {
  const BItem b = bitem(envs, items, starts, nitems, REV ? PD_CB_N2TILE : PD_CB_CTILE);
  const ContactDev& C = *b.C;
  scan_apply_body(REV ? C.rev_cnt : C.cell_count, REV ? C.rev_off : C.cell_start, REV ? C.n2 : C.ncell_cap,
                  scan_scratch(C), b.local);
}

__global__ __launch_bounds__(256) void kb_search_mark(const EnvDev* envs, const int2* items, const int32_t* starts,
                                                      int nitems) {
  const BItem b = bitem(envs, items, starts, nitems, PD_CB_N1T);
  csearch_mark(*b.E, *b.C, b.local * 256 + threadIdx.x);
}

__global__ __launch_bounds__(256) void kb_sort_mark(const EnvDev* envs, const int2* items, const int32_t* starts,
                                                    int nitems) {
  const BItem b = bitem(envs, items, starts, nitems, PD_CB_N2T);
  crev_mark(*b.C, b.local * 256 + threadIdx.x);
}

template <bool SORT>
__global__ __launch_bounds__(256) void kb_walk(const EnvDev* envs, const int2* items, const int32_t* starts,
                                               int nitems) {
  __shared__ int32_t cand[4][kCandCap];
  const int wv = threadIdx.x >> 6;
  const int cls = SORT ? PD_CB_SORT : PD_CB_SRCH;
  const BItem b = bitem(envs, items, starts, nitems, cls);
  const int32_t* st = starts + cls * (nitems + 1);
  const int k = item_of(st, nitems, blockIdx.x);
  const int nb = st[k + 1] - st[k];
  walk_list<SORT>(*b.E, *b.C, b.local * 4 + wv, nb * 4, threadIdx.x & 63, cand[wv]);
}

__global__ __launch_bounds__(256) void kb_rev_fill(const EnvDev* envs, const int2* items, const int32_t* starts,
                                                   int nitems) {
  const BItem b = bitem(envs, items, starts, nitems, PD_CB_N1W);
  crev_fill_body(*b.C, b.local * 4 + (threadIdx.x >> 6), threadIdx.x & 63);
}


inline int nb(int n) { return n > 0 ? (n + 255) / 256 : 1; }

}  // namespace

int pd_scan_scratch_ints(int n) { return (n + kTile - 1) / kTile + 1; }

void pd_scan_exclusive(hipStream_t s, const int32_t* in, int32_t* out, int n, int32_t* scratch) {
  if (n <= 0) return;
  const int nt = (n + kTile - 1) / kTile;
  hipLaunchKernelGGL(k_scan_tiles, dim3(nt), dim3(256), 0, s, in, n, scratch);
  hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(256), 0, s, scratch, nt);
  hipLaunchKernelGGL(k_scan_apply, dim3(nt), dim3(256), 0, s, in, out, n, (const int32_t*)scratch);
}

// Rebuilds the contact lists of contact model `cidx` of env `slot`.  `c` is the
// host copy of the ContactDev (device pointers, sizes).  Scratch inside c.box:
// [16, 16 + 12*kRedBlocks) partial min/max; the scan scratch follows.
void pd_launch_contact_update(hipStream_t s, EnvDev* envs, int slot, const EnvDev& host_env,
                              const ContactDev& c, int cidx) {
  (void)host_env;
  double* part = c.box + 16;
  int32_t* scan_tmp = reinterpret_cast<int32_t*>(c.box + 16 + 12 * kRedBlocks);
  int32_t* cell_cursor = c.cell_start + c.ncell_cap;
  hipLaunchKernelGGL(k_cbox1, dim3(kRedBlocks), dim3(256), 0, s, envs, slot, cidx, part);
  hipLaunchKernelGGL(k_cbox2, dim3(1), dim3(64), 0, s, envs, slot, cidx, part, kRedBlocks);
  hipLaunchKernelGGL(k_izero, dim3(nb(c.ncell_cap)), dim3(256), 0, s, c.cell_count, c.ncell_cap);
  hipLaunchKernelGGL(k_izero, dim3(nb(c.ncell_cap)), dim3(256), 0, s, cell_cursor, c.ncell_cap);
  hipLaunchKernelGGL(k_izero, dim3(nb(c.n2)), dim3(256), 0, s, c.rev_cursor, c.n2);
  hipLaunchKernelGGL(k_izero, dim3(nb(c.n2)), dim3(256), 0, s, c.rev_cnt, c.n2);
  hipLaunchKernelGGL(k_izero, dim3(1), dim3(256), 0, s, c.wcount, 2);
  hipLaunchKernelGGL(k_cbin, dim3(nb(c.n2)), dim3(256), 0, s, envs, slot, cidx);
  pd_scan_exclusive(s, c.cell_count, c.cell_start, c.ncell_cap, scan_tmp);
  hipLaunchKernelGGL(k_cfill, dim3(nb(c.n2)), dim3(256), 0, s, envs, slot, cidx);
  const int wgrid = c.n1 > 0 ? (c.n1 + 3) / 4 : 1;  // one wavefront per block-1 member
  hipLaunchKernelGGL(k_csearch_mark, dim3(nb(c.n1)), dim3(256), 0, s, envs, slot, cidx);
  if (c.n1 > 0) hipLaunchKernelGGL(k_walk<false>, dim3(list_blocks(c.n1, false)), dim3(256), 0, s, envs, slot, cidx);
  pd_scan_exclusive(s, c.rev_cnt, c.rev_off, c.n2, scan_tmp);
  hipLaunchKernelGGL(k_crev_fill, dim3(wgrid), dim3(256), 0, s, envs, slot, cidx);
  hipLaunchKernelGGL(k_crev_mark, dim3(nb(c.n2)), dim3(256), 0, s, envs, slot, cidx);
  if (c.n2 > 0) hipLaunchKernelGGL(k_walk<true>, dim3(list_blocks(c.n2, true)), dim3(256), 0, s, envs, slot, cidx);
}

int pd_contact_batch_blocks(int cls, int n1, int n2, int ncell_cap) {
  auto cdiv = [](int a, int b) { return a > 0 ? (a + b - 1) / b : 0; };
  switch (cls) {
    case PD_CB_BOX: return kRedBlocks;
    case PD_CB_ZERO: return cdiv(2 * ncell_cap + 2 * n2 + 2, 256);
    case PD_CB_N2T: return cdiv(n2, 256);
    case PD_CB_CTILE: return cdiv(ncell_cap, kTile);
    case PD_CB_N2TILE: return cdiv(n2, kTile);
    case PD_CB_N1W: return cdiv(n1, 4);
    case PD_CB_N1T: return cdiv(n1, 256);
    case PD_CB_SRCH: return list_blocks(n1, false);
    default: return list_blocks(n2, true);  // PD_CB_SORT
  }
}

void pd_launch_contact_update_batch(hipStream_t s, const EnvDev* envs, const int2* items, const int32_t* starts,
                                    int nitems, const int* grid) {
  if (nitems <= 0) return;
  auto go = [&](auto kern, int cls) {
    if (grid[cls] > 0) hipLaunchKernelGGL(kern, dim3(grid[cls]), dim3(256), 0, s, envs, items, starts, nitems);
  };
  go(kb_box1, PD_CB_BOX);
  hipLaunchKernelGGL(kb_box2, dim3(nitems), dim3(64), 0, s, envs, items, nitems);
  go(kb_zero, PD_CB_ZERO);
  go(kb_bin, PD_CB_N2T);
  go(kb_scan_tiles<false>, PD_CB_CTILE);
  hipLaunchKernelGGL(kb_scan_sums<false>, dim3(nitems), dim3(256), 0, s, envs, items, nitems);
  go(kb_scan_apply<false>, PD_CB_CTILE);
  go(kb_fill, PD_CB_N2T);
  go(kb_search_mark, PD_CB_N1T);
  go(kb_walk<false>, PD_CB_SRCH);
  go(kb_scan_tiles<true>, PD_CB_N2TILE);
  hipLaunchKernelGGL(kb_scan_sums<true>, dim3(nitems), dim3(256), 0, s, envs, items, nitems);
  go(kb_scan_apply<true>, PD_CB_N2TILE);
  go(kb_rev_fill, PD_CB_N1W);
  go(kb_sort_mark, PD_CB_N2T);
  go(kb_walk<true>, PD_CB_SORT);
}

void pd_launch_self_contact_update(hipStream_t s, EnvDev* envs, int slot, const ContactDev& c, int cidx) {
  int32_t* scan_tmp = reinterpret_cast<int32_t*>(c.box + 16 + 12 * kRedBlocks);
  hipLaunchKernelGGL(k_izero, dim3(nb(c.n2)), dim3(256), 0, s, c.rev_cursor, c.n2);
  hipLaunchKernelGGL(k_izero, dim3(nb(c.n2)), dim3(256), 0, s, c.rev_cnt, c.n2);
  hipLaunchKernelGGL(k_izero, dim3(1), dim3(256), 0, s, c.wcount, 2);
  hipLaunchKernelGGL(k_cself, dim3(nb(c.n1)), dim3(256), 0, s, envs, slot, cidx);
  pd_scan_exclusive(s, c.rev_cnt, c.rev_off, c.n2, scan_tmp);
  const int wgrid = c.n1 > 0 ? (c.n1 + 3) / 4 : 1;
  hipLaunchKernelGGL(k_crev_fill, dim3(wgrid), dim3(256), 0, s, envs, slot, cidx);
  hipLaunchKernelGGL(k_crev_mark, dim3(nb(c.n2)), dim3(256), 0, s, envs, slot, cidx);
  if (c.n2 > 0) hipLaunchKernelGGL(k_walk<true>, dim3(list_blocks(c.n2, true)), dim3(256), 0, s, envs, slot, cidx);
}
