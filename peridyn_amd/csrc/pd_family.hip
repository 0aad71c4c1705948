// Neighbour family on the GPU: cell list -> SELL-64 bond rows.
//
// Replaces get_cells / cal_family! / the GeneralMaterial truncation
// (src/peri.jl:500-576, src/materials/general_material.jl:85-90).  Cells here are
// at least `horizon` wide (w = horizon * (1 + 2^-30)), so every pair that passes
// the reference predicate
//     ((a1-a2)*(a1-a2) + (b1-b2)*(b1-b2)) + (c1-c2)*(c1-c2) < horizon*horizon
// (peri.jl:545, evaluated here unfused in the same order) lies in the 27-cell
// stencil; the resulting neighbour sets equal the reference's (which are the
// brute-force sets of the same predicate).  Rows are sorted ascending; the
// reference's descending ELL layout is produced on export.
#include "pd_device.h"
#include "pd_kernels.h"

namespace {

__device__ __forceinline__ int cell_coord(double v, double lo, double w, int nc) {
  int c = (int)floor((v - lo) / w);
  return c < 0 ? 0 : (c >= nc ? nc - 1 : c);
}

__global__ void k_cell_of(const double4* __restrict__ x, int first, int n, FamilyGrid g,
                          int32_t* __restrict__ cell_of, int32_t* __restrict__ cell_count) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double4 p = x[first + i];
  const int cx = cell_coord(p.x, g.lo[0], g.w, g.nc[0]);
  const int cy = cell_coord(p.y, g.lo[1], g.w, g.nc[1]);
  const int cz = cell_coord(p.z, g.lo[2], g.w, g.nc[2]);
  const int c = cx + g.nc[0] * (cy + g.nc[1] * cz);
  cell_of[i] = c;
  atomicAdd(&cell_count[c], 1);
}

__global__ void k_cell_fill(int n, const int32_t* __restrict__ cell_of, const int32_t* __restrict__ cell_start,
                            int32_t* __restrict__ cursor, int32_t* __restrict__ items, int first) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int c = cell_of[i];
  const int pos = cell_start[c] + atomicAdd(&cursor[c], 1);
  items[pos] = first + i;
}

template <bool FILL>
__global__ void k_pairs(const double4* __restrict__ x, int first, int n, FamilyGrid g, double horizon,
                        const int32_t* __restrict__ cell_of, const int32_t* __restrict__ cell_start,
                        const int32_t* __restrict__ cell_count, const int32_t* __restrict__ items,
                        const uint32_t* __restrict__ pslot0, int lgG, const int32_t* __restrict__ ref,
                        int32_t* __restrict__ deg, int32_t* __restrict__ col, uint8_t* __restrict__ state) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int p = first + i;
  const double4 xp = x[p];
  const double hh = horizon * horizon;
  const int c = cell_of[i];
  const int cx = c % g.nc[0];
  const int cy = (c / g.nc[0]) % g.nc[1];
  const int cz = c / (g.nc[0] * g.nc[1]);
  int k = 0;
  uint32_t base = 0;
  if (FILL) base = pslot0[p];
  for (int dz = -1; dz <= 1; ++dz) {
    const int zz = cz + dz;
    if (zz < 0 || zz >= g.nc[2]) continue;
    for (int dy = -1; dy <= 1; ++dy) {
      const int yy = cy + dy;
      if (yy < 0 || yy >= g.nc[1]) continue;
      for (int dx = -1; dx <= 1; ++dx) {
        const int xx = cx + dx;
        if (xx < 0 || xx >= g.nc[0]) continue;
        const int cc = xx + g.nc[0] * (yy + g.nc[1] * zz);
        const int s = cell_start[cc], e = s + cell_count[cc];
        for (int t = s; t < e; ++t) {
          const int q = items[t];
          if (q == p) continue;
          const double4 xq = x[q];
          const double a = xp.x - xq.x, b = xp.y - xq.y, cz2 = xp.z - xq.z;
          if ((a * a + b * b + cz2 * cz2) < hh) {
            if (FILL) col[pd_slot(base, k, lgG)] = q;
            ++k;
          }
        }
      }
    }
  }
  if (!FILL) {
    deg[p] = k;
    return;
  }
  // insertion sort of the row by reference index, ascending in k (so that a
  // descending walk over k visits the reference family order)
  for (int a = 1; a < k; ++a) {
    const int v = col[pd_slot(base, a, lgG)];
    const int rv = ref[v];
    int b = a - 1;
    while (b >= 0 && ref[col[pd_slot(base, b, lgG)]] > rv) {
      col[pd_slot(base, b + 1, lgG)] = col[pd_slot(base, b, lgG)];
      --b;
    }
    col[pd_slot(base, b + 1, lgG)] = v;
  }
  for (int a = 0; a < k; ++a) state[pd_slot(base, a, lgG)] = 1;
}

// The same pairs with one wavefront per particle (4 per workgroup): lanes 0..26
// read the ranges of the 27 cells around the particle's cell, and the cells'
// items, concatenated in stencil order, are dealt to the lanes 64 at a time
// with k_pairs' predicate.  The degree pass counts the hits; the fill pass
// gathers them into LDS and writes each at its rank by reference index
// (ascending; the indices are distinct), the row k_pairs' insertion sort gives.
// Rows of more than kFamRowCap bonds are left to k_pairs.
constexpr int kFamRowCap = 1024;

template <bool FILL>
__global__ __launch_bounds__(256) void k_pairs_wave(
    const double4* __restrict__ x, int first, int n, FamilyGrid g, double horizon,
    const int32_t* __restrict__ cell_of, const int32_t* __restrict__ cell_start,
    const int32_t* __restrict__ cell_count, const int32_t* __restrict__ items, const uint32_t* __restrict__ pslot0,
    int lgG, const int32_t* __restrict__ ref, int32_t* __restrict__ deg, int32_t* __restrict__ col,
    uint8_t* __restrict__ state) {
  __shared__ int32_t s_q[FILL ? 4 * kFamRowCap : 1];
  __shared__ int32_t s_r[FILL ? 4 * kFamRowCap : 1];
  const int wv = pd_wave_id(), lane = pd_lane();
  const int i = (int)blockIdx.x * 4 + wv;
  if (i >= n) return;
  const int p = first + i;
  const double4 xp = x[p];
  const double hh = horizon * horizon;
  const int c = cell_of[i];
  const int cx = c % g.nc[0];
  const int cy = (c / g.nc[0]) % g.nc[1];
  const int cz = c / (g.nc[0] * g.nc[1]);
  int my_s = 0, my_n = 0;
  if (lane < 27) {
    const int xx = cx + lane % 3 - 1, yy = cy + (lane / 3) % 3 - 1, zz = cz + lane / 9 - 1;
    if (xx >= 0 && xx < g.nc[0] && yy >= 0 && yy < g.nc[1] && zz >= 0 && zz < g.nc[2]) {
      const int cc = xx + g.nc[0] * (yy + g.nc[1] * zz);
      my_s = cell_start[cc];
      my_n = cell_count[cc];
    }
  }
  int incl = my_n;  // inclusive scan of the cell sizes over the lanes
  for (int o = 1; o < 64; o <<= 1) {
    const int v = __shfl_up(incl, o);
    if (lane >= o) incl += v;
  }
  const int total = __shfl(incl, 63);
  const int excl = incl - my_n;
  int32_t* qb = s_q + (FILL ? wv * kFamRowCap : 0);
  int32_t* rb = s_r + (FILL ? wv * kFamRowCap : 0);
  int k = 0;
  for (int t0 = 0; t0 < total; t0 += 64) {
    const int t = t0 + lane;
    // the cell holding concatenated index t: the last of lanes 0..26 whose
    // first index is <= t (empty cells share their successor's first index)
    int cl = 0;
    for (int b = 16; b > 0; b >>= 1) {
      const int cand = cl + b;
      const int ex = __shfl(excl, cand < 27 ? cand : 26);
      if (cand < 27 && ex <= t) cl = cand;
    }
    const int cs = __shfl(my_s, cl), ce = __shfl(excl, cl);
    bool hit = false;
    int q = -1;
    if (t < total) {
      q = items[cs + (t - ce)];
      if (q != p) {
        const double4 xq = x[q];
        const double a = xp.x - xq.x, b = xp.y - xq.y, cz2 = xp.z - xq.z;
        hit = (a * a + b * b + cz2 * cz2) < hh;
      }
    }
    const uint64_t m = __ballot(hit);
    if (FILL && hit) {
      const int pos = k + lanes_below(m);
      if (pos < kFamRowCap) {
        qb[pos] = q;
        rb[pos] = ref[q];
      }
    }
    k += __popcll(m);
  }
  if (!FILL) {
    if (lane == 0) deg[p] = k;
    return;
  }
  if (k > kFamRowCap) return;  // the host sends such blocks to k_pairs
  wave_sync();
  const uint32_t base = pslot0[p];
  for (int a = lane; a < k; a += 64) {
    const int32_t ra = rb[a];
    int r = 0;
    for (int b = 0; b < k; ++b) r += rb[b] < ra ? 1 : 0;
    col[pd_slot(base, r, lgG)] = qb[a];
    state[pd_slot(base, r, lgG)] = 1;
  }
}

// weighted_volume (src/peri.jl:342-354), summed in the family's descending order
__global__ void k_wvol(const double4* __restrict__ x, int first, int n, const uint32_t* __restrict__ pslot0,
                       int lgG, const int32_t* __restrict__ deg, const int32_t* __restrict__ col,
                       double* __restrict__ m) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int p = first + i;
  const double4 xp = x[p];
  const uint32_t base = pslot0[p];
  double acc = 0.0;
  for (int k = deg[p] - 1; k >= 0; --k) {
    const int j = col[pd_slot(base, k, lgG)];
    const double4 xj = x[j];
    const double d1 = xj.x - xp.x, d2 = xj.y - xp.y, d3 = xj.z - xp.z;
    const double w = 1.0 / pd_mag(d1, d2, d3);
    const double r = pd_mag(d1, d2, d3);
    acc += ((w * (r * r)) * 1.0) * xj.w;
  }
  m[p] = acc;
}

// reverse-bond slot for every bond (rows sorted ascending: binary search)
__global__ void k_rev(int first, int n, const uint32_t* __restrict__ pslot0, int lgG,
                      const int32_t* __restrict__ ref, const int32_t* __restrict__ deg,
                      const int32_t* __restrict__ col, uint32_t* __restrict__ rev) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int p = first + i;
  const uint32_t base = pslot0[p];
  const int d = deg[p];
  for (int k = 0; k < d; ++k) {
    const int j = col[pd_slot(base, k, lgG)];
    const uint32_t bj = pslot0[j];
    int lo = 0, hi = deg[j] - 1;
    int found = -1;
    const int rp = ref[p];
    while (lo <= hi) {
      const int mid = (lo + hi) >> 1;
      const int v = ref[col[pd_slot(bj, mid, lgG)]];
      if (v == rp) { found = mid; break; }
      if (v < rp) lo = mid + 1; else hi = mid - 1;
    }
    rev[pd_slot(base, k, lgG)] = (found >= 0) ? pd_slot(bj, found, lgG) : 0xffffffffu;
  }
}

inline int blocks_for(int n) { return (n + 255) / 256; }

}  // namespace

void pd_launch_cell_of(hipStream_t s, const double4* x, int first, int n, FamilyGrid g,
                       int32_t* cell_of, int32_t* cell_count) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_cell_of, dim3(blocks_for(n)), dim3(256), 0, s, x, first, n, g, cell_of, cell_count);
}

void pd_launch_cell_fill(hipStream_t s, int n, const int32_t* cell_of, const int32_t* cell_start,
                         int32_t* cell_cursor, int32_t* cell_items, int first) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_cell_fill, dim3(blocks_for(n)), dim3(256), 0, s, n, cell_of, cell_start, cell_cursor,
                     cell_items, first);
}

void pd_launch_degree(hipStream_t s, const double4* x, int first, int n, FamilyGrid g, double horizon,
                      const int32_t* cell_of, const int32_t* cell_start, const int32_t* cell_count,
                      const int32_t* cell_items, int32_t* deg) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_pairs_wave<false>, dim3((n + 3) / 4), dim3(256), 0, s, x, first, n, g, horizon, cell_of,
                     cell_start, cell_count, cell_items, (const uint32_t*)nullptr, 0, (const int32_t*)nullptr, deg,
                     (int32_t*)nullptr, (uint8_t*)nullptr);
}

void pd_launch_fill(hipStream_t s, const double4* x, int first, int n, FamilyGrid g, double horizon,
                    const int32_t* cell_of, const int32_t* cell_start, const int32_t* cell_count,
                    const int32_t* cell_items, const uint32_t* pslot0, int lgG, const int32_t* ref, int32_t* col,
                    uint8_t* state, int maxdeg) {
  if (n <= 0) return;
  if (maxdeg <= kFamRowCap)
    hipLaunchKernelGGL(k_pairs_wave<true>, dim3((n + 3) / 4), dim3(256), 0, s, x, first, n, g, horizon, cell_of,
                       cell_start, cell_count, cell_items, pslot0, lgG, ref, (int32_t*)nullptr, col, state);
  else
    hipLaunchKernelGGL(k_pairs<true>, dim3(blocks_for(n)), dim3(256), 0, s, x, first, n, g, horizon, cell_of,
                       cell_start, cell_count, cell_items, pslot0, lgG, ref, (int32_t*)nullptr, col, state);
}

void pd_launch_wvol(hipStream_t s, const double4* x, int first, int n, const uint32_t* pslot0, int lgG,
                    const int32_t* deg, const int32_t* col, double* m) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_wvol, dim3(blocks_for(n)), dim3(256), 0, s, x, first, n, pslot0, lgG, deg, col, m);
}

void pd_launch_rev(hipStream_t s, int first, int n, const uint32_t* pslot0, int lgG, const int32_t* ref,
                   const int32_t* deg, const int32_t* col, uint32_t* rev) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_rev, dim3(blocks_for(n)), dim3(256), 0, s, first, n, pslot0, lgG, ref, deg, col, rev);
}
