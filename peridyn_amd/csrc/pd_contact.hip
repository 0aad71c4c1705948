// Contact 1-2 neighbour search on the GPU (update_contact_neighs! RM12,
// src/contacts/contacts.jl:434-476, with collision_box :403-416).
//
// Semantics kept from the reference: the AABB intersection grown by `distance`
// gates both sides with strict inequalities; a pair (i, j) is a contact when
//   (a1-a2)*(a1-a2) + (b1-b2)*(b1-b2) + (c1-c2)*(c1-c2) < distance^2
// (block-1 particle first); each block-1 particle keeps the `max_neighs`
// LARGEST block-2 member indices in descending order (contacts.jl:468-469).
// The brute-force O(n1 n2) scan is replaced by a cell list (cells >= distance
// wide) over the in-box block-2 members.  For the gather-form force kernel the
// lists are also transposed: each block-2 member gets its block-1 partners in
// ascending order (the order in which the reference's loop over i scatters
// into f2[:, j]).
#include "pd_device.h"
#include "pd_kernels.h"

namespace {

constexpr int kRedBlocks = 128;

__device__ __forceinline__ const double4* cur_y(const EnvDev& E) { return E.y[E.cur]; }

// stage 1: per-block partial min/max over both member sets
__global__ __launch_bounds__(256) void k_cbox1(const EnvDev* __restrict__ envs, int slot, int cidx,
                                               double* __restrict__ part) {
  const EnvDev& E = envs[slot];
  const ContactDev& C = E.contacts[cidx];
  const double4* y = cur_y(E);
  double v[12];
  for (int d = 0; d < 3; ++d) {
    v[d] = INFINITY; v[3 + d] = -INFINITY; v[6 + d] = INFINITY; v[9 + d] = -INFINITY;
  }
  const int stride = gridDim.x * blockDim.x;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < C.n1; i += stride) {
    const double4 p = y[C.idx1[i]];
    v[0] = fmin(v[0], p.x); v[1] = fmin(v[1], p.y); v[2] = fmin(v[2], p.z);
    v[3] = fmax(v[3], p.x); v[4] = fmax(v[4], p.y); v[5] = fmax(v[5], p.z);
  }
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < C.n2; i += stride) {
    const double4 p = y[C.idx2[i]];
    v[6] = fmin(v[6], p.x); v[7] = fmin(v[7], p.y); v[8] = fmin(v[8], p.z);
    v[9] = fmax(v[9], p.x); v[10] = fmax(v[10], p.y); v[11] = fmax(v[11], p.z);
  }
  __shared__ double sh[12][4];
  for (int o = 32; o > 0; o >>= 1) {
    for (int d = 0; d < 12; ++d) {
      const double t = __shfl_down(v[d], o);
      v[d] = ((d / 3) % 2 == 0) ? fmin(v[d], t) : fmax(v[d], t);
    }
  }
  const int wv = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0)
    for (int d = 0; d < 12; ++d) sh[d][wv] = v[d];
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int d = 0; d < 12; ++d) {
      double r = sh[d][0];
      for (int q = 1; q < 4; ++q) r = ((d / 3) % 2 == 0) ? fmin(r, sh[d][q]) : fmax(r, sh[d][q]);
      part[blockIdx.x * 12 + d] = r;
    }
  }
}

// stage 2: box, overlap flag and cell grid (<= ncell_cap cells) into C.box
__global__ void k_cbox2(const EnvDev* __restrict__ envs, int slot, int cidx, const double* __restrict__ part,
                        int nparts) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const EnvDev& E = envs[slot];
  const ContactDev& C = E.contacts[cidx];
  double v[12];
  for (int d = 0; d < 12; ++d) v[d] = part[d];
  for (int b = 1; b < nparts; ++b)
    for (int d = 0; d < 12; ++d)
      v[d] = ((d / 3) % 2 == 0) ? fmin(v[d], part[b * 12 + d]) : fmax(v[d], part[b * 12 + d]);
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
  int nc[3];
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
__global__ void k_cbin(const EnvDev* __restrict__ envs, int slot, int cidx) {
  const EnvDev& E = envs[slot];
  const ContactDev& C = E.contacts[cidx];
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
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

// single-workgroup exclusive scan of n ints (chunked per thread)
__global__ __launch_bounds__(1024) void k_scan(const int32_t* __restrict__ in, int32_t* __restrict__ out, int n) {
  __shared__ int32_t sh[1024];
  const int t = threadIdx.x;
  const int per = (n + 1023) / 1024;
  const int b = t * per, e = min(n, b + per);
  int32_t s = 0;
  for (int i = b; i < e; ++i) s += in[i];
  sh[t] = s;
  __syncthreads();
  for (int o = 1; o < 1024; o <<= 1) {
    const int32_t add = (t >= o) ? sh[t - o] : 0;
    __syncthreads();
    sh[t] += add;
    __syncthreads();
  }
  int32_t run = sh[t] - s;
  for (int i = b; i < e; ++i) {
    const int32_t v = in[i];
    out[i] = run;
    run += v;
  }
}

__global__ void k_cfill(const EnvDev* __restrict__ envs, int slot, int cidx, int32_t* __restrict__ cursor) {
  const EnvDev& E = envs[slot];
  const ContactDev& C = E.contacts[cidx];
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= C.n2) return;
  const int c = C.cell_of[j];
  if (c < 0) return;
  C.cell_items[C.cell_start[c] + atomicAdd(&cursor[c], 1)] = j;
}

// search: per block-1 member, top max_neighs largest block-2 member ids
__global__ void k_csearch(const EnvDev* __restrict__ envs, int slot, int cidx) {
  const EnvDev& E = envs[slot];
  const ContactDev& C = E.contacts[cidx];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= C.n1) return;
  const double* box = C.box;
  const double4* y = cur_y(E);
  const double4 p = y[C.idx1[i]];
  int cnt = 0;
  const int M = C.max_neighs;
  int32_t* own = C.own + (size_t)i * M;
  if (box[6] != 0.0 && in_box(p, box)) {
    int cc[3];
    cell_id(p, box, cc);
    const int nc[3] = {(int)box[8], (int)box[9], (int)box[10]};
    const double d2max = C.distance * C.distance;
    for (int dz = -1; dz <= 1; ++dz) {
      const int zz = cc[2] + dz;
      if (zz < 0 || zz >= nc[2]) continue;
      for (int dy = -1; dy <= 1; ++dy) {
        const int yy = cc[1] + dy;
        if (yy < 0 || yy >= nc[1]) continue;
        for (int dx = -1; dx <= 1; ++dx) {
          const int xx = cc[0] + dx;
          if (xx < 0 || xx >= nc[0]) continue;
          const int cell = xx + nc[0] * (yy + nc[1] * zz);
          const int s = C.cell_start[cell], e = s + C.cell_count[cell];
          for (int t = s; t < e; ++t) {
            const int j = C.cell_items[t];
            const double4 q = y[C.idx2[j]];
            const double a = p.x - q.x, b = p.y - q.y, c = p.z - q.z;
            if (a * a + b * b + c * c < d2max) {
              // keep the M largest ids, descending
              if (cnt < M) {
                int u = cnt++;
                while (u > 0 && own[u - 1] < j) { own[u] = own[u - 1]; --u; }
                own[u] = j;
              } else if (j > own[M - 1]) {
                int u = M - 1;
                while (u > 0 && own[u - 1] < j) { own[u] = own[u - 1]; --u; }
                own[u] = j;
              }
            }
          }
        }
      }
    }
  }
  C.own_cnt[i] = cnt;
  for (int k = 0; k < cnt; ++k) atomicAdd(&C.rev_cnt[own[k]], 1);
}

__global__ void k_crev_fill(const EnvDev* __restrict__ envs, int slot, int cidx) {
  const EnvDev& E = envs[slot];
  const ContactDev& C = E.contacts[cidx];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= C.n1) return;
  const int cnt = C.own_cnt[i];
  const int32_t* own = C.own + (size_t)i * C.max_neighs;
  for (int k = 0; k < cnt; ++k) {
    const int j = own[k];
    C.rev[C.rev_off[j] + atomicAdd(&C.rev_cursor[j], 1)] = i;
  }
}

__global__ void k_crev_sort(const EnvDev* __restrict__ envs, int slot, int cidx) {
  const EnvDev& E = envs[slot];
  const ContactDev& C = E.contacts[cidx];
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= C.n2) return;
  int32_t* r = C.rev + C.rev_off[j];
  const int n = C.rev_cnt[j];
  for (int a = 1; a < n; ++a) {
    const int v = r[a];
    int b = a - 1;
    while (b >= 0 && r[b] > v) { r[b + 1] = r[b]; --b; }
    r[b + 1] = v;
  }
}

inline int nb(int n) { return n > 0 ? (n + 255) / 256 : 1; }

}  // namespace

// Rebuilds the contact lists of contact model `cidx` of env `slot`.  `c` is the
// host copy of the ContactDev (device pointers, sizes); `part` scratch of
// kRedBlocks*12 doubles lives in c.box + 16.
void pd_launch_contact_update(hipStream_t s, EnvDev* envs, int slot, const EnvDev& host_env,
                              const ContactDev& c, int cidx) {
  (void)host_env;
  double* part = c.box + 16;
  hipLaunchKernelGGL(k_cbox1, dim3(kRedBlocks), dim3(256), 0, s, envs, slot, cidx, part);
  hipLaunchKernelGGL(k_cbox2, dim3(1), dim3(64), 0, s, envs, slot, cidx, part, kRedBlocks);
  hipLaunchKernelGGL(k_izero, dim3(nb(c.ncell_cap)), dim3(256), 0, s, c.cell_count, c.ncell_cap);
  hipLaunchKernelGGL(k_izero, dim3(nb(c.n2)), dim3(256), 0, s, c.rev_cursor, c.n2);
  hipLaunchKernelGGL(k_izero, dim3(nb(c.n2)), dim3(256), 0, s, c.rev_cnt, c.n2);
  hipLaunchKernelGGL(k_cbin, dim3(nb(c.n2)), dim3(256), 0, s, envs, slot, cidx);
  hipLaunchKernelGGL(k_scan, dim3(1), dim3(1024), 0, s, c.cell_count, c.cell_start, c.ncell_cap);
  // cell cursor: the second half of the cell_start allocation
  int32_t* cell_cursor = c.cell_start + c.ncell_cap;
  hipLaunchKernelGGL(k_izero, dim3(nb(c.ncell_cap)), dim3(256), 0, s, cell_cursor, c.ncell_cap);
  hipLaunchKernelGGL(k_cfill, dim3(nb(c.n2)), dim3(256), 0, s, envs, slot, cidx, cell_cursor);
  hipLaunchKernelGGL(k_csearch, dim3(nb(c.n1)), dim3(256), 0, s, envs, slot, cidx);
  hipLaunchKernelGGL(k_scan, dim3(1), dim3(1024), 0, s, c.rev_cnt, c.rev_off, c.n2);
  hipLaunchKernelGGL(k_crev_fill, dim3(nb(c.n1)), dim3(256), 0, s, envs, slot, cidx);
  hipLaunchKernelGGL(k_crev_sort, dim3(nb(c.n2)), dim3(256), 0, s, envs, slot, cidx);
}
