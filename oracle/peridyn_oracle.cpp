// PeriDyn hot-path ORACLE — test infrastructure only.
//
// A single-threaded C++ restatement of the reference's explicit-dynamics loops,
// used by tests/, bench.py's cpu_baseline leg and __graft_entry__.smoke() as the
// checker.  The cpu_baseline leg times it with the reference's threaded loops
// on OpenMP threads (the *_omp entry points and the `par` loops). The product (peridyn_amd/) never links, loads or calls this file.
//
// Arithmetic rules (the reference is FP64 Julia with no contraction):
//   * compiled with -O2 -ffp-contract=off (no FMA fusion, no fast-math);
//   * every expression keeps the reference's association order, e.g.
//     get_magnitude(a) = sqrt((a1*a1 + a2*a2) + a3*a3)  (src/misc/const_macros.jl:116)
//     and `a * b * c * d` is ((a*b)*c)*d;
//   * loops visit particles and family rows in the reference's order, so the
//     non-atomic scatter of the single-threaded reference is reproduced exactly.
//
// Array layouts are the reference's: positions 3 x n column-major (x[3*i+d]),
// family/intact/edp M' x n column-major (a[k + M'*i]), 1-based neighbour ids with
// 0 padding, type ids global Int64, parameter matrices ntype x ntype column-major.
//
// Pinning (what checks each function; see tests/test_oracle_*.py):
//   orc_cal_family      brute force O(n^2) set equality; lattice shell counts
//   orc_weighted_volume closed form sum|xi| on an interior lattice point; brute force
//   orc_dilatation      homogeneous stretch => theta = 3*eps; rigid motion => 0
//   orc_td_norm2        elastic limit / hand-computed single bond
//   orc_force_density   2-particle closed forms (BB, OSB, EP), momentum balance,
//                       homogeneous-deformation interior zero force, break threshold
//   orc_contact_*       2-particle closed form; brute-force search + truncation
//   orc_julia_pow_int   exact small powers; agreement with correctly rounded x^n
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

inline double mag3(double a1, double a2, double a3) {
  // src/misc/const_macros.jl:116
  return std::sqrt(a1 * a1 + a2 * a2 + a3 * a3);
}

// Julia `fld(x::Float64, y::Float64)` = round((x - mod(x, y)) / y): the floor of
// the exact quotient (base/div.jl, base/float.jl `mod`).
inline double julia_mod(double x, double y) {
  double r = std::fmod(x, y);
  if (r == 0.0) return std::copysign(r, y);
  if ((r > 0) != (y > 0)) return r + y;
  return r;
}
inline double julia_fld(double x, double y) {
  return std::nearbyint((x - julia_mod(x, y)) / y);
}

// cell_number (src/peri.jl:484-486), 1-based in, 1-based out
inline int64_t cell_number(int64_t i, int64_t j, int64_t k, const int64_t* N) {
  return i + (j - 1) * N[0] + (k - 1) * N[0] * N[1];
}

inline double mat_at(const double* M, int64_t ntype, int64_t t1, int64_t t2) {
  // 1-based type-local indices into an ntype x ntype column-major matrix
  return M[(t1 - 1) + ntype * (t2 - 1)];
}

}  // namespace

extern "C" {

// Julia `^(x::Float64, n::Integer)` (base/math.jl pow_body): power by squaring
// carried in double-double, muladd taken as a fused multiply-add (x86-64 hosts
// with FMA). n == 0 -> 1, n == 3 -> x*x*x as Julia special-cases.
double orc_julia_pow_int(double x, int64_t n) {
  if (n == 0) return 1.0;
  double y = 1.0, xnlo = 0.0, ynlo = 0.0;
  if (n == 3) return x * x * x;
  if (n < 0) {
    double rx = 1.0 / x;
    if (n == -2) return rx * rx;
    if (std::isfinite(x)) xnlo = -std::fma(x, rx, -1.0) * rx;
    x = rx;
    n = -n;
  }
  while (n > 1) {
    if (n & 1) {
      double err = std::fma(y, xnlo, x * ynlo);
      double p = x * y;
      double plo = std::fma(x, y, -p);
      y = p;
      ynlo = plo + err;
    }
    double err = x * 2 * xnlo;
    double p = x * x;
    double plo = std::fma(x, x, -p);
    x = p;
    xnlo = plo + err;
    n >>= 1;
  }
  double err = std::fma(y, xnlo, x * ynlo);
  return (std::isfinite(x) && std::isfinite(err)) ? std::fma(x, y, err) : x * y;
}

// ---------------------------------------------------------------------------
// Neighbour family.  get_cells (src/peri.jl:500-520), neigh_cells (:459-471),
// cal_family! (:533-558), cal_family (:571-576), and the GeneralMaterial
// truncation + row reversal (src/materials/general_material.jl:85-90).
//
// fam_out must hold min(max_neigh, n) * n int64.  On success returns M' (the
// number of rows kept, = max degree + 1) and fam_out/intact_out hold the M' x n
// family (descending neighbour ids, zero padded) and its intact mask.
// Returns -1 where the reference raises BoundsError (degree > M while filling,
// or max degree == M at the `end:-1:k` truncation with k == 0).
// ---------------------------------------------------------------------------
int64_t orc_cal_family(int64_t n, const double* x, double horizon, int64_t max_neigh,
                       int64_t* fam_out, uint8_t* intact_out) {
  if (n <= 0 || max_neigh <= 0) return -2;
  const int64_t max_part = 30;
  double mn[3], mx[3];
  for (int d = 0; d < 3; ++d) { mn[d] = x[d]; mx[d] = x[d]; }
  for (int64_t i = 0; i < n; ++i)
    for (int d = 0; d < 3; ++d) {
      mn[d] = std::min(mn[d], x[3 * i + d]);
      mx[d] = std::max(mx[d], x[3 * i + d]);
    }
  int64_t N[3];
  for (int d = 0; d < 3; ++d) {
    double q = julia_fld(mx[d] - mn[d], horizon);
    N[d] = (int64_t)std::max(1.0, std::min((double)max_part, q));
  }
  const int64_t ncell = N[0] * N[1] * N[2];
  std::vector<std::vector<int64_t>> cells(ncell);
  std::vector<std::vector<int64_t>> cell_neighs(ncell);
  for (int64_t k = 1; k <= N[2]; ++k)
    for (int64_t j = 1; j <= N[1]; ++j)
      for (int64_t i = 1; i <= N[0]; ++i) {
        std::vector<int64_t>& a = cell_neighs[cell_number(i, j, k, N) - 1];
        for (int64_t kk = k - 1; kk <= k + 1; ++kk)
          for (int64_t jj = j - 1; jj <= j + 1; ++jj)
            for (int64_t ii = i - 1; ii <= i + 1; ++ii)
              if ((0 < ii * jj * kk) && ii <= N[0] && jj <= N[1] && kk <= N[2])
                a.push_back(cell_number(ii, jj, kk, N));
      }
  for (int64_t p = 0; p < n; ++p) {
    int64_t c[3];
    for (int d = 0; d < 3; ++d) {
      double q = julia_fld(x[3 * p + d] - mn[d], horizon);
      c[d] = (int64_t)std::max(1.0, std::min((double)max_part, q));
    }
    cells[cell_number(c[0], c[1], c[2], N) - 1].push_back(p + 1);
  }
  const int64_t M = std::min(max_neigh, n);
  std::vector<int64_t> family((size_t)M * n, 0);
  const double hh = horizon * horizon;
  for (int64_t ci = 0; ci < ncell; ++ci) {
    for (int64_t ca : cells[ci]) {
      int64_t ind = 1;
      const double a1 = x[3 * (ca - 1)], b1 = x[3 * (ca - 1) + 1], c1 = x[3 * (ca - 1) + 2];
      for (int64_t nc : cell_neighs[ci]) {
        for (int64_t fa : cells[nc - 1]) {
          const double a2 = x[3 * (fa - 1)], b2 = x[3 * (fa - 1) + 1], c2 = x[3 * (fa - 1) + 2];
          if ((ca != fa) && (((a1 - a2) * (a1 - a2) + (b1 - b2) * (b1 - b2) + (c1 - c2) * (c1 - c2)) < hh)) {
            if (ind > M) return -1;  // BoundsError in family[ind, ca]
            family[(size_t)(ca - 1) * M + (ind - 1)] = fa;
            ++ind;
          }
        }
      }
    }
  }
  int64_t maxdeg = 0;
  for (int64_t i = 0; i < n; ++i) {
    int64_t* col = &family[(size_t)i * M];
    std::sort(col, col + M);
    int64_t deg = 0;
    for (int64_t k = 0; k < M; ++k) deg += (col[k] > 0);
    maxdeg = std::max(maxdeg, deg);
  }
  const int64_t kk = M - maxdeg;  // general_material.jl:88
  if (kk < 1) return -1;          // intact[end:-1:0, :] -> BoundsError
  const int64_t Mp = M - kk + 1;
  for (int64_t i = 0; i < n; ++i)
    for (int64_t r = 0; r < Mp; ++r) {
      const int64_t v = family[(size_t)i * M + (M - 1 - r)];
      fam_out[(size_t)i * Mp + r] = v;
      intact_out[(size_t)i * Mp + r] = (uint8_t)(v > 0);
    }
  return Mp;
}

// weighted_volume (src/peri.jl:342-354):
// m_i += influence_function(dr) * |dr|^2 * horizon_correction * vol_j
void orc_weighted_volume(int64_t n, int64_t Mp, const int64_t* fam, const double* x,
                         const double* vol, double* m_out) {
  for (int64_t i = 0; i < n; ++i) {
    double m = 0.0;
    for (int64_t k = 0; k < Mp; ++k) {
      const int64_t j = fam[(size_t)i * Mp + k];
      if ((j > 0) && ((i + 1) != j)) {
        const double d1 = x[3 * (j - 1)] - x[3 * i];
        const double d2 = x[3 * (j - 1) + 1] - x[3 * i + 1];
        const double d3 = x[3 * (j - 1) + 2] - x[3 * i + 2];
        const double w = 1.0 / mag3(d1, d2, d3);
        const double r = mag3(d1, d2, d3);
        m += ((w * (r * r)) * 1.0) * vol[j - 1];
      }
    }
    m_out[i] = m;
  }
}

// The per-particle loops below take `par`: true runs them on OpenMP threads, for
// the CPU baseline's timing only (the reference runs them under Threads.@threads,
// src/peri.jl:158,420).  An iteration writes only particle i's entries, so the
// threaded loops give exactly the single-threaded results.

// dilatation! cpu (src/peri.jl:155-172):
// theta_i += c_i * w * xij * extension * vol_j * horizon_correction, c_i = 3/m_i
static void dilatation_loop(int64_t n, int64_t Mp, const int64_t* fam, const uint8_t* intact,
                            const double* x, const double* y, const double* vol, const double* m,
                            double* theta, bool par) {
  (void)par;
  for (int64_t i = 0; i < n; ++i) theta[i] = 0.0;
#pragma omp parallel for schedule(static) if (par)
  for (int64_t i = 0; i < n; ++i) {
    const double c_i = 3.0 / m[i];
    for (int64_t k = 0; k < Mp; ++k) {
      if (intact[(size_t)i * Mp + k]) {
        const int64_t j = fam[(size_t)i * Mp + k] - 1;
        const double X1 = x[3 * j] - x[3 * i], X2 = x[3 * j + 1] - x[3 * i + 1], X3 = x[3 * j + 2] - x[3 * i + 2];
        const double Y1 = y[3 * j] - y[3 * i], Y2 = y[3 * j + 1] - y[3 * i + 1], Y3 = y[3 * j + 2] - y[3 * i + 2];
        const double xij = mag3(X1, X2, X3);
        const double yij = mag3(Y1, Y2, Y3);
        const double extension = yij - xij;
        const double w = 1.0 / mag3(X1, X2, X3);
        theta[i] += ((((c_i * w) * xij) * extension) * vol[j]) * 1.0;
      }
    }
  }
}

void orc_dilatation(int64_t n, int64_t Mp, const int64_t* fam, const uint8_t* intact,
                    const double* x, const double* y, const double* vol, const double* m,
                    double* theta) {
  dilatation_loop(n, Mp, fam, intact, x, y, vol, m, theta, false);
}

// Material kinds.
enum { ORC_BB = 0, ORC_OSB = 1, ORC_EP = 2 };

struct OrcBlock {
  int64_t n, Mp, ntype, type_start;
  const int64_t* fam;
  uint8_t* intact;
  const double* x;     // 3 x n reference positions
  const double* vol;   // n
  const int64_t* type; // n global type ids
  const double* s0;    // ntype x ntype critical stretch
  const double* density;  // ntype
  const double* c;     // BB bond stiffness matrix
  const double* K;     // OSB/EP bulk modulus matrix
  const double* G;     // OSB/EP shear modulus matrix
  const double* psi;   // EP yield value matrix
  const double* m;     // weighted volume
  double* theta;       // n
  double* td_norm2;    // n
  double* edp;         // Mp x n
};

// EP td_norm2 pass: out_of_loop! (src/materials/elasto_plastic.jl:168-172) ->
// loop_over_neighs! (src/peri.jl:411-445) with td_norm2_ij! (elasto_plastic.jl:282-295).
// Breaks bonds with s >= s0 (peri.jl:439).
static void td_norm2_loop(const OrcBlock* b, const double* y, bool par) {
  (void)par;
  const int64_t n = b->n, Mp = b->Mp, nt = b->ntype;
  for (int64_t i = 0; i < n; ++i) b->td_norm2[i] = 0.0;
#pragma omp parallel for schedule(static) if (par)
  for (int64_t i = 0; i < n; ++i) {
    for (int64_t k = 0; k < Mp; ++k) {
      const size_t s_ik = (size_t)i * Mp + k;
      if (!b->intact[s_ik]) continue;
      const int64_t j = b->fam[s_ik] - 1;
      const double* x = b->x;
      const double X1 = x[3 * j] - x[3 * i], X2 = x[3 * j + 1] - x[3 * i + 1], X3 = x[3 * j + 2] - x[3 * i + 2];
      const double Y1 = y[3 * j] - y[3 * i], Y2 = y[3 * j + 1] - y[3 * i + 1], Y3 = y[3 * j + 2] - y[3 * i + 2];
      const double yij = mag3(Y1, Y2, Y3);
      const double xij = mag3(X1, X2, X3);
      const double extension = yij - xij;
      const double s = extension / xij;
      const int64_t t1 = b->type[i] - (b->type_start - 1);
      const int64_t t2 = b->type[j] - (b->type_start - 1);
      if (s < mat_at(b->s0, nt, t1, t2)) {
        const double wij = 1.0 / mag3(X1, X2, X3);
        const double wji = 1.0 / mag3(-X1, -X2, -X3);
        const double Gij = mat_at(b->G, nt, t1, t2);
        const double edpij = b->edp[s_ik];
        const double a = wij / b->m[i], bb = wji / b->m[j];
        const double w_m = a + bb;
        const double tw_m = b->theta[i] * a + b->theta[j] * bb;
        const double theta_ij = tw_m / w_m;
        const double ed = extension - (theta_ij * xij) / 3.0;
        const double td = ((15.0 * Gij) * (ed - edpij)) * w_m;
        b->td_norm2[i] += (td * td) * b->vol[j];
      } else {
        b->intact[s_ik] = 0;
      }
    }
  }
}

void orc_td_norm2(const OrcBlock* b, const double* y) { td_norm2_loop(b, y, false); }

// force_density_t_ij plug-ins.
// BB  src/materials/bond_based.jl:225-232 (items[1] = c[1,1], bond_force = s*c :195-197)
// OSB src/materials/ordinary_state_based.jl:176-187
// EP  src/materials/elasto_plastic.jl:228-255 (mutates edp[k,i])
static inline double orc_t_ij(int kind, const OrcBlock* b, int64_t i, int64_t j, size_t s_ik,
                              int64_t t1, int64_t t2, double xij, double yij, double extension,
                              double s, double wij, double wji) {
  (void)yij;
  const int64_t nt = b->ntype;
  if (kind == ORC_BB) {
    const double bond_stiffness = b->c[0];
    return s * bond_stiffness;
  }
  if (kind == ORC_OSB) {
    const double Kij = mat_at(b->K, nt, t1, t2), Gij = mat_at(b->G, nt, t1, t2);
    const double a = wij / b->m[i], bb = wji / b->m[j];
    return (3.0 * Kij - 5.0 * Gij) * (b->theta[i] * xij * a + b->theta[j] * xij * bb) +
           15.0 * Gij * extension * (a + bb);
  }
  // EP
  const double Kij = mat_at(b->K, nt, t1, t2), Gij = mat_at(b->G, nt, t1, t2);
  const double psiij = mat_at(b->psi, nt, t1, t2);
  const double C = std::sqrt((2.0 * psiij) / b->td_norm2[i]);
  const double a = wij / b->m[i], bb = wji / b->m[j];
  const double tw_m = b->theta[i] * a + b->theta[j] * bb;
  const double theta_ij = tw_m / (a + bb);
  const double ed = extension - (theta_ij * xij) / 3.0;
  const double ede = ed - b->edp[s_ik];
  const double td_trial = ((15.0 * Gij) * (a + bb)) * ede;
  double td;
  if (C > 1.0) {
    td = td_trial;
  } else {
    td = C * td_trial;
    b->edp[s_ik] = b->edp[s_ik] + ede * (1.0 - C);
  }
  const double tn = ((3.0 * Kij) * xij) * tw_m;
  return tn + td;
}

// out_of_loop! + force_density_T! cpu (src/materials/general_material.jl:296-369).
// f and y point at the block's first particle inside the global 3 x N arrays
// (the `limits` views).  The OSB/EP pre-passes run here as in the reference.
void orc_force_density(int kind, const OrcBlock* b, const double* y, double* f) {
  const int64_t n = b->n, Mp = b->Mp, nt = b->ntype;
  if (kind == ORC_OSB || kind == ORC_EP)
    orc_dilatation(n, Mp, b->fam, b->intact, b->x, y, b->vol, b->m, b->theta);
  if (kind == ORC_EP) orc_td_norm2(b, y);
  const double* x = b->x;
  for (int64_t i = 0; i < n; ++i) {
    const int64_t t1 = b->type[i] - (b->type_start - 1);
    const double den_i = b->density[t1 - 1];
    for (int64_t k = 0; k < Mp; ++k) {
      const size_t s_ik = (size_t)i * Mp + k;
      if (!b->intact[s_ik]) continue;
      const int64_t j = b->fam[s_ik] - 1;
      const int64_t t2 = b->type[j] - (b->type_start - 1);
      const double den_j = b->density[t2 - 1];
      const double X1 = x[3 * j] - x[3 * i], X2 = x[3 * j + 1] - x[3 * i + 1], X3 = x[3 * j + 2] - x[3 * i + 2];
      const double Y1 = y[3 * j] - y[3 * i], Y2 = y[3 * j + 1] - y[3 * i + 1], Y3 = y[3 * j + 2] - y[3 * i + 2];
      const double yij = mag3(Y1, Y2, Y3);
      const double xij = mag3(X1, X2, X3);
      const double extension = yij - xij;
      const double s = extension / xij;
      if (s < mat_at(b->s0, nt, t1, t2)) {
        const double wij = 1.0 / mag3(X1, X2, X3);
        const double wji = 1.0 / mag3(-X1, -X2, -X3);
        const double tij = orc_t_ij(kind, b, i, j, s_ik, t1, t2, xij, yij, extension, s, wij, wji);
        const double xx = 0.5 * tij / yij;
        f[3 * i + 0] += xx * Y1 * b->vol[j] / den_i;
        f[3 * i + 1] += xx * Y2 * b->vol[j] / den_i;
        f[3 * i + 2] += xx * Y3 * b->vol[j] / den_i;
        f[3 * j + 0] -= xx * Y1 * b->vol[i] / den_j;
        f[3 * j + 1] -= xx * Y2 * b->vol[i] / den_j;
        f[3 * j + 2] -= xx * Y3 * b->vol[i] / den_j;
      } else {
        b->intact[s_ik] = 0;
      }
    }
  }
}

// force_density for the CPU baseline's timing (never a parity source): the
// loops of orc_force_density with OpenMP over particles.  The pre-passes equal
// the single-threaded ones exactly; the force loop's scatter to j goes to one
// array per thread, summed afterwards (the race-free form of the reference's
// threaded scatter, general_material.jl:330), so f differs from the
// single-threaded sums by rounding.  Returns the threads used.
int orc_force_density_omp(int kind, const OrcBlock* b, const double* y, double* f) {
  const int64_t n = b->n, Mp = b->Mp, nt = b->ntype;
  if (kind == ORC_OSB || kind == ORC_EP)
    dilatation_loop(n, Mp, b->fam, b->intact, b->x, y, b->vol, b->m, b->theta, true);
  if (kind == ORC_EP) td_norm2_loop(b, y, true);
  const double* x = b->x;
  int nth = 1;
#ifdef _OPENMP
  nth = omp_get_max_threads();
#endif
  static std::vector<double> fl;  // per-thread force arrays, kept between calls (one caller at a time)
  const size_t stride = (size_t)3 * n;
  if (fl.size() < stride * nth) fl.resize(stride * nth);
  int used = 1;  // the team actually started; only its arrays are zeroed and summed
#pragma omp parallel num_threads(nth)
  {
    int tid = 0;
#ifdef _OPENMP
    tid = omp_get_thread_num();
#pragma omp single
    used = omp_get_num_threads();
#endif
    double* ft = fl.data() + stride * tid;
    std::fill(ft, ft + stride, 0.0);
#pragma omp for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
      const int64_t t1 = b->type[i] - (b->type_start - 1);
      const double den_i = b->density[t1 - 1];
      for (int64_t k = 0; k < Mp; ++k) {
        const size_t s_ik = (size_t)i * Mp + k;
        if (!b->intact[s_ik]) continue;
        const int64_t j = b->fam[s_ik] - 1;
        const int64_t t2 = b->type[j] - (b->type_start - 1);
        const double den_j = b->density[t2 - 1];
        const double X1 = x[3 * j] - x[3 * i], X2 = x[3 * j + 1] - x[3 * i + 1], X3 = x[3 * j + 2] - x[3 * i + 2];
        const double Y1 = y[3 * j] - y[3 * i], Y2 = y[3 * j + 1] - y[3 * i + 1], Y3 = y[3 * j + 2] - y[3 * i + 2];
        const double yij = mag3(Y1, Y2, Y3);
        const double xij = mag3(X1, X2, X3);
        const double extension = yij - xij;
        const double s = extension / xij;
        if (s < mat_at(b->s0, nt, t1, t2)) {
          const double wij = 1.0 / mag3(X1, X2, X3);
          const double wji = 1.0 / mag3(-X1, -X2, -X3);
          const double tij = orc_t_ij(kind, b, i, j, s_ik, t1, t2, xij, yij, extension, s, wij, wji);
          const double xx = 0.5 * tij / yij;
          ft[3 * i + 0] += xx * Y1 * b->vol[j] / den_i;
          ft[3 * i + 1] += xx * Y2 * b->vol[j] / den_i;
          ft[3 * i + 2] += xx * Y3 * b->vol[j] / den_i;
          ft[3 * j + 0] -= xx * Y1 * b->vol[i] / den_j;
          ft[3 * j + 1] -= xx * Y2 * b->vol[i] / den_j;
          ft[3 * j + 2] -= xx * Y3 * b->vol[i] / den_j;
        } else {
          b->intact[s_ik] = 0;
        }
      }
    }
  }
#pragma omp parallel for schedule(static) num_threads(nth)
  for (int64_t q = 0; q < (int64_t)stride; ++q) {
    double acc = 0.0;
    for (int t = 0; t < used; ++t) acc += fl[stride * t + q];
    f[q] += acc;
  }
  return used;
}

// ---------------------------------------------------------------------------
// Contact 1-2.  collision_box (src/contacts/contacts.jl:403-416) and
// update_contact_neighs! RM12 (:434-476).  y1 (3 x n1) and y2 (3 x n2) are the
// type-masked position views.  neighs_out (max_neighs x n1) receives, per
// block-1 particle, the `max_neighs` largest block-2 indices (1-based) within
// `distance`, descending, zero padded.
// ---------------------------------------------------------------------------
static void contact_neighs_loop(int64_t n1, const double* y1, int64_t n2, const double* y2,
                                double distance, int64_t max_neighs, int64_t* neighs_out, bool par) {
  (void)par;
  std::memset(neighs_out, 0, sizeof(int64_t) * (size_t)max_neighs * n1);
  if (n1 <= 0 || n2 <= 0) return;
  double mn1[3], mx1[3], mn2[3], mx2[3], bmin[3], bmax[3];
  for (int d = 0; d < 3; ++d) {
    mn1[d] = mx1[d] = y1[d];
    mn2[d] = mx2[d] = y2[d];
  }
  for (int64_t i = 0; i < n1; ++i)
    for (int d = 0; d < 3; ++d) { mn1[d] = std::min(mn1[d], y1[3 * i + d]); mx1[d] = std::max(mx1[d], y1[3 * i + d]); }
  for (int64_t i = 0; i < n2; ++i)
    for (int d = 0; d < 3; ++d) { mn2[d] = std::min(mn2[d], y2[3 * i + d]); mx2[d] = std::max(mx2[d], y2[3 * i + d]); }
  double boxvol = 1.0;
  for (int d = 0; d < 3; ++d) {
    bmin[d] = std::max(mn1[d] - distance, mn2[d] - distance);
    bmax[d] = std::min(mx1[d] + distance, mx2[d] + distance);
  }
  // prod(box_max - box_min) in Julia's left fold
  boxvol = ((bmax[0] - bmin[0]) * (bmax[1] - bmin[1])) * (bmax[2] - bmin[2]);
  if (boxvol < 0.0) return;
  std::vector<int64_t> in1, in2;
  for (int64_t i = 0; i < n1; ++i) {
    bool ok = true;
    for (int d = 0; d < 3; ++d) ok = ok && (y1[3 * i + d] > bmin[d]) && (y1[3 * i + d] < bmax[d]);
    if (ok) in1.push_back(i);
  }
  for (int64_t i = 0; i < n2; ++i) {
    bool ok = true;
    for (int d = 0; d < 3; ++d) ok = ok && (y2[3 * i + d] > bmin[d]) && (y2[3 * i + d] < bmax[d]);
    if (ok) in2.push_back(i);
  }
  const double dist2 = distance * distance;
  const int64_t nin1 = (int64_t)in1.size();
#pragma omp parallel if (par)
  {
  std::vector<int64_t> cand;
#pragma omp for schedule(dynamic, 64)
  for (int64_t q1 = 0; q1 < nin1; ++q1) {
    const int64_t i = in1[q1];
    cand.clear();
    const double a1 = y1[3 * i], b1 = y1[3 * i + 1], c1 = y1[3 * i + 2];
    for (int64_t j : in2) {
      const double a2 = y2[3 * j], b2 = y2[3 * j + 1], c2 = y2[3 * j + 2];
      if ((a1 - a2) * (a1 - a2) + (b1 - b2) * (b1 - b2) + (c1 - c2) * (c1 - c2) < dist2)
        cand.push_back(j + 1);
    }
    // sort ascending, then keep family[end:-1:end+1-max_neighs]
    std::sort(cand.begin(), cand.end());
    int64_t r = 0;
    for (int64_t q = (int64_t)cand.size() - 1; q >= 0 && r < max_neighs; --q, ++r)
      neighs_out[(size_t)i * max_neighs + r] = cand[q];
  }
  }
}

void orc_contact_neighs(int64_t n1, const double* y1, int64_t n2, const double* y2,
                        double distance, int64_t max_neighs, int64_t* neighs_out) {
  contact_neighs_loop(n1, y1, n2, y2, distance, max_neighs, neighs_out, false);
}

// The same with the block-1 loop on OpenMP threads (contacts.jl:459 threads it);
// rows are independent, so the lists equal the single-threaded ones.
void orc_contact_neighs_omp(int64_t n1, const double* y1, int64_t n2, const double* y2,
                            double distance, int64_t max_neighs, int64_t* neighs_out) {
  contact_neighs_loop(n1, y1, n2, y2, distance, max_neighs, neighs_out, true);
}

// strain^expo of the NonLinearSpring force (nonlinearspring.jl:140, expo a
// Float64): Julia's integer power for integer-valued exponents (see
// orc_julia_pow_int), else the C library's pow, which like Julia's general
// Float64 power is within one ulp of the exact value.
double orc_contact_power(double strain, double expo, int expo_is_int) {
  return expo_is_int ? orc_julia_pow_int(strain, (int64_t)expo) : std::pow(strain, expo);
}

// short_range_contact! RM12 cpu (src/contacts/contacts.jl:210-247) with the
// NonLinearSpring force function (src/contacts/nonlinearspring.jl:124-145).
// f1/f2 are the masked force views (3 x n1, 3 x n2) updated in place.
// expo_is_int selects Julia's integer power (pow_body) over a float power.
void orc_contact_force(int64_t n1, const double* y1, const double* vol1, const double* den1, double* f1,
                       int64_t n2, const double* y2, const double* vol2, const double* den2, double* f2,
                       int64_t M, const int64_t* neighs, double equi_size, double K,
                       double expo, int expo_is_int) {
  (void)n2;
  for (int64_t i = 0; i < n1; ++i) {
    const double _den1_i = 1.0 / den1[i];
    for (int64_t k = 0; k < M; ++k) {
      const int64_t j1 = neighs[(size_t)i * M + k];
      if (j1 > 0) {
        const int64_t j = j1 - 1;
        const double _den2_j = 1.0 / den2[j];
        const double d1 = y1[3 * i] - y2[3 * j], d2 = y1[3 * i + 1] - y2[3 * j + 1], d3 = y1[3 * i + 2] - y2[3 * j + 2];
        const double mag_dr = mag3(d1, d2, d3);
        const double del_x = equi_size - mag_dr;
        const double strain = del_x / equi_size;
        double F1 = 0.0, F2 = 0.0, F3 = 0.0;
        if (!(del_x < 0.0)) {
          const double p = orc_contact_power(strain, expo, expo_is_int);
          const double s = (K * p) / (mag_dr + 1e-10);
          F1 = s * d1; F2 = s * d2; F3 = s * d3;
        }
        const double a = vol2[j] * _den1_i;
        f1[3 * i + 0] += F1 * a;
        f1[3 * i + 1] += F2 * a;
        f1[3 * i + 2] += F3 * a;
        const double b = vol1[i] * _den2_j;
        f2[3 * j + 0] -= F1 * b;
        f2[3 * j + 1] -= F2 * b;
        f2[3 * j + 2] -= F3 * b;
      }
      if (j1 == 0) break;
    }
  }
}

}  // extern "C"
