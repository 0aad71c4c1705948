/* peridyn.h — C ABI of libperidyn_hip.so, the MI355X (gfx950) implementation of
 * PeriDyn's explicit-dynamics hot path.
 *
 * The reference is Julia; its GPU seam is multiple dispatch on a device tag
 * (DEVICE[], src/misc/const_macros.jl:10) reaching per-device methods such as
 * force_density_T!(f, y, limits, mat, Val{device}) (src/materials/general_material.jl:274-276).
 * A Julia shim (julia/PeriDynHIP.jl) binds these entry points with ccall.  Every
 * signature uses plain pointers and sizes; arrays use the reference's layouts:
 *   positions / velocities / forces  3 x n column-major Float64 (x[3*i + d])
 *   family / intact                  M' x n column-major, 1-based ids, 0 padded
 *   parameter matrices               ntype x ntype column-major Float64
 *
 * Ownership: host buffers are copied; nothing is retained after return.  Device
 * state belongs to the context and is freed by pd_ctx_destroy / pd_env_destroy.
 * Errors: every call returns 0 or a negative PD_E_* code; pd_last_error() gives
 * the message.  A context is not re-entrant: use it from one host thread.
 */
#ifndef PERIDYN_H
#define PERIDYN_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  PD_OK = 0,
  PD_E_HIP = -1,       /* HIP runtime error (message has the HIP error string) */
  PD_E_ARG = -2,       /* invalid argument */
  PD_E_OVERFLOW = -3,  /* family degree >= max_neigh: the reference raises BoundsError */
  PD_E_STATE = -4,     /* call not valid in the current state (e.g. env not committed) */
  PD_E_NOMEM = -5,     /* device memory exhausted */
  PD_E_NONFINITE = -6  /* NaN/Inf found in state (pd_env_check_finite) */
};

/* material kinds: BondBasedMaterial, OrdinaryStateBasedMaterial, ElastoPlasticSolidMaterial */
enum { PD_BB = 0, PD_OSB = 1, PD_EP = 2 };

/* arithmetic forms of the force passes */
enum {
  PD_PREC_EXACT = 0, /* FP64 in the single-threaded reference's summation order: bit-identical */
  PD_PREC_FP64 = 1,  /* FP64, one pass per bond in neighbour order (plain FP64 form) */
  PD_PREC_FP32C = 2  /* FP32 displacement-difference stretch + Kahan-compensated FP32 sums;
                        FP64 Verlet update and weighted volume */
};

enum { PD_BC_FIX = 0, PD_BC_TOFRO = 1 };

/* One material block = GeneralMaterial + *Specific after init()
 * (general_material.jl:43-56,81-98; bond_based.jl:21-27,118-122;
 *  ordinary_state_based.jl:18-24; elasto_plastic.jl:29-40,59-71). */
typedef struct {
  int32_t kind;                 /* PD_BB / PD_OSB / PD_EP */
  int32_t ntype;                /* number of types in the block (type range length) */
  int64_t n;                    /* particles */
  int64_t type_start;           /* mat.type.start */
  const double* x;              /* 3 x n reference positions */
  const double* y;              /* 3 x n initial deformed positions */
  const double* v;              /* 3 x n initial velocities */
  const double* volume;         /* n */
  const int64_t* type;          /* n global type ids */
  double horizon;
  double particle_size;         /* volume[1]^(1/3) as computed by the caller */
  int64_t max_neigh;            /* GeneralMaterial max_neigh */
  const double* density;        /* ntype */
  const double* critical_stretch; /* ntype x ntype (symmetric) */
  const double* bond_stiffness; /* PD_BB: 18K/pi/delta^4 matrix (only [1,1] is used, bond_based.jl:229) */
  const double* bulk_modulus;   /* PD_OSB / PD_EP */
  const double* shear_modulus;  /* PD_OSB / PD_EP */
  const double* psi;            /* PD_EP: 75/(8pi) sigma_e^2 / delta^5 */
  int32_t deformed;             /* GeneralMaterial.deformed (y0 != x) */
  int32_t reserved;
} pd_block_desc;

/* NonLinearSpringContactModel12 / LinearSpringContactModel (exponent 1) /
 * ShortRangeContactModel (contacts.jl:76-91, nonlinearspring.jl:63-76,124-145);
 * ntypes2 == 0 describes NonLinearSpringContactModel11, the self contact of one
 * block (contacts.jl:105-119, nonlinearspring.jl:78-122): types1 = mat.type,
 * equi_dist = particle_size * distanceD, types2 unused. */
typedef struct {
  const int64_t* types1; int64_t ntypes1; /* pair[1] type ids (self contact: the block's types) */
  const int64_t* types2; int64_t ntypes2; /* pair[2] type ids; 0 for a self contact */
  double exponent;
  int32_t exponent_is_int;      /* nonzero: integer-valued exponent, evaluated by Julia's compensated
                                   power by squaring (base/math.jl pow_body), as Julia 1.10's
                                   Float64^Float64 does for integer-valued y; 0: general pow */
  int32_t reserved;
  double stiffness;
  double equi_dist;             /* (ps1 + ps2)/2 * distanceD */
  double distance;              /* distanceX * equi_dist */
  int64_t max_neighs;           /* min(max_neighs, n2) */
} pd_contact_desc;

/* FixBC / ToFroBC / MoveBC (= ToFroBC with freq = Inf) (FixBC.jl:46-52,
 * ToFroBC.jl:51-62,79-81,107-111).  options:
 *   PD_BC_ONLYATSTART  the constructors' onlyatstart keyword: minimize!
 *       (minimize.jl:138-143) leaves the BC's particles free after applying it
 *       once; run! applies the BC every step regardless, as the reference does;
 *   PD_BC_DIMS(d)      the dims keyword as a 3-bit mask (bit k = dimension k+1;
 *       0 = all three).  The reference assigns its 3 x n update into
 *       y[dims, bool] (boundary_conditions.jl:108-110,123-125), which raises
 *       DimensionMismatch unless every dimension is selected, so a subset is
 *       refused here with PD_E_ARG. */
enum { PD_BC_ONLYATSTART = 1 };
#define PD_BC_DIMS(d) (((int32_t)(d) & 7) << 8)
typedef struct {
  int32_t kind;                 /* PD_BC_FIX / PD_BC_TOFRO */
  int32_t options;              /* PD_BC_ONLYATSTART | PD_BC_DIMS(mask); 0: defaults */
  const uint8_t* mask;          /* N (env particles) */
  double rate[3];               /* ToFroBC rate (a velocity) */
  double freq;                  /* ToFroBC freq (integer-valued or +Inf) */
  int64_t applyafter;
} pd_bc_desc;

/* ---- context --------------------------------------------------------- */
int pd_ctx_create(int device, void** ctx);
int pd_ctx_destroy(void* ctx);
const char* pd_last_error(void* ctx);
int pd_sync(void* ctx);
/* free / total device memory of the context's GPU (hipMemGetInfo; free
 * includes the memory the context keeps from destroyed envs for its next
 * ones, at most 8 GiB); no reference counterpart: the serving loop admits
 * simulations by footprint */
int pd_ctx_mem_info(void* ctx, int64_t* free_bytes, int64_t* total_bytes);

/* ---- environment (Env, src/environment.jl:166-235) ------------------- */
int pd_env_create(void* ctx, double dt, int32_t precision, int64_t* env);
int pd_env_add_block(void* ctx, int64_t env, const pd_block_desc* b, int64_t* block);
int pd_env_add_contact(void* ctx, int64_t env, const pd_contact_desc* c);
int pd_env_add_bc(void* ctx, int64_t env, const pd_bc_desc* bc);
/* Bond-loop schedule of the env: lanes per particle (power of two 1..64; 1 =
 * lane per particle, 64 = wavefront per particle; 0 = choose at commit: 4 for
 * envs that run tiled (pd_env_layout), else from the particle count).  Set
 * before commit; the reference-order form uses 1.  No
 * reference counterpart: it selects among layouts of the same computation
 * (peridyn_amd/autotune.py picks it by timing). */
int pd_env_set_schedule(void* ctx, int64_t env, int32_t lanes_per_particle);
int pd_env_schedule(void* ctx, int64_t env, int32_t* lanes_per_particle);
/* Bond layout of a committed env (no reference counterpart): lanes per
 * particle, whether the FP64 passes run tiled (halos staged in LDS; the
 * default for FP64 envs of single-type blocks with integer contact exponents
 * and 2..8 lanes per particle whose halos fit; PERIDYN_TILES=0 at commit turns
 * it off): *tiled = 0 untiled, 1 tiled, 2 tiled with compact halo records
 * (every reference position and volume is exactly an FP32 value, so they are
 * staged as FP32 and read back unchanged) -- and the largest tile halo.  Any
 * pointer may be NULL. */
int pd_env_layout(void* ctx, int64_t env, int32_t* lanes_per_particle, int32_t* tiled, int32_t* max_halo);
/* Uploads the env, builds every block's family on the GPU (cell list -> CSR),
 * the weighted volume m and the per-particle BC map.  PD_E_OVERFLOW where
 * cal_family/GeneralMaterial would raise BoundsError. */
int pd_env_commit(void* ctx, int64_t env);
int pd_env_destroy(void* ctx, int64_t env);
int pd_env_info(void* ctx, int64_t env, int64_t* nparticles, int64_t* nbonds, int64_t* device_bytes);

/* family in the reference ELL layout (general_material.jl:88-90): rows = max degree + 1 */
int pd_block_family_rows(void* ctx, int64_t env, int64_t block, int64_t* rows);
int pd_block_family_export(void* ctx, int64_t env, int64_t block, int64_t* family, uint8_t* intact, int64_t rows);
int pd_block_weighted_volume(void* ctx, int64_t env, int64_t block, double* m);
int pd_block_theta(void* ctx, int64_t env, int64_t block, double* theta);
int pd_block_edp(void* ctx, int64_t env, int64_t block, double* edp, int64_t rows);

/* ---- stepping ----------------------------------------------------------- */
/* run! physics cadence (solvers.jl:97-187) for the listed envs, stepped
 * together: update_neighs! (as every run! call does); then nsteps
 * velocity_verlet_step! (velocity_verlet.jl:57-89) with contact lists rebuilt
 * after every step i with i % nuf == 0, i counted from 1 in each call (run!'s
 * default start_at = 0).  Returns when the steps have completed. */
int pd_run(void* ctx, const int64_t* envs, int64_t nenv, int64_t nsteps, int64_t nuf);
/* pd_run in two halves, for a serving loop that admits new simulations while a
 * chunk of steps runs.  No counterpart in the reference (its run! returns when
 * done); the stepping is pd_run's, bit for bit.  pd_run_begin enqueues the
 * nsteps >= 1 steps of the listed envs and returns; pd_run_query sets *done to
 * 1 once they have completed (or when nothing is in flight), else 0;
 * pd_run_end waits for them and completes the call.  Between begin and end
 * the only calls accepted on the context are pd_env_create, pd_env_add_block
 * / _contact / _bc, pd_env_set_schedule, pd_env_schedule, pd_env_layout,
 * pd_env_commit, pd_env_info, pd_env_time_step (the value before the run),
 * pd_env_state_size, pd_ctx_mem_info and pd_env_destroy of an env that the
 * run does not step; their device work goes to a second stream of the context
 * and overlaps the run.  Every other call returns PD_E_STATE. */
int pd_run_begin(void* ctx, const int64_t* envs, int64_t nenv, int64_t nsteps, int64_t nuf);
int pd_run_query(void* ctx, int32_t* done);
int pd_run_end(void* ctx);
/* pd_run plus the printdata columns (environment.jl:308-320) that run! logs
 * after step 1 and after every `every`-th step (solvers.jl:158-167, every =
 * average_prop_freq): per logged step and env 8 doubles, time (time_step * dt),
 * px, py, pz (sum of p = mass .* v as update_misc! sets it inside update_acc!,
 * solvers.jl:289-291, i.e. from the step's half-step velocities), Fx, Fy, Fz
 * (sum of f) and the mean damage (cal_damage, peri.jl:587-594).  log holds
 * pd_log_rows(nsteps, every) x nenv x 8 doubles, rows in step order. */
int pd_run_logged(void* ctx, const int64_t* envs, int64_t nenv, int64_t nsteps, int64_t nuf, int64_t every,
                  double* log);
int64_t pd_log_rows(int64_t nsteps, int64_t every);
/* The same 8 columns for the env's current state, momentum from the current
 * velocities (as after update_acc!, which calls update_misc!). */
int pd_env_observables(void* ctx, int64_t env, double* out);
/* The 8 columns of pd_env_observables written into out_dev, 8 doubles of
 * device memory, without allocating: no hipMalloc, hipFree or other allocating
 * call, so it may be used under HIP graph capture or where a serving loop
 * forbids allocation.  No counterpart in the reference; it is for callers that
 * cannot allocate.  The partial sums go to a fixed 8 MiB scratch array per GPU
 * (one double4 per chunk of the env): envs of more than 262,144 chunks are
 * rejected with PD_E_ARG.  Asynchronous: it enqueues on the context's stream
 * and returns; pd_sync (or a wait on that stream) precedes reading out_dev.
 * Every call on one GPU shares that GPU's scratch, so calls that overlap on one
 * GPU (two contexts, or two host threads) race. */
int pd_env_observables_noalloc(void* ctx, int64_t env, double* out_dev);
/* update_acc! (solvers.jl:320-336) at the current positions; y may be NULL
 * (resident state) or 3 x N positions to upload first.  f receives 3 x N. */
int pd_env_update_acc(void* ctx, int64_t env, const double* y, double* f);
/* force_density_T!(f, y, limits, mat, Val{:hip}) for one block: adds the block's
 * material force density into f[:, lo:hi] (1-based inclusive limits, solvers.jl:236-237). */
int pd_force_density(void* ctx, int64_t env, int64_t block, const double* y, double* f, int64_t lo, int64_t hi);
/* update_contact_neighs! for every contact model of the env at the resident positions */
int pd_env_update_neighs(void* ctx, int64_t env);
/* update_contact_neighs!(y, type, RM) for contact model cidx of the env (its
 * index in pd_env_add_contact order): RM12 (contacts.jl:434-476) or RM11
 * (contacts.jl:539-591).  y may be NULL (resident positions) or 3 x N env
 * positions, uploaded into the resident state first. */
int pd_contact_update(void* ctx, int64_t env, int64_t cidx, const double* y);
/* RM.neighs as update_contact_neighs! leaves it: max_neighs x n1 column-major
 * (n1 = members of pair[1], in type-mask order), for each member the largest
 * pair[2] member indices within `distance` (1-based, descending, 0-padded);
 * for RM11, n1 = the block's particles and indices into the block.  max_neighs
 * must equal the model's min(max_neighs, n2). */
int pd_contact_neighs(void* ctx, int64_t env, int64_t cidx, int64_t* neighs, int64_t max_neighs);
/* short_range_contact!(y, f, type, bid, vol, den, RM) for contact model cidx
 * (RM12 contacts.jl:189-247, RM11 :357-380): adds its terms into f (3 x N, in
 * and out) one by one in the single-threaded reference's order, from the lists
 * of the last update; y may be NULL (resident positions) or 3 x N positions to
 * upload first.  vol and den are the env's own (environment.jl:173-214). */
int pd_contact_force(void* ctx, int64_t env, int64_t cidx, const double* y, double* f);

/* ---- state transfer ----------------------------------------------------- */
int pd_env_download(void* ctx, int64_t env, double* y, double* v, double* f);
int pd_env_upload(void* ctx, int64_t env, const double* y, const double* v, const double* f);
int pd_env_damage(void* ctx, int64_t env, double* damage);        /* cal_damage, peri.jl:587-594 */
int pd_env_time_step(void* ctx, int64_t env, int64_t* time_step);
int pd_env_check_finite(void* ctx, int64_t env);                  /* @check_nan, const_macros.jl:60-74 */
/* Quasi-static relaxation, QSDrag minimize! (src/solvers/minimize.jl:105-231):
 * BCs applied once (position, velocity), then up to max_iter iterations of
 * update_acc! followed by the drag update of the particles held by no BC
 *   force = f - drag*opt*(1 + drag/10 |opt|);  dy = clip(0.5 force step^2 +
 *   opt step, ps/10);  opt += step force;  y += dy
 * (ps: particle size of the first block), stopping at the first iteration with
 * max|f| < f_tol and max|dy| < x_tol; then check! and time_step += 1.  iters,
 * f_max, x_max (may be NULL) receive the iteration count and the last maxima. */
int pd_env_minimize(void* ctx, int64_t env, double step_size, double drag, int64_t max_iter, double x_tol,
                    double f_tol, int64_t* iters, double* f_max, double* x_max);
/* Resident-state snapshot (the restart role of save_to_file / load_from_file,
 * src/io/io.jl:170-187): every piece of state a run changes (positions,
 * velocities, forces, bond states, theta, td_norm2 and their per-particle
 * records, plastic extensions, deformed latches, time step) as an opaque blob
 * of pd_env_state_size bytes.  pd_env_state_load restores it into an env of the
 * same shape, form and schedule (e.g. built again from the same description);
 * a run resumed from it is bit-identical to the uninterrupted run. */
int pd_env_state_size(void* ctx, int64_t env, int64_t* bytes);
int pd_env_state_save(void* ctx, int64_t env, void* buf, int64_t bytes);
int pd_env_state_load(void* ctx, int64_t env, const void* buf, int64_t bytes);

/* ---- measurement ------------------------------------------------------------
 * No reference counterpart (the reference's only timing hook is TimerOutputs,
 * src/misc/const_macros.jl:7).  When enabled, every launch of a kernel class in
 * pd_run is bracketed by HIP events on the context stream; reads return the
 * summed device time and the launch count since the last enable. */
enum {
  PD_PROF_PROLOGUE = 0, PD_PROF_LATCH = 1, PD_PROF_THETA = 2, PD_PROF_TD = 3,
  PD_PROF_FORCE_BB = 4, PD_PROF_FORCE_OSB = 5, PD_PROF_FORCE_EP = 6,
  PD_PROF_CONTACT = 7, PD_PROF_BOOK = 8, PD_PROF_NCLASS = 9
};
int pd_ctx_profile(void* ctx, int32_t enable);
int pd_ctx_profile_read(void* ctx, int32_t cls, double* total_ms, int64_t* launches);

#ifdef __cplusplus
}
#endif
#endif /* PERIDYN_H */
