/*
 * fmmh.h — C ABI of the MI355X (gfx950) FMM force evaluator.
 *
 * Drop-in for the force evaluation of the reference solver `crates/fmm`
 * (Vedant9500/FMMS): bodies in (positions, masses, particle-id order),
 * accelerations and potentials out (particle-id order), at the reference's
 * expansion order p, MAC tolerance eps, leaf capacity n_leaf and centre mode.
 *
 * Conventions (as the reference): G = 1; potential is positive, pot_i =
 * sum_j m_j / r_ij; acceleration a_i = sum_j m_j (x_j - x_i) / r_ij^3.
 * All buffers are caller-owned; no allocation crosses the ABI. One context
 * per host thread; every call is synchronous (returns when results are in
 * the output buffers), like Fmm::run.
 *
 * Reference interfaces replaced (file:line in crates/fmm unless noted):
 *   fmmh_create/fmmh_build   Fmm::new, Fmm::with_center   fmm.rs:26-49
 *   fmmh_build (again)       Fmm::refresh                  fmm.rs:54-56
 *   fmmh_evaluate            Fmm::run / run_threaded       fmm.rs:60-67 -> Result fmm.rs:17-21
 *   fmmh_accel               Physics::accel (FmmBackend)   crates/visual/src/sim.rs:5-7,124-150
 *   fmmh_get_tree            Fmm.tree (Octree, Cell)       tree.rs:49-70
 *   fmmh_get_lists           Recorded {pairs, p2p}         walk.rs:334-340
 *   fmmh_plummer/hernquist   models::plummer/hernquist     models.rs:35-82
 *   fmmh_direct_sampled      direct + sampled_error        direct.rs:6-31, crates/bench/src/main.rs:34-78
 */
#ifndef FMMH_H
#define FMMH_H
#include <stddef.h>
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

/* Return codes. The reference panics on empty input (tree.rs:95) and on
 * p > 60 (rotation.rs:126); GpuError::{OverBudget, OverfullLeaf}
 * (crates/visual/src/gpufmm.rs:1656-1691) are the device-side failures. */
#define FMMH_OK 0
#define FMMH_E_EMPTY 1         /* n == 0 */
#define FMMH_E_PARAM 2         /* p out of range, n_leaf < 1, theta_max <= 0, ... */
#define FMMH_E_OVER_BUDGET 3   /* device memory would exceed the context budget */
#define FMMH_E_OVERFULL_LEAF 4 /* reserved: kernels here accept any leaf occupancy */
#define FMMH_E_NONFINITE 5     /* non-finite position or mass in the input */
#define FMMH_E_HIP 6           /* a HIP runtime call failed */
#define FMMH_E_STATE 7         /* evaluate before build, or n differs from the build */

/* Centre modes (tree.rs:19-25) */
#define FMMH_CENTER_COM 0
#define FMMH_CENTER_SPHERE 1

/* Precision forms */
#define FMMH_PREC_FP64 0      /* plain FP64 everywhere (the drop-in) */
#define FMMH_PREC_FP32BULK 1  /* mixed precision: P2P pair terms in FP32 on positions
                                 relative to the target leaf's centre; for p <= 10
                                 also P2M, M2M below the top three levels of each
                                 system, M2L, L2L and L2P in FP32 on radius-scaled
                                 operands (per-cell scale, scaled moments and locals,
                                 dimensionless offsets); FP64 positions, tree, top-level
                                 M2M, tile and chunk sums, softened-field trace terms,
                                 and a compensated (TwoSum) final per-body near + far
                                 sum. With a finite eps_mac the upward pass stays FP64
                                 (the error MAC reads every cell's FP64 moments); for
                                 p > 10 the far field runs the FP64 kernels. */

/* Flags */
#define FMMH_FLAG_TIMING 1u   /* record per-stage device times (hipEvents) */

typedef struct fmmh_params {
  int p;            /* expansion order (Fmm.p); 1..FMMH_MAX_P (rotation M2L from p = 20) */
  double eps_mac;   /* error-controlled MAC tolerance (Fmm.eps, mac.rs:8-11,92);
                       +INFINITY = theta-only MAC, no f pre-pass */
  double theta_max; /* opening angle bound; 1.0 = the reference's hard-coded theta<1 */
  int n_leaf;       /* leaf capacity (Fmm.n_leaf) */
  double softening; /* Plummer softening of the kernel (near field exact, far field through
                       the softened expansions); 0 = the reference (pure 1/r) */
  int center_mode;  /* FMMH_CENTER_* */
  int precision;    /* FMMH_PREC_* */
  unsigned flags;   /* FMMH_FLAG_* */
  double hbm_budget_bytes; /* device-memory budget of this context; 0 = no limit */
  double soft_mac;  /* softening guard of the MAC (extension, softening > 0 only): a pair
                       is rejected for M2L when softening^2 (rhoA + rhoB)^3 > soft_mac r^5,
                       the softened far field's leading residual relative to the pair's
                       field; 0 = off (default) */
} fmmh_params;

#define FMMH_MAX_P 40

typedef struct fmmh_ctx fmmh_ctx;

/* Fills the reference's parameters of Fmm::new(parts, p=8, eps=1e-5, n_leaf=16). */
void fmmh_default_params(fmmh_params* out);

int fmmh_create(const fmmh_params* params, fmmh_ctx** out);
void fmmh_destroy(fmmh_ctx* ctx);
const char* fmmh_last_error(const fmmh_ctx* ctx);

/* Host-buffer entry points. pos [n][3] and mass [n] in particle-id order. */
int fmmh_build(fmmh_ctx* ctx, size_t n, const double* pos, const double* mass);
int fmmh_evaluate(fmmh_ctx* ctx, size_t n, const double* pos, const double* mass,
                  double* accel_out /*[n][3]*/, double* pot_out /*[n] or NULL*/);
int fmmh_accel(fmmh_ctx* ctx, size_t n, const double* pos, const double* mass,
               double* accel_out, double* pot_out);

/* Device-buffer entry points (pointers into device memory, same layouts).
 * Results are complete in the output buffers when the call returns. */
int fmmh_build_device(fmmh_ctx* ctx, size_t n, const double* d_pos, const double* d_mass);
int fmmh_evaluate_device(fmmh_ctx* ctx, size_t n, const double* d_pos, const double* d_mass,
                         double* d_accel, double* d_pot);
int fmmh_accel_device(fmmh_ctx* ctx, size_t n, const double* d_pos, const double* d_mass,
                      double* d_accel, double* d_pot);

/* Tile shapes of the evaluation, settable per context (no rebuild needed;
 * they take effect at the next full evaluation). They replace the reference
 * GPU app's compile-time constants: its P2P split of leaves with many partner
 * bodies (FAT_THR / FAT_CHUNK, crates/visual/src/gpufmm.rs:686-689) and its
 * fixed preset (gpufmm.rs:55-58, whose p and n_leaf are fmmh_params here).
 * fmms_amd.autotune picks them (with p and n_leaf) by measurement. */
typedef struct fmmh_tuning {
  int m2l_chunk;          /* sources per M2L work chunk (one thread's serial work), 1..1024;
                             default 32 */
  uint64_t p2p_item_work; /* a leaf whose targets x sources exceed this is cut into P2P work
                             items of about equal source count, 64..2^40; default 2^18 */
} fmmh_tuning;
void fmmh_default_tuning(fmmh_tuning* out);
int fmmh_set_tuning(fmmh_ctx* ctx, const fmmh_tuning* t);
int fmmh_get_tuning(const fmmh_ctx* ctx, fmmh_tuning* out);

/* Device bytes this context holds (its arena of stage buffers); callers that
 * admit work under a memory budget read it after a call. No reference
 * counterpart (the reference's GPU app sizes buffers up front,
 * crates/visual/src/gpufmm.rs:1656-1691). */
size_t fmmh_device_bytes(const fmmh_ctx* ctx);

/* Tree of the last build (tree.rs:49-70), DFS-preorder cell ids. Any
 * output pointer may be NULL. children [nc][8], -1 where absent; cen [nc][3]. */
size_t fmmh_num_cells(const fmmh_ctx* ctx);

/* Batched ensembles (no reference counterpart: the reference evaluates one
 * system per Fmm). After this call the context treats its n = offsets[nsys]
 * bodies as nsys independent systems with bodies offsets[s]..offsets[s+1]-1
 * (host array, offsets[0] = 0, non-decreasing, nsys <= FMMH_MAX_SYSTEMS): one
 * forest whose top V = ceil(log8 nsys) levels select the system and whose
 * subtrees are the systems' own octrees (each keyed in its own bounding box,
 * at 21 - V bits per axis), walked from the systems' roots so no interaction
 * crosses systems; every stage runs once for the whole ensemble. Every
 * evaluation and leapfrog entry point then works on the ensemble (the
 * leapfrog's dt clamp uses the ensemble-wide maximum acceleration). Both
 * MACs work per system (the error-controlled MAC's f pre-pass walks each
 * body's own system only). nsys = 1 restores the single-system mode. */
#define FMMH_MAX_SYSTEMS 4096
int fmmh_set_systems(fmmh_ctx* ctx, size_t nsys, const uint64_t* offsets);

/* Temporal reuse — the frozen step of crates/visual/src/gpufmm.rs:2710-2741
 * (refresh model gpufmm.rs:9-20): evaluate on the tree and interaction lists
 * of this context's last full evaluation. Bodies are re-gathered into the
 * same slots, moments and locals are recomputed about the frozen cell
 * centres, no tree build and no walk run. Needs a previous fmmh_evaluate* or
 * fmmh_accel* with the same n (FMMH_E_STATE otherwise). The reference keeps
 * a tree for at most k <= 8 steps. */
int fmmh_evaluate_frozen(fmmh_ctx* ctx, size_t n, const double* pos, const double* mass,
                         double* accel, double* pot);
int fmmh_evaluate_frozen_device(fmmh_ctx* ctx, size_t n, const double* d_pos,
                                const double* d_mass, double* d_accel, double* d_pot);

/* Kick-drift-kick leapfrog on device-resident state (crates/visual/src/sim.rs:539-594;
 * the energy-gate loop of crates/fmm/tests/energy_gate.rs:156-177): per step
 *   dt = min(dt, dt_eta * sqrt(dt_eps_len / max|a|)) when dt_eta > 0 (sim.rs:19-20,553-557),
 *   v += a dt/2, x += v dt, a = accel(x), [cap |a| (sim.rs:14-33) when accel_cap > 0],
 *   v += a dt/2.
 * The tree and lists are rebuilt when they have served refresh_every
 * evaluations (1 = every step, as the energy gate does); steps in between are
 * frozen evaluations. Drift monitor (drift_tol > 0): a frozen step that
 * leaves any body farther than rho + drift_tol * max(rho, side/2) from its
 * leaf's frozen centre (rho, side: the leaf's radius and octant edge at the
 * build) makes the next step rebuild — expansions about frozen centres
 * diverge once bodies leave their cells (the reference's age rule,
 * gpufmm.rs:14-20, bounds the same hazard by step count only). With FMMH_LF_GRAPH the frozen step (dt, kicks, drift and
 * the whole force evaluation) is captured once per tree into a hipGraph and
 * replayed. FMMH_LF_ACC_VALID: accel/pot already hold a(x) of the entry
 * positions (else they are evaluated first); when this context has no tree
 * for them, the first step rebuilds. pos/vel/accel [n][3], pot [n], mass [n] are device pointers in id
 * order; on return accel/pot belong to the final positions and
 * *time_advanced (host, may be NULL) is the sum of the steps' dt. */
#define FMMH_LF_ACC_VALID 1u
#define FMMH_LF_GRAPH 2u
typedef struct fmmh_leapfrog_opts {
  double dt;
  int steps;
  int refresh_every;
  double dt_eta;     /* 0: fixed dt; sim.rs uses 0.25 */
  double dt_eps_len; /* sim.rs uses 0.02 */
  double accel_cap;  /* 0: off; sim.rs uses 50 */
  double drift_tol;  /* > 0: drift monitor (default 0.25), 0: off */
  unsigned flags;
} fmmh_leapfrog_opts;
void fmmh_leapfrog_default(fmmh_leapfrog_opts* out);
/* Counters of this context's leapfrog calls: [steps, tree rebuilds, frozen
 * steps, rebuilds forced by the drift monitor]. */
int fmmh_leapfrog_stats(const fmmh_ctx* ctx, uint64_t* out4);
int fmmh_leapfrog_device(fmmh_ctx* ctx, size_t n, double* d_pos, double* d_vel,
                         const double* d_mass, double* d_accel, double* d_pot,
                         const fmmh_leapfrog_opts* opts, double* time_advanced);
int fmmh_get_tree(fmmh_ctx* ctx, uint32_t* order, uint64_t* keys_sorted, uint32_t* first,
                  uint32_t* count, int32_t* children, uint8_t* leaf, uint32_t* level,
                  double* cen, double* mass, double* rho);

/* Interaction lists of the last evaluation (walk.rs:94-190): mutual M2L
 * pairs (a,b) and leaf-leaf P2P pairs (a,b) including self pairs (a,a),
 * each as [k][2] uint32 in an unspecified order. */
int fmmh_get_list_sizes(fmmh_ctx* ctx, size_t* n_m2l, size_t* n_p2p);
int fmmh_get_lists(fmmh_ctx* ctx, uint32_t* m2l, uint32_t* p2p);

/* Intermediates of the last evaluation (complex arrays interleaved re,im). */
int fmmh_get_moments(fmmh_ctx* ctx, double* out /*[nc][(p+1)^2][2]*/);
int fmmh_get_locals(fmmh_ctx* ctx, double* out /*after L2L*/);
int fmmh_get_pows(fmmh_ctx* ctx, double* out /*[nc][p+1]*/);
int fmmh_get_f(fmmh_ctx* ctx, double* f_slot /*[n]*/, double* fmin /*[nc]*/);

/* Per-stage device times (ms) of the last call when FMMH_FLAG_TIMING is set.
 * Writes up to max entries, returns the number of stages; names via
 * fmmh_stage_name. */
int fmmh_stage_times(const fmmh_ctx* ctx, double* ms, int max);
const char* fmmh_stage_name(int stage);
/* Work counts of the last evaluation: [n, ncells, nleaves, n_m2l_mutual,
 * n_p2p_pairs(incl. self), p2p_directed_interactions, max_level, walk_rounds]. */
int fmmh_work_counts(const fmmh_ctx* ctx, uint64_t* out8);
/* P2P work plan of the current lists (no counterpart in the reference, whose
 * threads take leaf pairs from a shared list): [work items, leaves cut into
 * several items]. A leaf whose targets x sources exceed 2^18 is cut into
 * items of about equal source count, added back in a fixed order. */
int fmmh_p2p_plan_counts(const fmmh_ctx* ctx, uint64_t* out2);

/* FP64 direct sum at sampled targets, on the device: acc [nt][3], pot [nt]
 * and f [nt] (= sum m/r^2) for target ids idx[nt]; host arrays. */
int fmmh_direct_sampled(size_t n, const double* pos, const double* mass, double softening,
                        size_t nt, const uint32_t* idx, double* acc, double* pot, double* f);

/* The same FP64 direct sum, allocating nothing. No counterpart in the
 * reference: for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Device pointers: pos [n][3], mass [n], idx [nt] (ids out of range give NaN
 * sums), acc [nt][3]; pot [nt] and f [nt] may be NULL. The kernels are
 * enqueued on `stream` (a hipStream_t; NULL = the null stream) and the call
 * returns without waiting: the results are in acc/pot/f once that stream's
 * work is done. Partial sums live in a static 8 MiB device array of the
 * library (one per GPU, held until the process ends), so
 * nt <= FMMH_DIRECT_NOALLOC_MAX_TARGETS (FMMH_E_PARAM beyond) and calls that
 * overlap on one GPU race. Each call overwrites every partial it reads.
 * Results equal fmmh_direct_sampled's bit for bit. */
#define FMMH_DIRECT_NOALLOC_MAX_TARGETS 209715
int fmmh_direct_sampled_noalloc(size_t n, const double* d_pos, const double* d_mass,
                                double softening, size_t nt, const uint32_t* d_idx, double* d_acc,
                                double* d_pot, double* d_f, void* stream);

/* Initial conditions (models.rs:13-82, seeded SplitMix64), host arrays. */
void fmmh_plummer(size_t n, const double center[3], uint64_t seed, double* pos, double* mass);
void fmmh_hernquist(size_t n, const double center[3], uint64_t seed, double* pos, double* mass);
/* Plummer truncated at r <= rmax (radius redrawn until inside); an extension
 * of models.rs:35-58 for the r <= 20 workloads. */
void fmmh_plummer_trunc(size_t n, const double center[3], double rmax, uint64_t seed, double* pos,
                        double* mass);

/* Initial-condition library with velocities (host arrays pos/vel [n][3],
 * mass [n]; G = 1, total mass 1; FMMH_OK or FMMH_E_PARAM). Extensions of
 * models.rs:35-82 (positions drawn as there) that replace the sigma(r)
 * Maxwellian of crates/fmm/tests/energy_gate.rs:39-79; no reference
 * counterpart beyond those.
 *   fmmh_ic_plummer    Plummer (a = 1) truncated at rmax, isotropic
 *                      distribution-function velocities (Aarseth, Henon &
 *                      Wielen 1974); positions equal fmmh_plummer_trunc's.
 *   fmmh_ic_hernquist  Hernquist (scale a) truncated at rmax, Maxwellian of
 *                      the isotropic Jeans dispersion, |v| < 0.95 v_esc.
 *   fmmh_ic_disk       exponential disk (scale length rd, sech^2 height zd,
 *                      R <= rmax) on circular orbits, v_c^2 = m_central/R +
 *                      Freeman's thin-disk term, plus isotropic Gaussian
 *                      dispersion sigma; m_central > 0 puts body 0 at rest at
 *                      the origin with that mass.
 *   fmmh_ic_halos_cube nhalo Hernquist halos carrying halo_frac of the mass
 *                      (masses log-uniform over mass_decades, scale radii
 *                      log-uniform in [a_min, a_max], centres uniform in the
 *                      unit cube, truncated at rmax_over_a * a) and a cold
 *                      uniform unit-cube background with the rest; equal
 *                      body masses 1/n, halo bodies first.
 * All but the halos have their mass-weighted mean velocity removed. */
int fmmh_ic_plummer(size_t n, double rmax, uint64_t seed, double* pos, double* vel, double* mass);
int fmmh_ic_hernquist(size_t n, double a, double rmax, uint64_t seed, double* pos, double* vel,
                      double* mass);
int fmmh_ic_disk(size_t n, double rd, double zd, double rmax, double m_central, double sigma,
                 uint64_t seed, double* pos, double* vel, double* mass);
int fmmh_ic_halos_cube(size_t n, int nhalo, double halo_frac, double a_min, double a_max,
                       double mass_decades, double rmax_over_a, uint64_t seed, double* pos,
                       double* vel, double* mass);

#ifdef __cplusplus
}
#endif
#endif
