/*
 * fmm_oracle.h — C API of the CPU oracle: a restatement of the reference
 * solver `crates/fmm` (Vedant9500/FMMS) in plain C++ / FP64.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker the GPU product is
 * compared against. Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it; the product never links or calls it.
 *
 * Arithmetic follows the reference operation by operation and is compiled
 * with -ffp-contract=off (Rust never fuses a*b+c), so trees, moments, MAC
 * decisions and interaction lists reproduce the reference bit for bit.
 * Extensions (all off by default, reference semantics when off):
 *   theta_max < 1  : r^2*theta_max^2 <= (rhoA+rhoB)^2 rejects (mac.rs:80-84 has 1)
 *   eps_mac = +inf : skip the error test and the f pre-pass (theta-only MAC)
 *   softening > 0  : Plummer-softened P2P / direct sum (reference is pure 1/r)
 */
#ifndef FMM_ORACLE_H
#define FMM_ORACLE_H
#include <stddef.h>
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_ctx orc_ctx;

/* models.rs:13-82 */
void orc_plummer(size_t n, const double center[3], uint64_t seed, double* pos, double* mass);
void orc_hernquist(size_t n, const double center[3], uint64_t seed, double* pos, double* mass);
double orc_rng_unit_first(uint64_t seed, size_t k, double* out); /* first k draws of Rng::unit */

/* Initial-condition library (ic_oracle.cpp; extensions of models.rs:35-82,
 * same arguments as the product's fmmh_ic_*; 0 = ok, 2 = bad argument). */
int orc_ic_plummer(size_t n, double rmax, uint64_t seed, double* pos, double* vel, double* mass);
int orc_ic_hernquist(size_t n, double a, double rmax, uint64_t seed, double* pos, double* vel,
                     double* mass);
int orc_ic_disk(size_t n, double rd, double zd, double rmax, double m_central, double sigma,
                uint64_t seed, double* pos, double* vel, double* mass);
int orc_ic_halos_cube(size_t n, int nhalo, double halo_frac, double a_min, double a_max,
                      double mass_decades, double rmax_over_a, uint64_t seed, double* pos,
                      double* vel, double* mass);
double orc_hernquist_sigma2(double r, double a);
/* crates/fmm/tests/energy_gate.rs:39-123: warm start, T, W, |sum m v| / sum m|v| */
void orc_warm_start(size_t n, const double* pos, const double* mass, const double* pot,
                    double* vel);
double orc_kinetic(size_t n, const double* mass, const double* vel);
double orc_potential_energy(size_t n, const double* mass, const double* pot);
double orc_momentum_rel(size_t n, const double* mass, const double* vel);

/* tree.rs:28-46 */
uint64_t orc_dilate(uint64_t x);
uint64_t orc_morton_key(double x, double y, double z, const double lo[3], double inv_side);

/* harmonic.rs:91-173 (tables are dim(p) complex, interleaved re,im) */
void orc_theta_table(const double r[3], int p, double* out);
void orc_upsilon_table(const double r[3], int p, double* out);

/* moments.rs:11-27 */
void orc_factorials(int p, double* out /*2p+1*/);
void orc_binomials(int p, double* out /*p+1*/);

/* translate.rs:47-136 mutual naive M2L on explicit centres/moments */
void orc_m2l_mutual(int p, const double cen_a[3], const double cen_b[3], const double* ma,
                    const double* mb, double* la, double* lb);
/* Softening extension of the far field: mutual M2L with the harmonic
 * derivatives of the Plummer kernel; one directed pair's trace terms (the
 * source-side field (q/6) Lap K to second order into f_dst, the target-side
 * Laplacian field of the source (L, grad L, Hessian of L: 10 doubles, see
 * SOFT_LAP_WORDS) added into lap10_dst); and the target-side L2P polynomial of
 * one cell's lap10 (pot += |d|^2 L / 6 + |d|^2 (d.G) / 10 + |d|^2 (d.H.d) / 28
 * - |d|^4 tr(H) / 280, acc = its gradient). */
void orc_m2l_mutual_soft(int p, const double cen_a[3], const double cen_b[3], const double* ma,
                         const double* mb, double eps2, double* la, double* lb);
void orc_soft_trace_pair(int p, const double cen_src[3], const double cen_dst[3], double q,
                         double mass, double eps2, double* f_dst, double* lap10_dst);
void orc_l2p_trace(const double cen[3], const double pos[3], const double* lap10, double* pot,
                   double acc[3]);
/* rotation.rs:262-435 mutual rotation M2L on explicit centres/moments */
void orc_m2l_rotation(int p, const double cen_a[3], const double cen_b[3], const double* ma,
                      const double* mb, double* la, double* lb);
/* rotation.rs:125-211 swap matrices B_n, block layout, blocksum(p) reals */
size_t orc_rot_blocksum(int p);
void orc_rot_swap(int p, double* out);
/* translate.rs:175-206 / 271-309 on explicit arrays */
void orc_l2l(int p, const double parent_cen[3], const double child_cen[3], const double* fpar,
             double* fchild);
void orc_l2p(int p, const double cen[3], const double pos[3], const double* f, double* pot,
             double acc[3]);

/* direct.rs:6-78; softening 0 = reference. acc [n][3], pot [n] (id order) */
void orc_direct(size_t n, const double* pos, const double* mass, double softening, double* acc,
                double* pot);
/* Targets-only direct sum (sampled error, crates/bench/src/main.rs:34-78):
 * acc[t][3], pot[t], f[t] = sum m/r^2 for target ids idx[t]. nthreads>=1. */
void orc_direct_targets(size_t n, const double* pos, const double* mass, double softening,
                        size_t nt, const uint32_t* idx, double* acc, double* pot, double* f,
                        int nthreads);
/* The bench's LCG sample picker (crates/bench/src/main.rs:41-47) */
void orc_sample_ids(size_t n, size_t samp, uint64_t seed, uint32_t* out);

/* Whole solver: Fmm::new / refresh / run / run_threaded (fmm.rs:26-142). */
orc_ctx* orc_new(void);
void orc_free(orc_ctx*);
int orc_build(orc_ctx*, size_t n, const double* pos, const double* mass, int n_leaf,
              int center_mode /*0 Com, 1 Sphere*/);
size_t orc_ncells(const orc_ctx*);
uint32_t orc_max_level(const orc_ctx*);
void orc_get_tree(const orc_ctx*, uint32_t* order, uint64_t* keys_sorted, uint32_t* first,
                  uint32_t* count, int32_t* children /*[nc][8]*/, uint8_t* leaf, uint32_t* level,
                  double* cen /*[nc][3]*/, double* mass, double* rho);
void orc_get_bbox(const orc_ctx*, double lo[3], double hi[3], double* inv_side);
/* nthreads<=1: serial reference order (dual_walk + inline mutual kernels).
 * nthreads>1 : run_threaded analogue (level-parallel passes, recorded walk,
 * threaded mutual replay with per-thread accumulators, reduce).
 * accel [n][3], pot [n] in particle-id order. Returns 0 on success. */
/* Softening guard of the MAC (extension, used by later orc_run calls with
 * softening > 0): a pair is rejected for M2L when
 * eps^2 (rhoA + rhoB)^3 > tau r^5; tau = 0 (default) turns it off. */
void orc_set_soft_mac(orc_ctx*, double tau);
int orc_run(orc_ctx*, const double* pos, const double* mass, int p, double eps_mac,
            double theta_max, double softening, int nthreads, double* accel, double* pot);
/* Intermediates of the last orc_run (complex arrays interleaved re,im). */
int orc_dim(const orc_ctx*);
void orc_get_moments(const orc_ctx*, double* out /*nc*dim*2*/);
void orc_get_pows(const orc_ctx*, double* out /*nc*(p+1)*/);
void orc_get_f(const orc_ctx*, double* out /*n, slot order*/);
void orc_get_fmin(const orc_ctx*, double* out /*nc*/);
size_t orc_n_m2l(const orc_ctx*);
size_t orc_n_p2p(const orc_ctx*);
void orc_get_m2l(const orc_ctx*, uint32_t* ab /*[n][2]*/);
void orc_get_p2p(const orc_ctx*, uint32_t* ab /*[n][2]*/);
void orc_get_locals_walk(const orc_ctx*, double* out); /* after walk, before L2L */
void orc_get_locals(const orc_ctx*, double* out);      /* after L2L */
void orc_get_near(const orc_ctx*, double* out /*[n][4] slot order ax,ay,az,pot*/);
/* Timings of the last orc_run, ms: moments, prepass, walk, l2l, l2p. */
void orc_get_times(const orc_ctx*, double* out5);

#ifdef __cplusplus
}
#endif
#endif
