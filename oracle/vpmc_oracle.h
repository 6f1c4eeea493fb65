/*
 * vpmc_oracle.h -- CPU restatement of the reference's vortex-particle hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.  The
 * product (vortexparticlemc_cuda_amd, libvpmc.so) never links or calls it.
 *
 * Every function cites the reference file:line whose behaviour it restates
 * (paths relative to the reference repository root).  Functions that have no
 * reference counterpart (the consistent estimator, the periodic kernel, the
 * FP64 direct sums, Philox) say so and name the mathematics they follow.
 *
 * Conventions shared with the HIP library (they define the build, not the
 * reference):
 *   - uniforms: u = (x >> 8) * 2^-24 in [0,1) from a 32-bit draw x;
 *   - Philox4x32-10 counter for a sample:
 *       c = { k, owner, step, (sim << 8) | (sign_idx << 7) | (branch << 4) | block }
 *       key = { seed & 0xffffffff, seed >> 32 }
 *     sign_idx 0 = positive set, 1 = negative set; branch 1 or 2;
 *   - replay uniforms (injected): per (owner, sign_idx) a block of
 *       N1*b1 + N2*b2 floats, b1 = 3 (1 for reference mode without periodicity),
 *       b2 = 2 (10 when the replayed stream carries the aux-target draws),
 *     laid out owner-major, sign-minor, in the order the reference's
 *     monte_carlo_bs_cpu consumes its std::mt19937 draws
 *     (src/cpu/monte_carlo_bs.cpp:44-119, aux = 0).
 */
#ifndef VPMC_ORACLE_H
#define VPMC_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- kernel W and its radial mass ---------------------------------------- */
float  orc_spline_f(float r, float h_w);            /* include/utilities.hpp:7-17 */
double orc_spline_d(double r, double h_w);
double orc_mass_fraction(double q);                  /* enclosed mass of W within q*h_w */
double orc_mass_inverse(double u);                   /* q with orc_mass_fraction(q) = u */

/* ---- spatial hash (src/cpu/hash.cpp) -------------------------------------- */
void orc_fill_cells(int rshx, int rshy, float Lx, float Ly, int n, const int *sub,
                    const float *px, const float *py, int *cell_num, int *cell_acc,
                    int *vp_cell);
void orc_counting_sort(int n, int *cell_num, const int *cell_acc, const int *vp_cell,
                       int *vp_map);

/* ---- vorticity query (src/cpu/evaluate_vorticity.cpp) --------------------- */
typedef struct {
    const float *px, *py, *w;   /* global particle arrays                       */
    int          N;             /* particles visited by the naive branch         */
    const int   *vp_sort;       /* NULL -> naive O(N) branch                     */
    const int   *cell_acc;
    int          rshx, rshy;
    float        Lx, Ly;
    int          cnt_view;      /* end of the last cell                          */
    const int   *sub;           /* local -> global                              */
    const float *F;             /* NULL, or per global particle the deformation
                                   F = (a b; c d) as 4 floats: the CUDA query's
                                   inv(F) d with minimum image and 2 h_w cutoff
                                   (include/vorticity_query.cuh:52-66)           */
} orc_view;

float orc_eval_vorticity(const orc_view *v, float px, float py, int sign, int use_signed,
                         int exclude_center, float h_w, int periodic);

/* ---- sign subsets and FP64 CDF (src/cpu/build_particle_sets.cpp:6-42) ------ */
void orc_build_subsets(int n, const float *omega, int *sub_pos, int *sub_neg,
                       double *cdf_pos, double *cdf_neg, int *cnt /*[2]*/,
                       double *cum /*[2]*/);
int  orc_sample_idx(float u, const double *cdf, int n);   /* monte_carlo_sampling.cpp:23-33 */

/* ---- Philox4x32-10 (Salmon et al., SC'11) --------------------------------- */
void  orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
float orc_u01(uint32_t x);

/* ---- periodic Biot-Savart kernel (no reference counterpart) -------------- */
void orc_kper_coeffs(int nterms, double *c);       /* c_m = G_{4m}(Z+iZ), m = 1..nterms */
void orc_kernel_smooth_d(double rx, double ry, double L, int nterms, const double *c,
                         double *kx, double *ky);

/* ---- Monte Carlo Biot-Savart estimator ------------------------------------ */
enum { ORC_MODE_REFERENCE = 0, ORC_MODE_CONSISTENT = 1 };

typedef struct {
    int      mode;
    int      N1, N2;
    float    h_w, Lx, Ly;
    int      periodic;
    uint64_t seed;
    uint32_t step, sim;
    const float *U;          /* injected uniforms or NULL (Philox)            */
    int      nterms;         /* periodic-kernel series length (consistent)    */
    const double *kper_c;    /* orc_kper_coeffs output                        */
    int      aux_draws;      /* replay stride: 1 -> each branch-2 sample owns 10
                                draws (main disk first), as the reference's
                                step_cpu with aux = 1 consumes them            */
} orc_mc_params;

typedef struct {
    int           n;          /* particles in the set                         */
    const int    *sub;        /* local -> global                              */
    const double *cdf;        /* normalised inclusive prefix of |w|           */
    double        omega;      /* sum |w|                                      */
    orc_view      view;       /* hashed view for the query                    */
} orc_sign_set;

/* velocity of one sign set at owner `owner` (dst = its position); out[2] */
void orc_mc_velocity(const orc_mc_params *P, const orc_sign_set *S, int sign_idx,
                     int owner, float dstx, float dsty, const float *px, const float *py,
                     float *out);
/* both signs for owners [0, n_owner), then forward Euler + wrap
   (src/cpu/simulation_step.cpp:28-71 with aux = 0).  vel[2*i] out. */
void orc_step(const orc_mc_params *P, const orc_sign_set *pos_set,
              const orc_sign_set *neg_set, int n_owner, float *px, float *py,
              float dt, float *vel, int nthreads);

/* ---- FP64 direct sums (oracle side) --------------------------------------- */
/* src/cuda/naive_biot_savart.cu:15-40 in FP64: eps-regularised, min image.
   skip_self >= 0 skips particle index == probe index (probe = particle). */
void orc_direct_eps(int n_probe, const double *qx, const double *qy, int skip_self,
                    int n, const double *px, const double *py, const double *g,
                    double L, int periodic, double eps, double *u, int nthreads);
/* exact velocity of the blob field sum_j g_j W(|y - x_j|) under the
   doubly-periodic kernel (consistent-mode expectation). */
void orc_direct_blob(int n_probe, const double *qx, const double *qy, int n,
                     const double *px, const double *py, const double *g, double h_w,
                     double L, int nterms, const double *kper_c, double *u, int nthreads);

/* ---- RBF Gram matvec and CG (src/cuda/cg_rbf_init.cu:11-116) -------------- */
void orc_rbf_matvec(int n, const double *px, const double *py, double h_w, double L,
                    int periodic, const double *x, double *y, int nthreads);
int  orc_cg_rbf(int n, const double *px, const double *py, double h_w, double L,
                int periodic, const double *b, double *w, double tol, int max_iter,
                double *rel_res, int nthreads);

/* ---- analytic fields ------------------------------------------------------ */
double orc_tg_vorticity(double x, double y);                 /* analytical_flow.hpp:24-31 */
void   orc_tg_velocity(double x, double y, double *u, double *v);
double orc_shear_vorticity(double x, double y, double rho, double delta);

#ifdef __cplusplus
}
#endif
#endif
