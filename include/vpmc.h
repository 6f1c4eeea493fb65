/*
 * vpmc.h -- C ABI of the MI355X (gfx950) vortex-particle Monte Carlo hot path.
 *
 * Every pointer argument is a device pointer unless stated otherwise; sizes
 * are element counts.  `stream` is a hipStream_t passed as void* (NULL is the
 * null stream).  Functions return 0 (VPMC_OK) or a vpmc_status code and never
 * throw.  Nothing here allocates: functions that need scratch memory take a
 * caller buffer sized by the matching *_workspace_bytes query.  No function
 * synchronises the stream unless its comment says so.
 *
 * Each entry point names the reference interface it replaces (file:line of
 * GargantuaOmni/VortexParticleMC_cuda).  float2 data are interleaved xy float
 * pairs; "points" are float4 (x, y, w, 0) records.
 */
#ifndef VPMC_H
#define VPMC_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum vpmc_status {
    VPMC_OK = 0,
    VPMC_ERR_ARG = 1,           /* bad size, NULL pointer, bad mode           */
    VPMC_ERR_HIP = 2,           /* a HIP call failed (vpmc_last_hip_error)    */
    VPMC_ERR_WORKSPACE = 3,     /* workspace smaller than the query result    */
    VPMC_ERR_GRID = 4,          /* cell grid cannot cover the 2 h_w cutoff
                                   without self-overlap (the reference's
                                   rsh < 3 case, include/vorticity_query.cuh:36) */
    VPMC_ERR_CG_BREAKDOWN = 5,  /* p.Ap <= 0 in CG (indefinite Gram matrix)   */
    VPMC_ERR_NOT_CONVERGED = 6  /* CG hit max_iter                            */
};

const char *vpmc_status_string(int status);
int vpmc_last_hip_error(void);          /* hipError_t of the last VPMC_ERR_HIP */
int vpmc_version(void);                 /* major*10000 + minor*100 + patch     */
/* Measurement aid, no counterpart in the reference: blocks x 256 lanes each
   run iters x 16 dependent-chain FP32 FMAs (32 * iters FLOP per lane) on
   the stream, for reading the unpacked FP32 issue rate of the device the
   rooflines are quoted against; out (blocks floats) is written only in a
   case that cannot occur. */
int vpmc_probe_fma(int blocks, int iters, float *out, void *stream);

/* ---------------------------------------------------------------------------
 * Spatial hash on the reference's grid (rsh = ResolutionX / int(hwr) / 2,
 * src/vortex_particle_mc.cu:125).  Cell key ix*rshy+iy with
 * ix = clamp(floor(x * (rshx/Lx)), 0, rshx-1) in FP32 (NaN maps to 0).
 */
size_t vpmc_hash_workspace_bytes(int n, int cells);

/* replaces FillCells_cuda (src/cuda/hash.cu:42-69, include/spatial_hash.hpp:60-67):
   vp_cell[i] = key(pos[sub[i]]), cell_num = histogram, cell_acc = exclusive
   scan.  sub may be NULL (identity). */
int vpmc_fill_cells(int rshx, int rshy, float Lx, float Ly, int n, const int *sub,
                    const float *pos_xy, int *cell_num, int *cell_acc, int *vp_cell,
                    void *workspace, size_t workspace_bytes, void *stream);

/* replaces CountingSort_cuda (src/cuda/hash.cu:71-91, include/spatial_hash.hpp:69-73).
   Stable: vp_map lists, per cell, local indices in ascending order -- the
   order of CountingSort_cpu (src/cpu/hash.cpp:36-47); cell_num ends at 0. */
int vpmc_counting_sort(int n, int cells, int *cell_num, const int *cell_acc,
                       const int *vp_cell, int *vp_map, void *workspace,
                       size_t workspace_bytes, void *stream);

/* ---------------------------------------------------------------------------
 * Sign subsets and the |w| CDF in FP64.
 * replaces Simulation::build_subsets_cuda (src/cuda/build_particle_sets.cu:11-89):
 * sub_pos/sub_neg list indices with w > 0 / w < 0 ascending; cdf_* is the
 * normalised inclusive prefix of |w| over the subset (FP64), counts[0..1] and
 * totals[0..1] (device) receive the subset sizes and sum |w|.
 */
size_t vpmc_subsets_workspace_bytes(int n);
int vpmc_build_subsets(int n, const float *omega, int *sub_pos, int *sub_neg,
                       double *cdf_pos, double *cdf_neg, int *counts, double *totals,
                       void *workspace, size_t workspace_bytes, void *stream);

/* Guide table for CDF inversion (no reference counterpart): guide[k] =
   upper_bound(cdf, count, k / 2^m_log2) for k in [0, 2^m_log2].  With it the
   sampler returns exactly std::upper_bound's index (sample_vorticity_idx_cdf,
   src/cpu/monte_carlo_sampling.cpp:23-33) after a short in-bucket search.
   count_dev is a device int (e.g. counts from vpmc_build_subsets). */
int vpmc_build_guide(const double *cdf, const int *count_dev, int m_log2, int *guide,
                     void *stream);

/* replaces sample_vorticity_idx_cdf (src/cpu/monte_carlo_sampling.cpp:23-33,
   include/monte_carlo_bs.cuh:105-115) for a batch of draws u in [0,1):
   idx_out[i] = upper_bound(cdf, u[i]) clamped to count-1 (-1 if count is 0). */
int vpmc_sample_idx(int nu, const float *u, const double *cdf, const int *count_dev,
                    const int *guide, int m_log2, int *idx_out, void *stream);

/* ---------------------------------------------------------------------------
 * Vorticity query on the reference's hashed view.
 */
typedef struct vpmc_view {            /* include/vortex_particle_mc.hpp:77-94 */
    const float *pos_xy;              /* global positions (float2)             */
    const float *w;                   /* global strengths                      */
    int          N;
    const int   *vp_sort;             /* local indices grouped by cell         */
    const int   *cell_acc;            /* rshx*rshy exclusive offsets           */
    int          rshx, rshy;
    float        Lx, Ly;
    int          cnt_view;            /* end of the last cell                  */
    const int   *sub;                 /* local -> global                       */
} vpmc_view;

/* replaces queryVorticityAbs (include/vorticity_query.cuh:15-79) with the
   semantics of evalVorticityAbs_cpu (src/cpu/evaluate_vorticity.cpp:8-75):
   out[i] = sum over the 3x3 cells of (signed or |w|) W(|d|), d folded across
   the periodic box, cut at h_w; sign filters w > 0 / w < 0 / none (0). */
int vpmc_query_view(const vpmc_view *view, int npts, const float *pts_xy, int sign,
                    int use_signed, int exclude_center, float h_w, int periodic, float *out,
                    void *stream);

/* ---------------------------------------------------------------------------
 * Fine cell grid ("field") used by the estimator and the RBF matvec: the
 * particles of one set, sorted by cell (stable, ascending local index), as
 * float4 (x, y, weight, 0) records with cell_start offsets.  The grid is
 * chosen so that cells are <= h_w / cells_per_hw wide; queries visit, per
 * cell column, only the rows within h_w of the point.
 */
typedef struct vpmc_field {
    int          nx, ny;
    float        Lx, Ly;
    int          n;
    const float *pts;                 /* n float4 records, cell order          */
    const int   *cell_start;          /* nx*ny + 1                             */
    const int   *order;               /* sorted position -> local index        */
    /* deformed kernels (F in the query, include/vorticity_query.cuh:52-66):
       gq (optional) holds per record the quadratic form G = F^-T F^-1 of
       its particle's deformation F as float4 (G11, 2 G12, G22, sigma_max(F)^2), so that
       the kernel is W(sqrt(d^T G d)) = W(|F^-1 d|); queries then search
       gq_bound * h_w (0: 2, the reference's 2 h_w cutoff), the largest
       stretch sigma_max(F) the records may carry. */
    const float *gq;
    float        gq_bound;
} vpmc_field;

int vpmc_field_grid(float h_w, float Lx, float Ly, int cells_per_hw, int *nx, int *ny);
size_t vpmc_field_workspace_bytes(int n, int nx, int ny);
/* weight = w[sub[j]] (abs_w: |w|); sub NULL = identity; set_xy (optional)
   receives pos[sub[j]] in local order. */
int vpmc_field_build(int nx, int ny, float Lx, float Ly, int n, const int *sub,
                     const float *pos_xy, const float *w, int abs_w, float *pts,
                     int *cell_start, int *order, float *set_xy, void *workspace,
                     size_t workspace_bytes, void *stream);
/* vpmc_field_build with deformations: jac (n float4 F per particle, global
   index) gives gq[k] = G of record k (field order) and set_jac[j] = F of
   local index j (either output may be NULL). */
int vpmc_field_build_jac(int nx, int ny, float Lx, float Ly, int n, const int *sub,
                         const float *pos_xy, const float *w, int abs_w, const float *jac,
                         float *pts, float *gq, int *cell_start, int *order, float *set_xy,
                         float *set_jac, void *workspace, size_t workspace_bytes, void *stream);
/* periodic query of the field's weighted kernel sum at npts points */
int vpmc_field_query(const vpmc_field *f, float h_w, int npts, const float *pts_xy,
                     float *out, void *stream);

/* ---------------------------------------------------------------------------
 * Monte Carlo Biot-Savart estimator.
 *   mode VPMC_MODE_REFERENCE: the formulas of monte_carlo_bs_cpu
 *     (src/cpu/monte_carlo_bs.cpp:11-128, aux = 0): CDF index + periodic
 *     image shift, area-uniform disk of radius sqrt(Lx^2+Ly^2), mis_contrib
 *     with cum = 9 * Omega.  Reproduces the reference on injected uniforms.
 *   mode VPMC_MODE_CONSISTENT: the same two samplers with the densities
 *     matched (blob jitter drawn from W, disk radius uniform so p2 = alpha/r)
 *     on the min-image cell, with the doubly periodic kernel; unbiased for
 *     the blob field's periodic velocity.
 * Uniform source: Philox4x32-10 keyed by (seed; k, owner, step,
 * sim/sign/branch) or, when `uniforms` is non-NULL, injected draws in the
 * order monte_carlo_bs_cpu consumes them.
 */
enum { VPMC_MODE_REFERENCE = 0, VPMC_MODE_CONSISTENT = 1 };

typedef struct vpmc_sign_set {
    int           n;                  /* particles in the set                  */
    const int    *count_dev;          /* device copy of n (NULL: use n)        */
    const double *cdf;                /* FP64 normalised prefix of |w|         */
    const double *total_dev;          /* device sum |w| (FP64)                 */
    const int    *guide;              /* vpmc_build_guide table                */
    int           guide_log2;
    const float  *set_xy;             /* positions of the set in local order   */
    vpmc_field    field;              /* the set on the fine grid, weight |w|  */
    const float  *set_jac;            /* optional: F of the set in local order
                                         (float4); branch-1 jitter F_j u
                                         (src/cuda/monte_carlo_bs.cu:210-212)  */
} vpmc_sign_set;

typedef struct vpmc_mc_desc {
    int          mode;
    int          N1, N2;
    float        h_w, Lx, Ly;
    int          periodic;
    uint64_t     seed;
    uint32_t     step, sim;
    const float *uniforms;            /* replay stream or NULL                 */
    int          aux_draws;           /* replay stride: 10 draws per branch-2
                                         sample (main disk first)              */
    int          n_owner;
    const float *owner_xy;            /* destinations (float2)                 */
    vpmc_sign_set set[2];             /* 0: w > 0, 1: w < 0                    */
    float       *vel_sign;            /* out: 2*n_owner float2, sign-major     */
    unsigned long long *eval_count;   /* optional: += kernel evaluations within
                                         h_w and candidates visited ([0],[1])  */
    int          algo;                /* VPMC_ALGO_BINNED (default) or _DIRECT */
    void        *eval_events[2];      /* optional hipEvent_t pair, recorded on the
                                         stream right before and after the
                                         evaluation kernel (k_bin_eval, or
                                         k_mc_estimate for the direct algorithm);
                                         NULL: not recorded                    */
    int          spl;                 /* binned evaluation: samples per lane (1,
                                         2 or 4; 0 means 2).  A tuning knob only:
                                         every sample sums its candidates in the
                                         same order, so results are identical */
    int          eval_threads;        /* binned evaluation workgroup size: 64,
                                         128 or 256 (0: 256), a tuning knob with
                                         identical results; the stencil's
                                         columns must fit (VPMC_ERR_GRID)     */
    /* auxiliary targets (src/cuda/monte_carlo_bs.cu:166-169, 213-259;
       src/cpu/monte_carlo_bs.cpp:60-119): with aux != 0 the binned estimator
       also estimates at dst -/+ aux_dx e_x (L, R) and dst +/- aux_dx e_y (T, B).
       Branch-1 samples are shared by the five targets (same point and
       vorticity); each aux target has N2 disk samples of its own (Philox
       branch ids 3..6; in a replay stream, draws 2..9 of each branch-2
       sample's ten, aux_draws = 1).  vel_aux: 8 * n_owner float2, index
       (sign * 4 + target) * n_owner + owner, targets L, R, T, B. */
    int          aux;
    float        aux_dx;
    float       *vel_aux;
    /* vpmc_step only: per-particle deformation F (n float4, row-major
       F.x F.y; F.z F.w) or NULL.  The step builds the fields with it (the
       sets' set_jac and field.gq buffers must be given) and, with aux,
       replaces it by update_jacobian_from5 of this step's estimates. */
    float       *jac;
    /* Owner ranges (no reference counterpart; the reference launches one
       block per particle, src/cuda/simulation_step.cu:13,51-58, and so has
       no sample buffer to bound).  owner_base is the global index of owner
       0: the Philox counter and the replay stream use owner_base + i for
       owner i, so an estimate over owners [b, b + n) with owner_base = b
       equals rows b .. b+n-1 of the estimate over all owners.  The binned
       algorithm runs its owners in chunks of at most chunk_samples sample
       slots (0: 2^30; at most 2^31 - 1), so its workspace is bounded by the
       chunk, not by n_owner; results do not depend on the chunk size. */
    int          owner_base;
    uint64_t     chunk_samples;
} vpmc_mc_desc;

enum { VPMC_ALGO_BINNED = 0, VPMC_ALGO_DIRECT = 1 };

/* replaces monte_carlo_bs_kernel_safe (src/cuda/monte_carlo_bs.cu:144-271)
   for both signs: vel_sign[s*n_owner + i] = estimate for owner i, sign s.
   Direct algorithm: one workgroup per (owner, sign), no workspace. */
int vpmc_mc_estimate(const vpmc_mc_desc *d, void *stream);

/* Same estimate, binned algorithm (periodic boxes): samples are generated,
   counting-sorted by fine cell and evaluated cell by cell, then reduced per
   owner in sample order, one owner chunk after the other (chunk_samples).
   Needs vpmc_mc_binned_workspace_bytes(d) bytes (about 48 bytes per sample
   slot of one chunk).  Results equal the direct algorithm's up to FP32
   summation order.  With several chunks, eval_events bracket the
   evaluations of all chunks (first evaluation start to last evaluation
   end). */
size_t vpmc_mc_binned_workspace_bytes(const vpmc_mc_desc *d);
int vpmc_mc_estimate_binned(const vpmc_mc_desc *d, void *workspace, size_t workspace_bytes,
                            void *stream);

/* replaces update_jacobian_from5 + renorm_det1 (src/cuda/monte_carlo_bs.cu:274-288,
   include/monte_carlo_bs.cuh:140-144; src/cuda/simulation_step.cu:69-73):
   from the centre velocities vel_sign and the aux targets' vel_aux of one
   estimate (u = v+ - v- per target), F = I + (dt / 2dx) [uR - uL | uT - uB]
   rescaled to det F = 1; jac_out: n float4 (F.x F.y; F.z F.w) row-major. */
int vpmc_update_jacobian(int n, const float *vel_sign, const float *vel_aux, float dt, float dx,
                         float *jac_out, void *stream);

/* replaces the combine + Euler + wrap of step_cuda
   (src/cuda/simulation_step.cu:63-67,77-89) and step_cpu
   (src/cpu/simulation_step.cpp:41-71): vel = v+ - v-, pos += vel*dt, wrap. */
int vpmc_advect(int n, float *pos_xy, const float *vel_sign, float *vel_xy, float dt,
                float Lx, float Ly, int periodic, void *stream);

/* ---------------------------------------------------------------------------
 * One plain time step (replaces Simulation::step_cuda,
 * src/cuda/simulation_step.cu:9-90, with aux = 0): rebuild both sign fields
 * from the current positions, estimate, advect.  Subset, CDF and guide arrays
 * are inputs (w is constant in the inviscid step).  Field buffers are
 * caller-owned in `d->set[s].field` (cast away const) and `set_xy`.  Every
 * particle is an owner (n_owner == n_total, owner_xy == pos_xy): the step
 * moves all of them; the binned estimator runs them in owner chunks of
 * d->chunk_samples sample slots, so any N with aux targets fits a bounded
 * workspace.
 */
size_t vpmc_step_workspace_bytes(const vpmc_mc_desc *d);
int vpmc_step(vpmc_mc_desc *d, int n_total, float *pos_xy, const float *omega,
              const int *sub_pos, const int *sub_neg, float *vel_xy, float dt,
              void *workspace, size_t workspace_bytes, void *stream);

/* No counterpart in the reference: vpmc_step for callers that cannot
   allocate (code under HIP graph capture, serving loops that forbid
   hipMalloc).  It takes no workspace and allocates nothing: its scratch is
   a static device array of VPMC_NOALLOC_SCRATCH_BYTES per device, held by
   every GPU the library runs on until the process ends, and the binned
   estimator's owner chunks are sized to fit it (at most d->chunk_samples
   slots), which leaves every result bit-identical to vpmc_step's.  Returns
   VPMC_ERR_WORKSPACE when a field build or one owner's samples need more than
   the array.  Calls that overlap on one device race on the array. */
#define VPMC_NOALLOC_SCRATCH_BYTES (256u << 20)
int vpmc_step_noalloc(vpmc_mc_desc *d, int n_total, float *pos_xy, const float *omega,
                      const int *sub_pos, const int *sub_neg, float *vel_xy, float dt,
                      void *stream);

/* ---------------------------------------------------------------------------
 * Ensemble of independent periodic simulations (no reference counterpart: the
 * reference runs one Simulation per process).  All simulations advance in
 * one set of launches per step: one field sort, one sample generation, one
 * sample sort, one evaluation, one reduction and one advection for the whole
 * ensemble.  Simulation s is the particle range [h_offset[s], h_offset[s+1])
 * of pos_xy / omega, with its own h_w and box side (Lx = Ly = h_L[s]).  Its
 * trajectory equals vpmc_step's for that simulation alone with sim = s, the
 * same seed, cells_per_hw and spl, bit for bit.  The workspace holds the
 * ensemble's state (subsets, CDFs, per-simulation tables, the step counter):
 * keep it and the descriptor unchanged between init and the steps.
 */
typedef struct vpmc_ensemble_desc {
    int          n_sim;
    const int   *h_offset;            /* host: n_sim + 1 particle offsets      */
    const float *h_h_w;               /* host: per-simulation h_w               */
    const float *h_L;                 /* host: per-simulation box side          */
    int          N1, N2;
    int          mode;                /* VPMC_MODE_*                            */
    int          cells_per_hw;        /* fine grid (vpmc_field_grid)            */
    int          spl;                 /* samples per lane of the evaluation     */
    int          eval_threads;        /* evaluation workgroup size (64/128/256) */
    uint64_t     seed;
    float        dt;
    float       *pos_xy;              /* device: all particles (float2)         */
    const float *omega;               /* device: strengths                      */
    float       *vel_xy;              /* device out (float2), or NULL           */
    void        *eval_events[2];      /* optional hipEvent_t pair, recorded on the
                                         stream right before and after the
                                         evaluation kernel (k_ens_eval)         */
    /* device: per-particle deformation F (float4, row-major), or NULL.  With
       it every step is the CUDA step's (vpmc_step with jac): fields carry the
       quadratic forms of F, branch-1 jitter is F u, each owner is also
       estimated at its four aux targets (dx = h_w / 2), and F is replaced by
       update_jacobian_from5 of the five estimates before the move. */
    float       *jac;
} vpmc_ensemble_desc;

size_t vpmc_ensemble_workspace_bytes(const vpmc_ensemble_desc *e);
/* subsets, FP64 CDFs, guides and tables; step counter 0.  Synchronises. */
int vpmc_ensemble_init(const vpmc_ensemble_desc *e, void *workspace, size_t workspace_bytes,
                       void *stream);
/* one step of every simulation; no host synchronisation, capturable */
int vpmc_ensemble_step(const vpmc_ensemble_desc *e, void *workspace, size_t workspace_bytes,
                       void *stream);
/* sets the step counter (the Philox step of the next step) */
int vpmc_ensemble_set_step(const vpmc_ensemble_desc *e, void *workspace, size_t workspace_bytes,
                           uint32_t step, void *stream);
/* captures `steps` consecutive vpmc_ensemble_step calls on `stream` (not the
   null stream) into an executable HIP graph; replaying it advances the
   ensemble by `steps` steps (the step counter lives in the workspace). */
int vpmc_ensemble_graph_create(const vpmc_ensemble_desc *e, void *workspace,
                               size_t workspace_bytes, int steps, void *stream,
                               void **graph_exec);
int vpmc_graph_launch(void *graph_exec, void *stream);
int vpmc_graph_destroy(void *graph_exec);

/* ---------------------------------------------------------------------------
 * Adjoint (reverse-mode derivative) of the consistent estimator and of the
 * advection step under common random numbers (no reference counterpart: the
 * reference has no derivatives; the forward is monte_carlo_bs_kernel_safe,
 * src/cuda/monte_carlo_bs.cu:144-271, and step_cuda,
 * src/cuda/simulation_step.cu:9-90).  Samples are frozen: the Philox draws,
 * the particle each branch-1 sample was drawn from and the jitter / disk
 * offsets stay fixed, so sample points move with their source particle
 * (branch 1) or owner (branch 2).  The MIS denominator is the density the
 * samples were drawn from: it follows the positions but not the strengths
 * (the frozen indices do not follow them), so the estimate and its
 * derivative stay unbiased for the exact velocity and its derivative.
 * Periodic, consistent mode, Philox draws only; owners are the particles
 * (n_owner = particle count, owner_xy = their positions) and the fields of
 * `d` are built from those positions.  Gradients are accumulated (+=).
 */
size_t vpmc_mc_adjoint_workspace_bytes(const vpmc_mc_desc *d);
/* adj_vel: n_owner float2, dL/du_i of u = v+ - v- (vel_sign[0] - vel_sign[1]);
   grad_pos (float2) and grad_w receive dL/dx and dL/dw of every particle. */
int vpmc_mc_adjoint(const vpmc_mc_desc *d, const int *sub_pos, const int *sub_neg,
                    const float *adj_vel, float *grad_pos, float *grad_w, void *workspace,
                    size_t workspace_bytes, void *stream);
/* One step backwards through x' = wrap(x + dt u(x, w)) at the step index in
   d->step: rebuilds the fields from pos_xy (the state the forward step
   started from), then lambda (dL/dx', float2) becomes dL/dx and grad_w
   accumulates dL/dw of the step. */
size_t vpmc_step_adjoint_workspace_bytes(const vpmc_mc_desc *d);
int vpmc_step_adjoint(vpmc_mc_desc *d, int n_total, const float *pos_xy, const float *omega,
                      const int *sub_pos, const int *sub_neg, float dt, float *lambda,
                      float *grad_w, void *workspace, size_t workspace_bytes, void *stream);
/* FP64 forward of the same frozen-sample step (checks by finite differences):
   positions pos_xy (double2, advanced in place) and strengths w (double) at
   which the velocity is evaluated; the sampling tables of `d` (subsets, CDF,
   guide, totals) and the MIS density come from the nominal strengths omega0. */
size_t vpmc_step_f64_workspace_bytes(const vpmc_mc_desc *d);
int vpmc_step_f64(vpmc_mc_desc *d, int n_total, double *pos_xy, const double *w,
                  const float *omega0, const int *sub_pos, const int *sub_neg, double dt,
                  void *workspace, size_t workspace_bytes, void *stream);
/* The deformed step (step_cuda with kernel deformation: the four auxiliary
   targets and F, src/cuda/simulation_step.cu:51-73, src/cuda/monte_carlo_bs.cu:
   166-169, 210-212, 274-288; include/vorticity_query.cuh:52-66) backwards:
   d->jac holds the F the forward step started from (read only), d->aux = 1
   with d->vel_aux, and the sets carry field.gq / set_jac buffers (rebuilt
   here).  The derivative includes F's: a particle's blob is W(|F^-1 d|) in
   every query and its branch-1 jitter is F u (u frozen), and the step's
   F' = renorm(I + dt/(2 dx) [u_R - u_L | u_T - u_B]) follows the four aux
   estimates.  lambda: dL/dx' -> dL/dx; jac_bar (n float4): dL/dF' -> dL/dF;
   grad_w accumulates dL/dw.  Runs the step's forward estimate first (into
   d->vel_sign, d->vel_aux). */
size_t vpmc_step_adjoint_deformed_workspace_bytes(const vpmc_mc_desc *d);
int vpmc_step_adjoint_deformed(vpmc_mc_desc *d, int n_total, const float *pos_xy,
                               const float *omega, const int *sub_pos, const int *sub_neg,
                               float dt, float *lambda, float *jac_bar, float *grad_w,
                               void *workspace, size_t workspace_bytes, void *stream);
/* FP64 forward of the deformed frozen-sample step (checks): jac64 (n double4,
   F per particle) is replaced by this step's F' before the move; d->aux = 1
   and the sets' field.gq pointers set (they give the fields' 2 h_w search
   radius; their contents are not read).  Workspace: vpmc_step_f64_workspace_bytes. */
int vpmc_step_f64_deformed(vpmc_mc_desc *d, int n_total, double *pos_xy, const double *w,
                           const float *omega0, const int *sub_pos, const int *sub_neg,
                           double dt, double *jac64, void *workspace, size_t workspace_bytes,
                           void *stream);

/* ---------------------------------------------------------------------------
 * Vorticity sampled by the B-spline kernel on a grid (no single reference
 * entry point: the host loop of dump_vorticity_png, include/viz_utils.hpp:26-82,
 * and the grid of test_vorticity_grid, src/vortex_particle_mc.cu:332-456):
 * out[a*gny + b] = sum_j w_j W(|z_ab - x_j|), z_ab = ((a+1/2) Lx/gnx,
 * (b+1/2) Ly/gny), signed strengths, periodic.
 */
size_t vpmc_grid_vorticity_workspace_bytes(int n, int nx, int ny);
int vpmc_grid_vorticity(int n, const float *pos_xy, const float *w, float h_w, float Lx,
                        float Ly, int cells_per_hw, int gnx, int gny, float *out, void *workspace,
                        size_t workspace_bytes, void *stream);
/* adjoint: grad_pos / grad_w += the gradient of sum_g grad_out[g] out[g] */
int vpmc_grid_vorticity_adjoint(int n, const float *pos_xy, const float *w, float h_w, float Lx,
                                float Ly, int gnx, int gny, const float *grad_out,
                                float *grad_pos, float *grad_w, void *stream);
/* FP64, all particles per grid point (checks) */
int vpmc_grid_vorticity_f64(int n, const double *pos_xy, const double *w, double h_w, double Lx,
                            double Ly, int gnx, int gny, double *out, void *stream);

/* ---------------------------------------------------------------------------
 * RBF strength initialisation.
 * replaces Ax_cuda / matvec_kernel (src/cuda/cg_rbf_init.cu:11-50):
 *   y_i = sum_j x_j W(|x_i - x_j|) over all particles (diagonal included).
 * Order "cell" means x and y are in the field's sorted order.
 */
size_t vpmc_rbf_workspace_bytes(int n, int nx, int ny);
int vpmc_rbf_matvec(const vpmc_field *f_all, float h_w, const float *x_sorted,
                    float *y_sorted, void *stream);
int vpmc_rbf_matvec_f64(const vpmc_field *f_all, float h_w, const double *x_sorted,
                        double *y_sorted, void *stream);
/* the same operator in particle order (Ax_cuda's contract): builds the
   all-particle field from pos_xy in the workspace (vpmc_cg_workspace_bytes)
   and applies it to x. */
int vpmc_rbf_matvec_particles(int n, const float *pos_xy, const float *x, float *y, float h_w,
                              float Lx, float Ly, int cells_per_hw, void *workspace,
                              size_t workspace_bytes, void *stream);

/* replaces conjugate_gradient_rbf (src/cuda/cg_rbf_init.cu:53-116): solves
   A w = b.  refine == 0: the reference's FP32 CG from w0 = b / W(0) (FP32
   matvec, FP64 dot products), VPMC_ERR_CG_BREAKDOWN when p.Ap <= 0 (the Gram
   matrix is indefinite).  refine != 0: MINRES (symmetric indefinite) on FP64
   vectors with the FP64 matvec from w0 = 0, until ||b - A w|| / ||b|| < tol,
   else VPMC_ERR_NOT_CONVERGED after max_iter.  The strengths are returned in
   FP32, so tol below ~1e-7 amplifies the rounding noise of b in the near-null
   modes of A.  Synchronises the stream each iteration (host loop).
   iters/rel_res (host pointers) report the iterations and the final FP64
   relative residual of the FP64 iterate. */
int vpmc_cg_rbf(int n, const float *pos_xy, const float *b, float *w_out, float h_w,
                float Lx, float Ly, int cells_per_hw, double tol, int max_iter, int refine,
                int *iters, double *rel_res, void *workspace, size_t workspace_bytes,
                void *stream);
size_t vpmc_cg_workspace_bytes(int n, int nx, int ny);
/* The same solve (MINRES on FP64 vectors, as refine != 0) for an FP64
   right-hand side b, e.g. a vorticity field evaluated in FP64: an FP32 b
   carries rounding noise into every lattice mode, which bounds the reachable
   residual near 1e-8; an FP64 b of a smooth field does not.  w_out receives
   the strengths in FP32, w64_out (optional) the FP64 iterate whose relative
   residual *rel_res reports. */
int vpmc_cg_rbf_f64(int n, const float *pos_xy, const double *b, float *w_out, double *w64_out,
                    float h_w, float Lx, float Ly, int cells_per_hw, double tol, int max_iter,
                    int *iters, double *rel_res, void *workspace, size_t workspace_bytes,
                    void *stream);

/* vpmc_cg_rbf_f64 for particles on the cell-centred nx_l x ny_l lattice
   (particle i*ny_l + j near ((i+1/2) Lx/nx_l, (j+1/2) Ly/ny_l); the matvec
   uses pos_xy as given): MINRES preconditioned by the lattice's circulant
   Gram matrix, |lambda(k)| in the lattice Fourier basis (floored at
   1e-10 lambda(0)), applied by direct DFTs along lattice lines (nx_l, ny_l
   <= 2048).  Replaces conjugate_gradient_rbf (src/cuda/cg_rbf_init.cu:53-116)
   where the reference initialises on its lattice; reaches 1e-10 relative
   residuals that the unpreconditioned solve does not. */
size_t vpmc_cg_lattice_workspace_bytes(int nx_l, int ny_l, float h_w, float Lx, float Ly,
                                       int cells_per_hw);
int vpmc_cg_rbf_lattice(int nx_l, int ny_l, const float *pos_xy, const double *b, float *w_out,
                        double *w64_out, float h_w, float Lx, float Ly, int cells_per_hw,
                        double tol, int max_iter, int *iters, double *rel_res, void *workspace,
                        size_t workspace_bytes, void *stream);

/* Reinitialisation, described in the reference's README.md:319-321 and
   :416-420 (step 16: positions to a uniform grid, Jacobians to I, strengths
   re-fitted by the RBF solve) with no code behind it.  The current field
   sum_j w_j W(|F_j^-1 (x - x_j)|) (jac NULL: F = I; signed strengths, FP64
   sums, deformed search radius 2 h_w) is evaluated at the nx_l x ny_l
   cell-centred lattice ((i+1/2) Lx/nx_l, (j+1/2) Ly/ny_l), index i*ny_l + j,
   written to pos_out (2 nx_l ny_l floats); w_out receives strengths on that
   lattice from the FP64 MINRES solve of vpmc_cg_rbf_f64 to relative
   residual tol.  The caller resets its F to I. */
size_t vpmc_reinit_workspace_bytes(int n, int nx_l, int ny_l, float h_w, float Lx, float Ly,
                                   int cells_per_hw, int deformed);
int vpmc_reinit(int n, const float *pos_xy, const float *w, const float *jac, float h_w,
                float Lx, float Ly, int cells_per_hw, int nx_l, int ny_l, double tol,
                int max_iter, float *pos_out, float *w_out, int *iters, double *rel_res,
                void *workspace, size_t workspace_bytes, void *stream);

/* ---------------------------------------------------------------------------
 * FP64 direct sums (error oracles on the device).
 * replaces biot_savart_naive_kernel (src/cuda/naive_biot_savart.cu:15-40) in
 * FP64: u(q) = sum_j g_j perp(d)/(2 pi (|d|^2+eps^2)), d = minimg(q - x_j);
 * skip_self skips j == probe index. */
int vpmc_direct_sum_eps(int nprobe, const double *probe_xy, int skip_self, int n,
                        const double *pos_xy, const double *g, double L, int periodic,
                        double eps, double *u_out, void *stream);
/* no reference counterpart: exact periodic velocity of the blob field
   sum_j g_j W(|y - x_j|) (consistent-mode expectation). */
int vpmc_direct_sum_blob(int nprobe, const double *probe_xy, int n, const double *pos_xy,
                         const double *g, double h_w, double L, double *u_out, void *stream);

#ifdef __cplusplus
}
#endif
#endif
