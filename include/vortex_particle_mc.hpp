// vortex_particle_mc.hpp -- simulation parameters, particle buffers, views and
// the Simulation driver of the GPU path.
//
// Replaces include/vortex_particle_mc.hpp:16-201.  Names, fields and the
// SimParam construction semantics are the reference's; the implementation
// (src/cxx_api.cpp in the package's csrc/) calls the HIP kernels through the C
// ABI of vpmc.h.  The reference's CPU-backend entry points keep their
// semantics with the arithmetic on the GPU: step_cpu draws the CPU step's
// std::mt19937 stream (monte_carlo_bs_cpu per particle and sign, in its order),
// build_subsets_cpu builds the subsets (FP64 CDFs), evalVorticityAbs_cpu
// queries a view.
#pragma once

#include <cassert>
#include <cmath>
#include <cstdint>
#include <random>
#include <vector>

#include "device_vector_alias.hpp"
#include "spatial_hash.hpp"
#include "spatial_hasher.hpp"

struct ParticleSet {
    int N_max{0};
    int N_cur{0};

    dvec<float2> pos;
    dvec<float2> pos_left, pos_right, pos_top, pos_bottom;
    dvec<float2> pos_temp;
    dvec<float2> vel;

    dvec<float> omega;
    dvec<float> omega_field;
    dvec<float> jac;

    dvec<float4> F;

    void resize(int cap)
    {
        N_max = cap;
        N_cur = cap;
        pos.resize(cap);
        vel.resize(cap);
        pos_left.resize(cap);
        pos_right.resize(cap);
        pos_top.resize(cap);
        pos_bottom.resize(cap);
        pos_temp.resize(cap);
        omega.resize(cap);
        omega_field.resize(cap);
        jac.resize(cap);
        F.resize(cap);
    }
    void set_n_current(const int N_cur_) { N_cur = N_cur_; }

    // device pointers (callable from host and device code, as in the reference)
    HD float2 *d_p() { return raw_ptr(pos); }
    HD float2 *d_pt() { return raw_ptr(pos_temp); }
    HD float2 *d_v() { return raw_ptr(vel); }
    HD float4 *d_F() { return raw_ptr(F); }
    HD float *d_omega() { return raw_ptr(omega); }
    HD float *d_omega_field() { return raw_ptr(omega_field); }
    HD float *d_jac() { return raw_ptr(jac); }
    HD const float2 *d_p() const { return raw_ptr(pos); }
    HD const float2 *d_pt() const { return raw_ptr(pos_temp); }
    HD const float2 *d_v() const { return raw_ptr(vel); }
    HD const float4 *d_F() const { return raw_ptr(F); }
    HD const float *d_omega() const { return raw_ptr(omega); }
    HD const float *d_omega_field() const { return raw_ptr(omega_field); }
    HD const float *d_jac() const { return raw_ptr(jac); }
};

struct VorticityView {
    const float2 *pos;
    const float4 *F;
    const float *w;
    int N;

    const int *vp_sort;   // local id
    const int *cell_acc;
    int rshx{}, rshy{};
    float Lx, Ly;
    float InvLx, InvLy;
    int cnt_view;

    const int *sub;       // local id -> global id
};

struct MCBSCtx;

struct SimParam {
    float dt = 0.001f;
    float Lx = 1.f;
    float Ly = 1.f;
    int ResolutionX = 200;
    float hwr = 8.f;
    int N_max = 6000;
    int N1 = 1000;
    int N2 = 1000;

    int ResolutionY = 0;
    float h = 0.f;
    float h_w = 0.f;
    int length_iCDF = 10;

    float dx = 0.f;

    bool periodic = true;

    // the constructor's defaults win over the member initialisers
    explicit SimParam(float dt_ = 0.01f, float Lx_ = 1.f, float Ly_ = 1.f, int resX = 200,
                      float hwr_ = 16.f)
        : dt(dt_), Lx(Lx_), Ly(Ly_), ResolutionX(resX), hwr(hwr_)
    {
        assert(ResolutionX > 0 && "ResolutionX must be positive");
        h = Lx / float(ResolutionX);
        ResolutionY = int(std::round(Ly / h));
        if (ResolutionY == 0) ResolutionY = 1;
        h_w = h * hwr;
        dx = 0.5f * h_w;
    }
};

// Estimator selection of the GPU step (no reference counterpart: the reference
// has one estimator per backend).  Reference reproduces monte_carlo_bs_cpu's
// formulas; Consistent is the unbiased estimator of the periodic velocity.
enum class EstimatorMode : int { Reference = 0, Consistent = 1 };

class Simulation {
public:
    ParticleSet particles_;
    SimParam P_;

    explicit Simulation(SimParam p);
    ~Simulation();
    // copies own their device buffers (deep copy, as thrust::device_vector
    // members copy in the reference); the copy's views point at its own
    // buffers
    Simulation(const Simulation &o);
    Simulation &operator=(const Simulation &o);

    // src/vortex_particle_mc.cu:77-120: num_of_p uniform random positions
    // (std::mt19937(42)), strengths fitted by CG to the Taylor-Green field
    void init(int num_of_p);
    // positions and target vorticity given by the caller (e.g. a lattice);
    // strengths fitted to b (FP64 MINRES to relative residual tol)
    void init_with(const std::vector<float2> &pos, const std::vector<float> &b,
                   float tol = 1e-6f, int max_iter = 1000);
    // positions and strengths given directly
    void set_particles(const std::vector<float2> &pos, const std::vector<float> &omega);

    void do_spatial_hashing();
    void build_global_hash();

    void build_subsets_cpu();   // the same subsets as build_subsets_cuda (FP64 CDFs)
    void build_subsets_cuda();
    void prepare_rhs_b();

    std::mt19937 rng;
    // src/cpu/simulation_step.cpp:7-71: monte_carlo_bs_cpu for sign +1 then -1
    // per particle with the aux targets pos_left .. pos_bottom, on a copy of
    // rng (the member is not advanced, as in the reference), then x += u dt
    // for the particle and its four aux points, wrapped when periodic.  One
    // GPU launch replays the whole step's draws.
    void step_cpu(bool periodic);
    void step_cuda(bool periodic);

    void compute_velocity_naive(bool periodic, float eps_core);
    void step_cuda_naive(bool periodic);

    // include/vortex_particle_mc.hpp:160,204-208: w+(p) - w-(p) by the
    // reference-grid query of each sign's view
    inline float eval_vorticity_cpu(float2 p, bool periodic) const;

    void test_vorticity_grid(int res);

    void ensure_icdf_kernel(int L);

    VorticityView makePosView(const float *w_ptr = nullptr) const;
    VorticityView makeNegView(const float *w_ptr = nullptr) const;
    VorticityView makePosViewNaive(const float *w_ptr = nullptr) const;
    VorticityView makeNegViewNaive(const float *w_ptr = nullptr) const;
    VorticityView makeAllView(const float *w_ptr = nullptr) const;

    // ---- build extensions (no reference counterpart) ----
    EstimatorMode mode = EstimatorMode::Consistent;
    uint64_t seed = 0;        // Philox key
    uint32_t sim_id = 0;      // ensemble member (Philox counter)
    uint32_t step_index = 0;  // advanced by step_cuda
    int cells_per_hw = 8;     // fine grid of the estimator
    // step_cuda with the CUDA step's aux targets and kernel deformation
    // (particles.F: W(|F^-1 d|) in the query, F u in the jitter, F reset
    // from the five estimates every step); off: aux = 0, F = I
    bool kernel_deformation = false;
    float deformation_bound = 0.f;   // largest sigma_max(F) searched (0: 2, the 2 h_w cutoff)
    hipStream_t stream = nullptr;
    int cg_iterations = 0;
    double cg_rel_residual = 0.0;
    // reinitialisation (README.md:319-321, 416-420): particles to the
    // nx_l x ny_l cell-centred lattice (0: ResolutionX x ResolutionY),
    // strengths re-fitted to the current field (deformed when
    // kernel_deformation), F = I
    void reinitialize(int nx_l = 0, int ny_l = 0, float tol = 1e-6f, int max_iter = 1000);
    // velocity estimate at the current positions without advancing
    void estimate_velocity(bool periodic);
    // One step backwards (the adjoint of step_cuda under common random
    // numbers, vpmc_step_adjoint): the current positions are the state the
    // forward step `step` started from; lambda (device float2[N_cur]) holds
    // dL/dx after that step and becomes dL/dx before it; grad_w (device
    // float[N_cur]) accumulates dL/dw.  Consistent periodic estimator.  With
    // kernel_deformation (vpmc_step_adjoint_deformed) particles.F is the F
    // the step started from and jac_bar (device float4[N_cur]) holds dL/dF
    // after the step and becomes dL/dF before it.
    void step_adjoint(uint32_t step, float2 *lambda, float *grad_w, float4 *jac_bar = nullptr);
    // The FP64 forward of the same frozen-sample step (vpmc_step_f64): pos
    // (device double2[N_cur]) advances in place by the velocity at pos with
    // strengths w (device double[N_cur]); the draws, the sampling tables and
    // the MIS density are those of the current strengths; with
    // kernel_deformation, jac64 (device double4[N_cur]) is the FP64 F,
    // advanced in place (vpmc_step_f64_deformed).  No counterpart in the
    // reference: the accurate trajectory for step_adjoint's gradients.
    void step_f64(uint32_t step, double2 *pos, const double *w, double4 *jac64 = nullptr);
    int subset_count(int sign) const { return sign > 0 ? cnt_pos_ : cnt_neg_; }

    struct Impl;
    friend MCBSCtx make_mcbs_ctx(const Simulation &sim, bool periodic);

private:
    SpatialHash hash_pos_, hash_neg_, hash_all_;

    dvec<int> sub_all_;
    dvec<int> sub_pos_, sub_neg_;
    dvec<double> cdf_pos_, cdf_neg_;   // FP64 (the reference keeps FP32)
    dvec<float> cdf_pos_f32_, cdf_neg_f32_;   // the same CDFs rounded to FP32
    int cnt_pos_{}, cnt_neg_{};
    double cum_pos_{}, cum_neg_{};

    dvec<float> cdf_kernel, icdf_kernel;
    dvec<float> icdf_kernel_;
    int icdf_L_ = 0;

    VorticityView pos_view{};
    VorticityView neg_view{};

    Impl *impl_ = nullptr;
};

// src/cpu/evaluate_vorticity.cpp:8-75: the (signed or |w|) kernel sum at p over
// the view's 3x3 reference cells; sign filters w > 0 / w < 0 / none (0).  The
// view holds device pointers; the query runs on the GPU (vpmc_query_view).
float evalVorticityAbs_cpu(const VorticityView &v, const float2 &p, int sign, bool use_signed,
                           bool exclude_center, float h_w, bool periodic);

inline float Simulation::eval_vorticity_cpu(float2 p, bool periodic) const
{
    const float w_pos = evalVorticityAbs_cpu(pos_view, p, 1, false, false, P_.h_w, periodic);
    const float w_neg = evalVorticityAbs_cpu(neg_view, p, -1, false, false, P_.h_w, periodic);
    return w_pos - w_neg;
}
