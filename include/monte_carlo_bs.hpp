// monte_carlo_bs.hpp -- Monte Carlo Biot-Savart estimator interface.
// Replaces include/monte_carlo_bs.hpp:7-102.  MCBSEstimate, MCBSCtx,
// monte_carlo_bs_cpu, the std::mt19937 samplers and the inline kernel helpers
// keep the reference's names, signatures and meaning; MCBSCtx points at
// device buffers.  monte_carlo_bs_cpu consumes the caller's generator exactly
// as the reference does and evaluates the estimate on the GPU (reference
// mode, replaying those draws); monte_carlo_bs_gpu evaluates the estimator for
// every particle of a Simulation at once on Philox streams.
#pragma once

#include <cmath>
#include <random>

#include "device_vector_alias.hpp"
#include "utilities.hpp"
#include "vortex_particle_mc.hpp"

struct MCBSEstimate {
    float2 main{};
    float2 left{};
    float2 right{};
    float2 top{};
    float2 bottom{};
};

struct MCBSCtx {
    // the reference's members and types (include/monte_carlo_bs.hpp:15-35)
    const dvec<int> *sub_pos = nullptr;
    const dvec<int> *sub_neg = nullptr;
    const dvec<float> *cdf_pos = nullptr;   // normalised CDF (FP32)
    const dvec<float> *cdf_neg = nullptr;
    const dvec<float> *icdf_kernel = nullptr;
    int kernel_L = 0;
    float h_w = 1.0f;
    bool periodic = false;
    float Lx = 1.0f;
    float Ly = 1.0f;
    float InvLx = 1.0f;
    float InvLy = 1.0f;
    float cum_pos = 1.0f;   // sum |w| per sign
    float cum_neg = 1.0f;
    VorticityView pos_view;
    VorticityView neg_view;
    // Added members (no reference counterpart).  FP64 CDF and total of a
    // sign: when cdf_pos64 / cdf_neg64 is set it is used, with cum_pos64 /
    // cum_neg64, in place of the FP32 pair above.  An FP32 CDF is widened to
    // FP64 exactly, so a caller that fills only the reference's members gets
    // the reference's particle choices.
    const dvec<double> *cdf_pos64 = nullptr;
    const dvec<double> *cdf_neg64 = nullptr;
    double cum_pos64 = 1.0;
    double cum_neg64 = 1.0;
    // subset sizes (the reference's CDF vectors are exactly that long);
    // -1: the size of the CDF in use
    int cnt_pos = -1;
    int cnt_neg = -1;
};

// src/cpu/monte_carlo_bs.cpp:28-128: the estimate at dst for one sign (+1 / -1)
// and, with auxiliary != 0, at the four given targets, from N1 branch-1 and N2
// branch-2 samples per target.  rng advances by exactly the draws the
// reference takes, in its order: per branch-1 sample the CDF draw and, when
// periodic, two image draws; per branch-2 sample two draws for the disk at dst
// and, with auxiliary, two for each of left, right, top, bottom.  The sums run
// on the GPU (VPMC_MODE_REFERENCE replaying those draws; ctx.pos_view and
// neg_view are not read).  An empty sign set gives zero (the reference reads
// sub[-1]).  Blocking: returns after the device work.
MCBSEstimate monte_carlo_bs_cpu(const float2 &dst, int sign, int auxiliary,
                                const float2 &dst_left, const float2 &dst_right,
                                const float2 &dst_top, const float2 &dst_bottom,
                                const ParticleSet &P, const MCBSCtx &ctx, int N1, int N2,
                                std::mt19937 &rng);

// src/cpu/monte_carlo_sampling.cpp:9-45, host utilities over host arrays
// (the reference's CPU samplers; no GPU path calls them):
// kernel radius by linear interpolation of the L-entry inverse CDF, times h
float sample_kernel_radius(std::mt19937 &gen, const float *icdf, int L, float h);
// upper_bound(cdf, cdf + N, u) for one draw u, N - 1 when past the end
int sample_vorticity_idx_cdf(std::mt19937 &gen, const float *cdf, int N);
// area-uniform point of the disk (center, R): angle draw first, then radius
float2 sample_disk_biot_savart(std::mt19937 &gen, const float2 &center, float R);

// the context src/cpu/simulation_step.cpp:13-27 builds from a Simulation
// (subsets, the FP32 CDFs and totals of the reference's members and the FP64
// ones of the added members, views; counts set), for monte_carlo_bs_cpu
MCBSCtx make_mcbs_ctx(const Simulation &sim, bool periodic);

// K(x, y) = (-(x-y)_y, (x-y)_x) / (2 pi (|x-y|^2 + 1e-8))   (:60-66)
HD inline float2 biot_savart_kernel(const float2 &x, const float2 &y)
{
    float2 r{x.x - y.x, x.y - y.y};
    float inv_r2 = 1.f / (r.x * r.x + r.y * r.y + 1e-8f);
    const float two_pi = 2.f * 3.14159265358979323846f;
    return make_float2(-r.y * inv_r2 / two_pi, r.x * inv_r2 / two_pi);
}

static inline void accumulate(float2 &acc, const float2 &add)
{
    acc.x += add.x;
    acc.y += add.y;
}

// largest distance from dst to its 9 periodic images: sqrt(Lx^2 + Ly^2)  (:73-89)
HD inline float SamplingDiskRadius(const float2 &dst, bool periodic, const float &Lx,
                                   const float &Ly)
{
    if (!periodic) return sqrtf(Lx * Lx + Ly * Ly);
    float maxR2 = 0.f;
    for (int dx = -1; dx <= 1; ++dx)
        for (int dy = -1; dy <= 1; ++dy) {
            float2 img{dst.x + dx * Lx, dst.y + dy * Ly};
            float2 d{img.x - dst.x, img.y - dst.y};
            float r2 = d.x * d.x + d.y * d.y;
            maxR2 = maxR2 > r2 ? maxR2 : r2;
        }
    return sqrtf(maxR2);
}

// periodic image offset from two draws: u < 1/3 -> -1, u > 2/3 -> +1  (:92-102)
static inline float2 sample_periodic_deplace(std::mt19937 &gen)
{
    std::uniform_real_distribution<float> U(0.f, 1.f);
    float2 d{0, 0};
    float ux = U(gen), uy = U(gen);
    if (ux < 1.f / 3.f) d.x = -1.f;
    else if (ux > 2.f / 3.f) d.x = 1.f;
    if (uy < 1.f / 3.f) d.y = -1.f;
    else if (uy > 2.f / 3.f) d.y = 1.f;
    return d;
}

// All particles of `sim`, both signs: est[i].main = v+ - v- at particle i
// (aux targets are not evaluated: est[i].left.. are zero).  Uses sim.mode,
// sim.seed, sim.step_index; N1/N2 samples per sign.
void monte_carlo_bs_gpu(Simulation &sim, int N1, int N2, std::vector<MCBSEstimate> &est);
