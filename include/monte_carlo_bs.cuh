// monte_carlo_bs.cuh -- device-side helpers of the GPU estimator.
// Replaces include/monte_carlo_bs.cuh:12-163 for HIP kernels written against
// the C++ API: the same names and meaning for the device helpers.  The
// reference's kernels are launch units of its step, replaced by C ABI entry
// points (vpmc.h):
//   monte_carlo_bs_kernel(_safe)   vpmc_mc_estimate / vpmc_mc_estimate_binned
//   update_jacobian_from5          vpmc_update_jacobian
//   biot_savart_naive_kernel       Simulation::compute_velocity_naive (FP64:
//                                  vpmc_direct_sum_eps)
// Its curandStatePhilox4_32_10_t samplers have no counterpart: the library
// draws from its own Philox4x32-10 streams (or a caller's replay stream), and
// cuRAND's state layout is not reproduced.
#pragma once

#include <cmath>

#include "monte_carlo_bs.hpp"
#include "vorticity_query.cuh"

// |w|-weighted (or signed) kernel sum at p over the view's 3 x 3 cells with
// the periodic fold (monte_carlo_bs.cuh:12-15; queryVorticityAbs with
// periodic = true; the deformation F is not applied here, see vpmc_field_query
// for deformed kernels)
__device__ inline float evalVorticityAbs_gpu(const VorticityView &v, float2 p, int sign,
                                             bool use_signed, bool exclude_center, float h_w)
{
    return queryVorticityAbs(v, p, sign, use_signed, exclude_center, h_w, true);
}

__device__ inline float length(float2 v) { return sqrtf(v.x * v.x + v.y * v.y); }

// F v for F = (F.x F.y; F.z F.w) row-major (:99-103)
HD inline float2 mul_mat2(float4 F, float2 v)
{
    return make_float2(F.x * v.x + F.y * v.y, F.z * v.x + F.w * v.y);
}

// d reduced by the nearest multiple of L when periodic (:134-138, :153-156)
HD inline float wrap_delta(float d, float L, float InvL, bool periodic)
{
    (void)InvL;
    return periodic ? d - roundf(d / L) * L : d;
}
HD inline float wrap_min_image(float d, float L, bool periodic)
{
    return periodic ? d - rintf(d / L) * L : d;
}

// F / sqrt(det F) (det floored at 1e-12): det -> 1 (:140-144)
HD inline float4 renorm_det1(float4 F)
{
    const float det = F.x * F.w - F.y * F.z;
    const float s = 1.f / sqrtf(fmaxf(det, 1e-12f));
    return make_float4(F.x * s, F.y * s, F.z * s, F.w * s);
}

// regularised Biot-Savart kernel perp(d) / (2 pi (|d|^2 + eps^2)) (:158-163)
HD inline float2 bs_kernel_2d(float2 d, float eps2)
{
    const float c = 1.f / (2.f * 3.14159265358979323846f * (d.x * d.x + d.y * d.y + eps2));
    return make_float2(-d.y * c, d.x * c);
}
