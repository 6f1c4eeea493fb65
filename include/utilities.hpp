// utilities.hpp -- the cubic-spline kernel W and the reference's jitter tables.
// Replaces include/utilities.hpp (:7-17 cubicSplinePDF) and the table
// builders of src/utilities.cu:12-96.
#pragma once

#include "device_vector_alias.hpp"

// W(r; h_w) = 40/(7 pi h_w^2) * {6(q^3 - q^2) + 1 (q <= 1/2); 2(1-q)^3 (q < 1); 0}
HD inline float cubicSplinePDF(float r, float h_w)
{
    const float pi = 3.141592653589793f;
    const float sig2 = 40.f / (7.f * pi * h_w * h_w);
    float q = r / h_w;
    if (q >= 0.f && q <= 0.5f) return sig2 * (6.f * (q * q * q - q * q) + 1.f);
    if (q > 0.5f && q < 1.f) return sig2 * 2.f * (1.f - q) * (1.f - q) * (1.f - q);
    return 0.f;
}

// The reference's r*W(r) tables (GPU branch-1 jitter only, src/utilities.cu).
// Built on the host; kept for interface compatibility -- the estimator's
// consistent mode inverts the radial mass exactly instead.
void build_cubic_rcdf(float h_w, int length, dvec<float> &cdf);
void cdf_to_icdf(const dvec<float> &cdf, float h_w, dvec<float> &icdf);
