// field_query.hpp -- kernel sum over a cell-sorted particle set (device code).
//
// For a query point p the sum runs over every particle within h_w (periodic
// images by column/row wrap).  Per cell column within reach, the rows whose
// cells can hold a particle within the search radius form one contiguous run
// of the sorted array (cell key = column * ny + row), split in at most three
// pieces by the periodic wrap; each piece is streamed as float4 records.
//
// The cubic spline of include/utilities.hpp:7-17 is evaluated branch-free:
//   W(q) = sig2 * (2 t^3 - 8 u^3),  t = max(1 - q, 0),  u = max(1/2 - q, 0),
// which equals sig2 (1 - 6 q^2 + 6 q^3) for q <= 1/2, 2 sig2 (1-q)^3 for
// 1/2 < q < 1 and 0 beyond; the two cubes are accumulated separately.
#pragma once
#include "vpmc_internal.hpp"

namespace vpmc {

struct FieldD {
    const float4 *pts;
    const int *cell_start;
    FieldGeom g;
    const float4 *gq;   // deformed records: quadratic forms (G11, 2 G12, G22, sigma_max^2), or NULL
};

struct EvalCount {
    unsigned in_support = 0, visited = 0;
};

template <bool COUNT>
__device__ __forceinline__ void seg_accum(const float4 *__restrict__ pts, int beg, int end,
                                          float qx, float qy, float inv_hw, float &a1,
                                          float &a2, EvalCount &ec,
                                          const float4 *__restrict__ gq = nullptr)
{
    for (int k = beg; k < end; ++k) {
        float4 c = pts[k];
        float dx = c.x - qx, dy = c.y - qy;
        float r2 = fmaf(dx, dx, dy * dy);
        if (gq) {   // |F^-1 d|^2
            const float4 G = gq[k];
            r2 = fmaxf(fmaf(dx, fmaf(G.y, dy, G.x * dx), G.z * dy * dy), 0.f);
        }
        float q = __builtin_amdgcn_sqrtf(r2) * inv_hw;
        float t = fmaxf(1.f - q, 0.f);   // q >= 0: the upper clamp never binds
        float u = fmaxf(0.5f - q, 0.f);
        a1 = fmaf(c.z * t, t * t, a1);   // record = (x, y, |w|, 0)
        a2 = fmaf(c.z * u, u * u, a2);
        if (COUNT) {
            ec.in_support += (q < 1.f) ? 1u : 0u;
            ec.visited += 1u;
        }
    }
}

// sum_j w_j W(|p - x_j|) / sig2, p inside [0,Lx) x [0,Ly) for a periodic field
template <bool COUNT, bool PERIODIC = true>
__device__ __forceinline__ float field_sum(const FieldD &f, float px, float py, float inv_hw,
                                           EvalCount &ec)
{
    const FieldGeom &g = f.g;
    int cx = min((int)(px * g.inv_cx), g.nx - 1);
    float a1 = 0.f, a2 = 0.f;
    for (int ox = -g.reach; ox <= g.reach; ++ox) {
        int col = cx + ox;
        float gx = 0.f;
        if (ox < 0) gx = px - (float)(col + 1) * g.cw;
        else if (ox > 0) gx = (float)col * g.cw - px;
        gx = fmaxf(gx, 0.f);
        float rem = g.rs2 - gx * gx;
        if (rem < 0.f) continue;
        float ry = __builtin_amdgcn_sqrtf(rem);
        int r0 = (int)floorf((py - ry) * g.inv_cy);
        int r1 = (int)floorf((py + ry) * g.inv_cy);
        float qx = px;
        int colw = col;
        if (PERIODIC) {
            if (colw < 0) { colw += g.nx; qx = px + g.Lx; }
            else if (colw >= g.nx) { colw -= g.nx; qx = px - g.Lx; }
        } else {
            if (colw < 0 || colw >= g.nx) continue;
            r0 = max(r0, 0);
            r1 = min(r1, g.ny - 1);
            if (r1 < r0) continue;
        }
        const int *cs = f.cell_start + (size_t)colw * g.ny;
        if (PERIODIC) {
            if (r0 < 0) {   // rows ny+r0 .. ny-1 seen through the image y - Ly
                seg_accum<COUNT>(f.pts, cs[g.ny + r0], cs[g.ny], qx, py + g.Ly, inv_hw, a1, a2,
                                 ec, f.gq);
                r0 = 0;
            }
            if (r1 >= g.ny) {   // rows 0 .. r1-ny seen through the image y + Ly
                seg_accum<COUNT>(f.pts, cs[0], cs[r1 - g.ny + 1], qx, py - g.Ly, inv_hw, a1, a2,
                                 ec, f.gq);
                r1 = g.ny - 1;
            }
        }
        seg_accum<COUNT>(f.pts, cs[r0], cs[r1 + 1], qx, py, inv_hw, a1, a2, ec, f.gq);
    }
    return 2.f * a1 - 8.f * a2;
}

// FP64 variant with weights from a separate array (RBF residual refresh), or
// with the records' own weights when wts is NULL; deformed records (f.gq)
// use r^2 = d^T G d
__device__ __forceinline__ double field_sum_f64(const FieldD &f, const double *__restrict__ wts,
                                                double px, double py, double h_w)
{
    const FieldGeom &g = f.g;
    int cx = min((int)((float)px * g.inv_cx), g.nx - 1);
    double acc = 0.0;
    const double hw2 = (double)h_w * h_w;
    for (int ox = -g.reach; ox <= g.reach; ++ox) {
        int col = cx + ox;
        float gx = 0.f;
        if (ox < 0) gx = (float)px - (float)(col + 1) * g.cw;
        else if (ox > 0) gx = (float)col * g.cw - (float)px;
        gx = fmaxf(gx, 0.f);
        float rem = g.rs2 - gx * gx;
        if (rem < 0.f) continue;
        float ry = sqrtf(rem);
        int r0 = (int)floorf(((float)py - ry) * g.inv_cy);
        int r1 = (int)floorf(((float)py + ry) * g.inv_cy);
        double qx = px;
        int colw = col;
        if (colw < 0) { colw += g.nx; qx = px + g.Lx; }
        else if (colw >= g.nx) { colw -= g.nx; qx = px - g.Lx; }
        const int *cs = f.cell_start + (size_t)colw * g.ny;
        int b[3], e[3];
        double sy[3];
        int np = 0;
        if (r0 < 0) { b[np] = cs[g.ny + r0]; e[np] = cs[g.ny]; sy[np++] = py + g.Ly; r0 = 0; }
        if (r1 >= g.ny) { b[np] = cs[0]; e[np] = cs[r1 - g.ny + 1]; sy[np++] = py - g.Ly; r1 = g.ny - 1; }
        b[np] = cs[r0]; e[np] = cs[r1 + 1]; sy[np++] = py;
        for (int p = 0; p < np; ++p)
            for (int k = b[p]; k < e[p]; ++k) {
                float4 c = f.pts[k];
                double dx = (double)c.x - qx, dy = (double)c.y - sy[p];
                double r2 = dx * dx + dy * dy;
                if (f.gq) {
                    const float4 G = f.gq[k];
                    r2 = (double)G.x * dx * dx + (double)G.y * dx * dy + (double)G.z * dy * dy;
                }
                if (r2 < hw2) acc += (wts ? wts[k] : (double)c.z) * spline_d(sqrt(r2), h_w);
            }
    }
    return acc;
}

// FP32 variant with weights from a separate array (RBF matvec)
__device__ __forceinline__ float field_sum_wts(const FieldD &f, const float *__restrict__ wts,
                                               float px, float py, float inv_hw)
{
    const FieldGeom &g = f.g;
    int cx = min((int)(px * g.inv_cx), g.nx - 1);
    float a1 = 0.f, a2 = 0.f;
    for (int ox = -g.reach; ox <= g.reach; ++ox) {
        int col = cx + ox;
        float gx = 0.f;
        if (ox < 0) gx = px - (float)(col + 1) * g.cw;
        else if (ox > 0) gx = (float)col * g.cw - px;
        gx = fmaxf(gx, 0.f);
        float rem = g.rs2 - gx * gx;
        if (rem < 0.f) continue;
        float ry = __builtin_amdgcn_sqrtf(rem);
        int r0 = (int)floorf((py - ry) * g.inv_cy);
        int r1 = (int)floorf((py + ry) * g.inv_cy);
        float qx = px;
        int colw = col;
        if (colw < 0) { colw += g.nx; qx = px + g.Lx; }
        else if (colw >= g.nx) { colw -= g.nx; qx = px - g.Lx; }
        const int *cs = f.cell_start + (size_t)colw * g.ny;
        int b[3], e[3];
        float sy[3];
        int np = 0;
        if (r0 < 0) { b[np] = cs[g.ny + r0]; e[np] = cs[g.ny]; sy[np++] = py + g.Ly; r0 = 0; }
        if (r1 >= g.ny) { b[np] = cs[0]; e[np] = cs[r1 - g.ny + 1]; sy[np++] = py - g.Ly; r1 = g.ny - 1; }
        b[np] = cs[r0]; e[np] = cs[r1 + 1]; sy[np++] = py;
        for (int p = 0; p < np; ++p)
            for (int k = b[p]; k < e[p]; ++k) {
                float4 c = f.pts[k];
                float dx = c.x - qx, dy = c.y - sy[p];
                float r2 = fmaf(dx, dx, dy * dy);
                float q = __builtin_amdgcn_sqrtf(r2) * inv_hw;
                float t = fmaxf(1.f - q, 0.f);   // q >= 0: the upper clamp never binds
                float u = fmaxf(0.5f - q, 0.f);
                float wv = wts[k];
                a1 = fmaf(wv * t, t * t, a1);
                a2 = fmaf(wv * u, u * u, a2);
            }
    }
    return 2.f * a1 - 8.f * a2;
}

}  // namespace vpmc
