// vorticity_query.cuh -- device-side vorticity query on a VorticityView.
// Replaces include/vorticity_query.cuh:7-103 for HIP kernels written against
// the C++ API.  queryVorticityAbs follows evalVorticityAbs_cpu
// (src/cpu/evaluate_vorticity.cpp:8-75): 3x3 cells, displacements folded by
// L beyond L - 2 cell widths, cut at h_w; the deformation matrix F is not
// applied (the CPU path ignores it).
#pragma once

#include "utilities.hpp"
#include "vortex_particle_mc.hpp"

HD inline void minimum_image(float2 &d, float Lx, float Ly)
{
    if (d.x > 0.5f * Lx) d.x -= Lx;
    if (d.x < -0.5f * Lx) d.x += Lx;
    if (d.y > 0.5f * Ly) d.y -= Ly;
    if (d.y < -0.5f * Ly) d.y += Ly;
}

__device__ inline float queryVorticityAbs(const VorticityView &v, float2 p, int sign,
                                          bool use_signed, bool exclude_center, float h_w,
                                          bool periodic)
{
    if (periodic) {
        p.x = p.x - floorf(p.x / v.Lx) * v.Lx;
        p.y = p.y - floorf(p.y / v.Ly) * v.Ly;
    } else if (p.x < 0.f || p.x > v.Lx || p.y < 0.f || p.y > v.Ly) {
        return 0.f;
    }
    const float cellx = v.Lx / v.rshx, celly = v.Ly / v.rshy;
    int cx = min((int)(p.x * v.rshx / v.Lx), v.rshx - 1);
    int cy = min((int)(p.y * v.rshy / v.Ly), v.rshy - 1);
    float sum = 0.f;
    for (int ox = -1; ox <= 1; ++ox)
        for (int oy = -1; oy <= 1; ++oy) {
            int gx = cx + ox, gy = cy + oy;
            if (periodic) {
                gx = (gx % v.rshx + v.rshx) % v.rshx;
                gy = (gy % v.rshy + v.rshy) % v.rshy;
            } else if (gx < 0 || gy < 0 || gx >= v.rshx || gy >= v.rshy) {
                continue;
            }
            int c = gx * v.rshy + gy;
            int beg = v.cell_acc[c];
            int end = (c + 1 < v.rshx * v.rshy) ? v.cell_acc[c + 1] : v.cnt_view;
            for (int k = beg; k < end; ++k) {
                int I = v.sub[v.vp_sort[k]];
                float dx = v.pos[I].x - p.x, dy = v.pos[I].y - p.y;
                if (periodic) {
                    if (dx > v.Lx - 2 * cellx) dx -= v.Lx;
                    if (dx < -v.Lx + 2 * cellx) dx += v.Lx;
                    if (dy > v.Ly - 2 * celly) dy -= v.Ly;
                    if (dy < -v.Ly + 2 * celly) dy += v.Ly;
                }
                float r2 = dx * dx + dy * dy;
                if (r2 > h_w * h_w || (exclude_center && r2 < 1e-12f)) continue;
                float wv = v.w[I];
                if ((sign > 0 && wv <= 0.f) || (sign < 0 && wv >= 0.f)) continue;
                sum += (use_signed ? wv : fabsf(wv)) * cubicSplinePDF(sqrtf(r2), h_w);
            }
        }
    return sum;
}
