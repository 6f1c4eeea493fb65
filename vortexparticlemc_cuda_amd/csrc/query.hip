// query.hip -- vorticity query on the reference's hashed view (3x3 cells).
//
// Semantics of evalVorticityAbs_cpu (src/cpu/evaluate_vorticity.cpp:8-75):
// wrap the point into the box, visit the 3x3 neighbouring cells, fold each
// displacement by L when it exceeds L - 2 cell widths, cut at r^2 > h_w^2,
// filter by sign, accumulate |w| (or w) * W(r).  One thread per point; this
// is the drop-in for code that holds the reference's VorticityView.  The
// estimator uses the fine-grid field instead (field_query.hpp).
#include "vpmc_internal.hpp"

// no FMA contraction: the arithmetic matches the CPU path operation by operation
#pragma clang fp contract(off)

namespace vpmc {

struct ViewD {
    const float2 *pos;
    const float *w;
    int N;
    const int *vp_sort, *cell_acc, *sub;
    int rshx, rshy;
    float Lx, Ly;
    int cnt_view;
};

__device__ float spline_ref(float r, float h_w)
{
    // include/utilities.hpp:7-17 as written
    const float pi = 3.141592653589793f;
    const float sig2 = 40.f / (7.f * pi * h_w * h_w);
    float q = r / h_w;
    if (q >= 0.f && q <= 0.5f) return sig2 * (6.f * (q * q * q - q * q) + 1.f);
    if (q > 0.5f && q < 1.f) return sig2 * 2.f * (1.f - q) * (1.f - q) * (1.f - q);
    return 0.f;
}

__global__ void k_query_view(ViewD v, int npts, const float2 *__restrict__ q, int sign,
                             int use_signed, int exclude_center, float h_w, int periodic,
                             float *__restrict__ out)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npts) return;
    float2 p = q[i];
    if (!(isfinite(p.x) && isfinite(p.y))) {
        out[i] = 0.f;
        return;
    }
    if (periodic) {
        p.x = p.x - floorf(p.x / v.Lx) * v.Lx;
        p.y = p.y - floorf(p.y / v.Ly) * v.Ly;
    } else if (p.x < 0.f || p.x > v.Lx || p.y < 0.f || p.y > v.Ly) {
        out[i] = 0.f;
        return;
    }
    const float cellx = v.Lx / v.rshx, celly = v.Ly / v.rshy;
    const float hh = h_w * h_w;
    float sum = 0.f;
    auto visit = [&](int I) {
        float2 xj = v.pos[I];
        float dx = xj.x - p.x, dy = xj.y - p.y;
        if (periodic && !v.vp_sort) {
            // all-particle view: no cell grid to fold against (the CPU path
            // divides by rshx = 0 here); plain minimum image
            if (dx > 0.5f * v.Lx) dx -= v.Lx;
            if (dx < -0.5f * v.Lx) dx += v.Lx;
            if (dy > 0.5f * v.Ly) dy -= v.Ly;
            if (dy < -0.5f * v.Ly) dy += v.Ly;
        } else if (periodic) {
            if (dx > v.Lx - 2 * cellx) dx -= v.Lx;
            if (dx < -v.Lx + 2 * cellx) dx += v.Lx;
            if (dy > v.Ly - 2 * celly) dy -= v.Ly;
            if (dy < -v.Ly + 2 * celly) dy += v.Ly;
        }
        float r2 = __fadd_rn(__fmul_rn(dx, dx), __fmul_rn(dy, dy));
        if (r2 > hh || (exclude_center && r2 < 1e-12f)) return;
        float wv = v.w[I];
        if ((sign > 0 && wv <= 0.f) || (sign < 0 && wv >= 0.f)) return;
        float wval = use_signed ? wv : fabsf(wv);
        sum += wval * spline_ref(sqrtf(r2), h_w);
    };
    if (v.vp_sort && v.cell_acc) {
        int cx = min((int)(p.x * v.rshx / v.Lx), v.rshx - 1);
        int cy = min((int)(p.y * v.rshy / v.Ly), v.rshy - 1);
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
                for (int k = beg; k < end; ++k) visit(v.sub[v.vp_sort[k]]);
            }
    } else {
        for (int I = 0; I < v.N; ++I) visit(I);
    }
    out[i] = sum;
}

}  // namespace vpmc

using namespace vpmc;

extern "C" int vpmc_query_view(const vpmc_view *view, int npts, const float *pts_xy, int sign,
                               int use_signed, int exclude_center, float h_w, int periodic,
                               float *out, void *stream)
{
    if (!view || npts < 0 || !(h_w > 0.f)) return VPMC_ERR_ARG;
    if (npts == 0) return VPMC_OK;
    if (!pts_xy || !out || !view->pos_xy || !view->w) return VPMC_ERR_ARG;
    if (view->vp_sort && (view->rshx <= 0 || view->rshy <= 0 || !view->cell_acc || !view->sub))
        return VPMC_ERR_ARG;
    ViewD v{(const float2 *)view->pos_xy, view->w, view->N, view->vp_sort, view->cell_acc,
            view->sub, view->rshx, view->rshy, view->Lx, view->Ly, view->cnt_view};
    hipLaunchKernelGGL(k_query_view, dim3((npts + 255) / 256), dim3(256), 0, as_stream(stream), v,
                       npts, (const float2 *)pts_xy, sign, use_signed, exclude_center, h_w,
                       periodic, out);
    VPMC_CHECK_LAUNCH();
    return VPMC_OK;
}
