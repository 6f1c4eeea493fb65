// cg.hip -- RBF Gram matvec and conjugate-gradient strength initialisation.
//
// The Gram matrix A_ij = W(|x_i - x_j|) (all particles, diagonal included,
// src/cuda/cg_rbf_init.cu:11-31) is applied matrix-free on the fine-grid
// field.  CG runs on vectors kept in the field's cell order (A is permutation
// invariant), so the weights of a cell's neighbours are read contiguously.
// Dot products accumulate in FP64 with a fixed reduction tree; with
// refinement, an outer loop recomputes the residual with an FP64 matvec and
// solves for the correction in FP32 (mixed-precision iterative refinement).
#include <math.h>

#include <vector>

#include "field_query.hpp"

namespace vpmc {

constexpr int DOT_BLOCKS = 256;
constexpr int DOT_THREADS = 256;

__global__ void k_rbf_matvec(FieldD f, float inv_hw, float sig2, int n,
                             const float *__restrict__ x, float *__restrict__ y)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    float4 p = f.pts[t];
    y[t] = sig2 * field_sum_wts(f, x, p.x, p.y, inv_hw);
}

__global__ void k_rbf_matvec_f64(FieldD f, float h_w, int n, const double *__restrict__ x,
                                 double *__restrict__ y)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    float4 p = f.pts[t];
    y[t] = field_sum_f64(f, x, (double)p.x, (double)p.y, (double)h_w);
}

template <class A, class B>
__global__ void __launch_bounds__(DOT_THREADS) k_dot(const A *__restrict__ a,
                                                     const B *__restrict__ b, int n,
                                                     double *__restrict__ part)
{
    double s = 0.0;
    for (int i = blockIdx.x * DOT_THREADS + threadIdx.x; i < n; i += DOT_BLOCKS * DOT_THREADS)
        s += (double)a[i] * (double)b[i];
    __shared__ double w[DOT_THREADS / WAVE];
    s = wave_sum(s);
    if ((threadIdx.x & (WAVE - 1)) == 0) w[threadIdx.x / WAVE] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = w[0];
        for (int k = 1; k < DOT_THREADS / WAVE; ++k) t += w[k];
        part[blockIdx.x] = t;
    }
}

__global__ void k_sum_parts(const double *__restrict__ part, double *__restrict__ out)
{
    double s = part[threadIdx.x];
    __shared__ double w[DOT_BLOCKS / WAVE];
    s = wave_sum(s);
    if ((threadIdx.x & (WAVE - 1)) == 0) w[threadIdx.x / WAVE] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = w[0];
        for (int k = 1; k < DOT_BLOCKS / WAVE; ++k) t += w[k];
        *out = t;
    }
}

__global__ void k_permute_in(int n, const int *__restrict__ order, const float *__restrict__ b,
                             float *__restrict__ bs, double *__restrict__ b64)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    float v = b[order[t]];
    bs[t] = v;
    b64[t] = (double)v;
}

__global__ void k_permute_out(int n, const int *__restrict__ order, const double *__restrict__ ws,
                              float *__restrict__ w)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    w[order[t]] = (float)ws[t];
}

// elementwise helpers, one kernel per CG update
__global__ void k_scale_f(int n, const float *__restrict__ x, float a, float *__restrict__ y)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = x[i] * a;
}
__global__ void k_sub_f(int n, const float *__restrict__ a, const float *__restrict__ b,
                        float *__restrict__ y)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = a[i] - b[i];
}
__global__ void k_cg_update(int n, float alpha, const float *__restrict__ p,
                            const float *__restrict__ Ap, float *__restrict__ w,
                            float *__restrict__ r)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    w[i] += alpha * p[i];
    r[i] -= alpha * Ap[i];
}
__global__ void k_xpby(int n, const float *__restrict__ r, float beta, float *__restrict__ p)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = r[i] + beta * p[i];
}
__global__ void k_res64(int n, const double *__restrict__ b, const double *__restrict__ Aw,
                        double *__restrict__ r64, float *__restrict__ r32)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double v = b[i] - Aw[i];
    r64[i] = v;
    r32[i] = (float)v;
}
__global__ void k_add_corr(int n, const float *__restrict__ d, double *__restrict__ w64)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) w64[i] += (double)d[i];
}
__global__ void k_f2d(int n, const float *__restrict__ x, double *__restrict__ y)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = (double)x[i];
}

struct CgBufs {
    float *pts;
    int *cell_start, *order;
    float *bs, *w, *r, *p, *Ap, *d;
    double *b64, *w64, *r64, *Aw64, *part, *scal;
    void *fws;
    size_t fws_bytes;
};

static int carve_cg(int n, int nx, int ny, void *ws, size_t bytes, CgBufs *B)
{
    Carve cv(ws, bytes);
    size_t m = (size_t)(n > 0 ? n : 1);
    B->pts = cv.take<float>(4 * m);
    B->cell_start = cv.take<int>((size_t)nx * ny + 1);
    B->order = cv.take<int>(m);
    B->bs = cv.take<float>(m);
    B->w = cv.take<float>(m);
    B->r = cv.take<float>(m);
    B->p = cv.take<float>(m);
    B->Ap = cv.take<float>(m);
    B->d = cv.take<float>(m);
    B->b64 = cv.take<double>(m);
    B->w64 = cv.take<double>(m);
    B->r64 = cv.take<double>(m);
    B->Aw64 = cv.take<double>(m);
    B->part = cv.take<double>(DOT_BLOCKS);
    B->scal = cv.take<double>(4);
    B->fws_bytes = vpmc_field_workspace_bytes(n, nx, ny);
    B->fws = cv.take<char>(B->fws_bytes);
    if (!B->pts || !B->cell_start || !B->order || !B->bs || !B->w || !B->r || !B->p || !B->Ap ||
        !B->d || !B->b64 || !B->w64 || !B->r64 || !B->Aw64 || !B->part || !B->scal || !B->fws)
        return VPMC_ERR_WORKSPACE;
    return VPMC_OK;
}

template <class A, class Bt>
static int dot_host(const A *a, const Bt *b, int n, CgBufs &B, hipStream_t s, double *out)
{
    hipLaunchKernelGGL((k_dot<A, Bt>), dim3(DOT_BLOCKS), dim3(DOT_THREADS), 0, s, a, b, n, B.part);
    hipLaunchKernelGGL(k_sum_parts, dim3(1), dim3(DOT_BLOCKS), 0, s, (const double *)B.part, B.scal);
    VPMC_CHECK_LAUNCH();
    hipError_t e = hipMemcpyAsync(out, B.scal, sizeof(double), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) { set_hip_error(e); return VPMC_ERR_HIP; }
    return VPMC_OK;
}

// FP32 CG on A x = rhs from x0 = rhs / W(0) (src/cuda/cg_rbf_init.cu:70-74),
// stopping at ||r|| / ||rhs|| < tol.  Returns iterations via *it.
static int cg_fp32(const FieldD &f, float inv_hw, float sig2, float diag, int n, const float *rhs,
                   double rhs_norm, float *x, CgBufs &B, double tol, int max_iter,
                   hipStream_t s, int *it, bool *broke)
{
    unsigned gb = (unsigned)((n + 255) / 256);
    *broke = false;
    *it = 0;
    hipLaunchKernelGGL(k_scale_f, dim3(gb), dim3(256), 0, s, n, rhs, 1.f / diag, x);
    hipLaunchKernelGGL(k_rbf_matvec, dim3(gb), dim3(256), 0, s, f, inv_hw, sig2, n,
                       (const float *)x, B.Ap);
    hipLaunchKernelGGL(k_sub_f, dim3(gb), dim3(256), 0, s, n, rhs, (const float *)B.Ap, B.r);
    hipLaunchKernelGGL(k_scale_f, dim3(gb), dim3(256), 0, s, n, (const float *)B.r, 1.f, B.p);
    VPMC_CHECK_LAUNCH();
    double r2_old;
    VPMC_TRY(dot_host(B.r, B.r, n, B, s, &r2_old));
    if (rhs_norm == 0.0 || sqrt(r2_old) / rhs_norm < tol) return VPMC_OK;
    for (int k = 0; k < max_iter; ++k) {
        hipLaunchKernelGGL(k_rbf_matvec, dim3(gb), dim3(256), 0, s, f, inv_hw, sig2, n,
                           (const float *)B.p, B.Ap);
        VPMC_CHECK_LAUNCH();
        double pAp;
        VPMC_TRY(dot_host(B.p, B.Ap, n, B, s, &pAp));
        *it = k + 1;
        if (!(pAp > 0.0)) { *broke = true; return VPMC_OK; }
        float alpha = (float)(r2_old / pAp);
        hipLaunchKernelGGL(k_cg_update, dim3(gb), dim3(256), 0, s, n, alpha, (const float *)B.p,
                           (const float *)B.Ap, x, B.r);
        VPMC_CHECK_LAUNCH();
        double r2_new;
        VPMC_TRY(dot_host(B.r, B.r, n, B, s, &r2_new));
        if (sqrt(r2_new) / rhs_norm < tol) return VPMC_OK;
        float beta = (float)(r2_new / r2_old);
        hipLaunchKernelGGL(k_xpby, dim3(gb), dim3(256), 0, s, n, (const float *)B.r, beta, B.p);
        VPMC_CHECK_LAUNCH();
        r2_old = r2_new;
    }
    return VPMC_OK;
}

}  // namespace vpmc

using namespace vpmc;

extern "C" {

size_t vpmc_rbf_workspace_bytes(int n, int nx, int ny) { return vpmc_cg_workspace_bytes(n, nx, ny); }

size_t vpmc_cg_workspace_bytes(int n, int nx, int ny)
{
    size_t m = (size_t)(n > 0 ? n : 1);
    return align256(16 * m) + align256(((size_t)nx * ny + 1) * 4) + 7 * align256(4 * m) +
           4 * align256(8 * m) + align256(8 * DOT_BLOCKS) + align256(32) +
           vpmc_field_workspace_bytes(n, nx, ny) + 4096;
}

int vpmc_rbf_matvec(const vpmc_field *f, float h_w, const float *x, float *y, void *stream)
{
    FieldGeom g;
    VPMC_TRY(field_geom(f, h_w, &g));
    if (f->n <= 0) return VPMC_OK;
    if (!x || !y || !f->pts) return VPMC_ERR_ARG;
    FieldD d{(const float4 *)f->pts, f->cell_start, g};
    hipLaunchKernelGGL(k_rbf_matvec, dim3((f->n + 255) / 256), dim3(256), 0, as_stream(stream), d,
                       1.f / h_w, 40.f / (7.f * PI_F * h_w * h_w), f->n, x, y);
    VPMC_CHECK_LAUNCH();
    return VPMC_OK;
}

int vpmc_rbf_matvec_f64(const vpmc_field *f, float h_w, const double *x, double *y, void *stream)
{
    FieldGeom g;
    VPMC_TRY(field_geom(f, h_w, &g));
    if (f->n <= 0) return VPMC_OK;
    if (!x || !y || !f->pts) return VPMC_ERR_ARG;
    FieldD d{(const float4 *)f->pts, f->cell_start, g};
    hipLaunchKernelGGL(k_rbf_matvec_f64, dim3((f->n + 255) / 256), dim3(256), 0, as_stream(stream),
                       d, h_w, f->n, x, y);
    VPMC_CHECK_LAUNCH();
    return VPMC_OK;
}

int vpmc_cg_rbf(int n, const float *pos_xy, const float *b, float *w_out, float h_w, float Lx,
                float Ly, int cells_per_hw, double tol, int max_iter, int refine, int *iters,
                double *rel_res, void *workspace, size_t workspace_bytes, void *stream)
{
    if (n < 0 || !(h_w > 0.f) || max_iter < 0) return VPMC_ERR_ARG;
    if (iters) *iters = 0;
    if (rel_res) *rel_res = 0.0;
    if (n == 0) return VPMC_OK;
    if (!pos_xy || !b || !w_out) return VPMC_ERR_ARG;
    hipStream_t s = as_stream(stream);
    int nx, ny;
    VPMC_TRY(vpmc_field_grid(h_w, Lx, Ly, cells_per_hw, &nx, &ny));
    CgBufs B;
    VPMC_TRY(carve_cg(n, nx, ny, workspace, workspace_bytes, &B));
    VPMC_TRY(field_build_impl(nx, ny, Lx, Ly, n, nullptr, nullptr, pos_xy, nullptr, 0, B.pts,
                              B.cell_start, B.order, nullptr, B.fws, B.fws_bytes, s));
    vpmc_field fh{nx, ny, Lx, Ly, n, B.pts, B.cell_start, B.order};
    FieldGeom g;
    VPMC_TRY(field_geom(&fh, h_w, &g));
    FieldD f{(const float4 *)B.pts, B.cell_start, g};
    const float inv_hw = 1.f / h_w, sig2 = 40.f / (7.f * PI_F * h_w * h_w);
    const float diag = sig2;   // W(0)
    unsigned gb = (unsigned)((n + 255) / 256);
    hipLaunchKernelGGL(k_permute_in, dim3(gb), dim3(256), 0, s, n, (const int *)B.order, b, B.bs,
                       B.b64);
    VPMC_CHECK_LAUNCH();
    double bb;
    VPMC_TRY(dot_host(B.b64, B.b64, n, B, s, &bb));
    const double bn = sqrt(bb);
    int total_it = 0;
    bool broke = false;
    double rr = 0.0;
    if (!refine) {
        int it;
        VPMC_TRY(cg_fp32(f, inv_hw, sig2, diag, n, B.bs, bn, B.w, B, tol, max_iter, s, &it,
                         &broke));
        total_it = it;
        hipLaunchKernelGGL(k_f2d, dim3(gb), dim3(256), 0, s, n, (const float *)B.w, B.w64);
        VPMC_CHECK_LAUNCH();
    } else {
        // w64 = 0; outer: r = b - A64 w64 ; solve A d = r in FP32 ; w64 += d
        hipLaunchKernelGGL(k_scale_f, dim3(gb), dim3(256), 0, s, n, (const float *)B.bs, 0.f, B.d);
        hipLaunchKernelGGL(k_f2d, dim3(gb), dim3(256), 0, s, n, (const float *)B.d, B.w64);
        VPMC_CHECK_LAUNCH();
        for (int outer = 0; outer < 30; ++outer) {
            hipLaunchKernelGGL(k_rbf_matvec_f64, dim3(gb), dim3(256), 0, s, f, h_w, n,
                               (const double *)B.w64, B.Aw64);
            hipLaunchKernelGGL(k_res64, dim3(gb), dim3(256), 0, s, n, (const double *)B.b64,
                               (const double *)B.Aw64, B.r64, B.bs);
            VPMC_CHECK_LAUNCH();
            double r2;
            VPMC_TRY(dot_host(B.r64, B.r64, n, B, s, &r2));
            rr = bn > 0 ? sqrt(r2) / bn : 0.0;
            if (rr < tol || total_it >= max_iter) break;
            int it;
            VPMC_TRY(cg_fp32(f, inv_hw, sig2, diag, n, B.bs, sqrt(r2), B.d, B, 1e-4,
                             max_iter - total_it, s, &it, &broke));
            total_it += it;
            hipLaunchKernelGGL(k_add_corr, dim3(gb), dim3(256), 0, s, n, (const float *)B.d, B.w64);
            VPMC_CHECK_LAUNCH();
            if (broke) break;
        }
    }
    // true FP64 residual of the returned strengths
    hipLaunchKernelGGL(k_rbf_matvec_f64, dim3(gb), dim3(256), 0, s, f, h_w, n,
                       (const double *)B.w64, B.Aw64);
    hipLaunchKernelGGL(k_res64, dim3(gb), dim3(256), 0, s, n, (const double *)B.b64,
                       (const double *)B.Aw64, B.r64, B.d);
    hipLaunchKernelGGL(k_permute_out, dim3(gb), dim3(256), 0, s, n, (const int *)B.order,
                       (const double *)B.w64, w_out);
    VPMC_CHECK_LAUNCH();
    double r2;
    VPMC_TRY(dot_host(B.r64, B.r64, n, B, s, &r2));
    rr = bn > 0 ? sqrt(r2) / bn : 0.0;
    if (iters) *iters = total_it;
    if (rel_res) *rel_res = rr;
    if (broke) return VPMC_ERR_CG_BREAKDOWN;
    if (!(rr < tol) && tol > 0.0) {
        // the FP32 recursion may stop on its own residual before the FP64
        // residual of the cast result meets tol
        if (total_it >= max_iter) return VPMC_ERR_NOT_CONVERGED;
    }
    return VPMC_OK;
}

}  // extern "C"
