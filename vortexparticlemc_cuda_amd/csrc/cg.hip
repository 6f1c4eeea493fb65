// cg.hip -- RBF Gram matvec and conjugate-gradient strength initialisation.
//
// The Gram matrix A_ij = W(|x_i - x_j|) (all particles, diagonal included,
// src/cuda/cg_rbf_init.cu:11-31) is applied matrix-free on the fine-grid
// field.  CG runs on vectors kept in the field's cell order (A is permutation
// invariant), so the weights of a cell's neighbours are read contiguously.
// Dot products accumulate in FP64 with a fixed reduction tree.  Without
// refinement the solve is the reference's FP32 CG; with it, MINRES runs on
// FP64 vectors and an FP64 matvec from zero.  (FP32 corrections on an FP64
// residual stall near 1e-7: the residual of a rounded iterate spreads over
// every lattice mode, and the indefinite Gram matrix has eigenvalues on both
// sides of zero, so each correction would need thousands of iterations.)
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

__global__ void k_permute_in64(int n, const int *__restrict__ order, const double *__restrict__ b,
                               float *__restrict__ bs, double *__restrict__ b64)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    double v = b[order[t]];
    bs[t] = (float)v;
    b64[t] = v;
}

__global__ void k_permute_out(int n, const int *__restrict__ order, const double *__restrict__ ws,
                              float *__restrict__ w, double *__restrict__ w64)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    w[order[t]] = (float)ws[t];
    if (w64) w64[order[t]] = ws[t];
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
                        double *__restrict__ r64)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) r64[i] = b[i] - Aw[i];
}
__global__ void k_f2d(int n, const float *__restrict__ x, double *__restrict__ y)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = (double)x[i];
}

struct CgBufs {
    float *pts;
    int *cell_start, *order;
    float *bs, *w, *r, *p, *Ap;          // FP32 CG
    double *b64, *w64, *r64, *Aw64;      // FP64 right-hand side, iterate, residual
    double *m[7];                        // FP64 MINRES vectors
    double *part, *scal;
    void *fws;
    size_t fws_bytes;
};

constexpr int CG_F32_VECS = 5, CG_F64_VECS = 4 + 7;

static int carve_cg(int n, int nx, int ny, void *ws, size_t bytes, CgBufs *B)
{
    Carve cv(ws, bytes);
    size_t m = (size_t)(n > 0 ? n : 1);
    B->pts = cv.take<float>(4 * m);
    B->cell_start = cv.take<int>((size_t)nx * ny + 1);
    B->order = cv.take<int>(m);
    float **f32[CG_F32_VECS] = {&B->bs, &B->w, &B->r, &B->p, &B->Ap};
    double **f64[CG_F64_VECS] = {&B->b64, &B->w64, &B->r64, &B->Aw64, &B->m[0], &B->m[1],
                                 &B->m[2], &B->m[3], &B->m[4], &B->m[5], &B->m[6]};
    bool ok = B->pts && B->cell_start && B->order;
    for (float **q : f32) ok = ok && (*q = cv.take<float>(m)) != nullptr;
    for (double **q : f64) ok = ok && (*q = cv.take<double>(m)) != nullptr;
    B->part = cv.take<double>(DOT_BLOCKS);
    B->scal = cv.take<double>(4);
    B->fws_bytes = vpmc_field_workspace_bytes(n, nx, ny);
    B->fws = cv.take<char>(B->fws_bytes);
    if (!ok || !B->part || !B->scal || !B->fws) return VPMC_ERR_WORKSPACE;
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

template <class T>
__global__ void k_scale(int n, const T *__restrict__ x, T a, T *__restrict__ y)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = x[i] * a;
}
template <class T>
__global__ void k_axpy(int n, T a, const T *__restrict__ x, T *__restrict__ y)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] += a * x[i];
}
template <class T>
__global__ void k_minres_dir(int n, const T *__restrict__ v, const T *__restrict__ w1,
                             const T *__restrict__ w2, T oldeps, T delta, T inv_gamma, T phi,
                             T *__restrict__ w, T *__restrict__ x)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    T wn = (v[i] - oldeps * w1[i] - delta * w2[i]) * inv_gamma;
    w[i] = wn;
    x[i] += phi * wn;
}

// ---- circulant preconditioner for particles on a uniform periodic lattice ----
// On the cell-centred nx_l x ny_l lattice, particle i*ny_l + j at
// ((i+1/2) hx, (j+1/2) hy), the Gram matrix is block-circulant with circulant
// blocks (up to the FP32 rounding of the positions): its eigenvectors are the
// lattice Fourier modes, with eigenvalues
//   lambda(k1, k2) = sum_o W(|(o1 hx, o2 hy)|) cos(2 pi k1 o1 / nx_l) cos(2 pi k2 o2 / ny_l)
// (o over the offsets within h_w; W is even in each component, so the sines
// cancel).  lambda changes sign (the matrix is indefinite) and is tiny at
// high |k|, which is where unpreconditioned MINRES stalls.  MINRES takes the
// SPD preconditioner M = F^-1 diag(max(|lambda|, 1e-10 lambda(0))) F, applied
// with four passes of a direct DFT along lattice lines (FP64, twiddles in LDS).

constexpr int DFT_THREADS = 256;
constexpr int DFT_LDS_BYTES = 65536;
constexpr int DFT_MAX_LINE = DFT_LDS_BYTES / 32;   // twiddles + one line

__global__ void k_lattice_lambda(int nxl, int nyl, double hx, double hy, double h_w, int rx,
                                 int ry, double *__restrict__ lam)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nxl * nyl) return;
    const int k1 = i / nyl, k2 = i - k1 * nyl;
    double s = 0.0;
    for (int o1 = -rx; o1 <= rx; ++o1) {
        const long long m1 = ((long long)k1 * (o1 < 0 ? -o1 : o1)) % nxl;
        const double c1 = cospi(2.0 * (double)m1 / (double)nxl);
        for (int o2 = -ry; o2 <= ry; ++o2) {
            const double r = hypot(o1 * hx, o2 * hy);
            if (!(r < h_w)) continue;
            const long long m2 = ((long long)k2 * (o2 < 0 ? -o2 : o2)) % nyl;
            s += spline_d(r, h_w) * c1 * cospi(2.0 * (double)m2 / (double)nyl);
        }
    }
    lam[i] = s;
}

// scale[i] = 1 / (max(|lambda_i|, 1e-10 lambda(0)) nx ny); lambda(0) is the
// largest eigenvalue (W >= 0)
__global__ void k_lattice_scale(int n, const double *__restrict__ lam, double *__restrict__ scale)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double floor_ = 1e-10 * fabs(lam[0]);
    scale[i] = 1.0 / (fmax(fabs(lam[i]), floor_) * (double)n);
}

// out[k * rows + r] = sum_j in[r * L + j] exp(sign 2 pi i k j / L) for the
// rows lines of length L (REAL_IN: in holds doubles; REAL_OUT: the real
// part only).  A workgroup takes `per` lines; LDS: L twiddles + per lines.
template <bool REAL_IN, bool REAL_OUT>
__global__ void __launch_bounds__(DFT_THREADS)
    k_dft_lines(int rows, int L, int per, double sign, const void *__restrict__ in,
                void *__restrict__ out)
{
    extern __shared__ double2 dft_lds[];
    double2 *tw = dft_lds, *line = dft_lds + L;
    for (int m = threadIdx.x; m < L; m += DFT_THREADS) {
        double sn, cs;
        sincospi(2.0 * (double)m / (double)L, &sn, &cs);
        tw[m] = make_double2(cs, sign * sn);
    }
    const int r0 = blockIdx.x * per;
    const int nr = min(per, rows - r0);
    for (int t = threadIdx.x; t < nr * L; t += DFT_THREADS) {
        const size_t src = (size_t)r0 * L + t;
        line[t] = REAL_IN ? make_double2(((const double *)in)[src], 0.0)
                          : ((const double2 *)in)[src];
    }
    __syncthreads();
    for (int t = threadIdx.x; t < nr * L; t += DFT_THREADS) {
        const int k = t / nr, rr = t - k * nr;   // consecutive lanes: consecutive lines
        const double2 *ln = line + (size_t)rr * L;
        double ar = 0.0, ai = 0.0;
        int idx = 0;
        for (int j = 0; j < L; ++j) {
            const double2 x = ln[j], w = tw[idx];
            ar = fma(x.x, w.x, fma(-x.y, w.y, ar));
            ai = fma(x.x, w.y, fma(x.y, w.x, ai));
            idx += k;
            if (idx >= L) idx -= L;
        }
        const size_t dst = (size_t)k * rows + r0 + rr;
        if (REAL_OUT) ((double *)out)[dst] = ar;
        else ((double2 *)out)[dst] = make_double2(ar, ai);
    }
}

__global__ void k_to_lattice(int n, const int *__restrict__ order, const double *__restrict__ x,
                             double *__restrict__ lat)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) lat[order[t]] = x[t];
}
__global__ void k_from_lattice(int n, const int *__restrict__ order,
                               const double *__restrict__ lat, double *__restrict__ x)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) x[t] = lat[order[t]];
}
__global__ void k_scale_spectrum(int n, const double *__restrict__ scale, double2 *__restrict__ c)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double a = scale[i];
    c[i] = make_double2(c[i].x * a, c[i].y * a);
}

struct LatticePc {
    int nxl = 0, nyl = 0;
    const int *order = nullptr;   // field position -> particle (= lattice) index
    double *scale = nullptr;      // [k1 * ny + k2]
    double *lat = nullptr;
    double *z = nullptr;          // MINRES: M^-1 of the newest Lanczos vector
    double2 *c1 = nullptr, *c2 = nullptr;
};

inline bool lattice_pc_ok(int nxl, int nyl)
{
    return nxl >= 1 && nyl >= 1 && nxl <= DFT_MAX_LINE && nyl <= DFT_MAX_LINE;
}

static size_t lattice_pc_bytes(int nxl, int nyl)
{
    const size_t n = (size_t)nxl * nyl;
    return 3 * align256(8 * n) + 2 * align256(16 * n);
}

static int lattice_dft(int rows, int L, bool real_in, bool real_out, double sign, const void *in,
                       void *out, hipStream_t s)
{
    int per = DFT_LDS_BYTES / (16 * L) - 1;
    per = per < 1 ? 1 : (per > 16 ? 16 : per);
    const size_t lds = (size_t)16 * L * (1 + per);
    const dim3 grid((unsigned)((rows + per - 1) / per)), block(DFT_THREADS);
    if (real_in)
        hipLaunchKernelGGL((k_dft_lines<true, false>), grid, block, lds, s, rows, L, per, sign,
                           in, out);
    else if (real_out)
        hipLaunchKernelGGL((k_dft_lines<false, true>), grid, block, lds, s, rows, L, per, sign,
                           in, out);
    else
        hipLaunchKernelGGL((k_dft_lines<false, false>), grid, block, lds, s, rows, L, per, sign,
                           in, out);
    VPMC_CHECK_LAUNCH();
    return VPMC_OK;
}

static int lattice_pc_init(LatticePc &P, int nxl, int nyl, float Lx, float Ly, float h_w,
                           const int *order, void *ws, size_t bytes, hipStream_t s)
{
    Carve cv(ws, bytes);
    const size_t n = (size_t)nxl * nyl;
    P.nxl = nxl;
    P.nyl = nyl;
    P.order = order;
    P.scale = cv.take<double>(n);
    P.lat = cv.take<double>(n);
    P.z = cv.take<double>(n);
    P.c1 = cv.take<double2>(n);
    P.c2 = cv.take<double2>(n);
    if (!P.scale || !P.lat || !P.z || !P.c1 || !P.c2) return VPMC_ERR_WORKSPACE;
    const double hx = (double)Lx / nxl, hy = (double)Ly / nyl;
    const int rx = (int)floor((double)h_w / hx), ry = (int)floor((double)h_w / hy);
    const unsigned gb = (unsigned)((n + 255) / 256);
    hipLaunchKernelGGL(k_lattice_lambda, dim3(gb), dim3(256), 0, s, nxl, nyl, hx, hy,
                       (double)h_w, rx, ry, P.lat);
    hipLaunchKernelGGL(k_lattice_scale, dim3(gb), dim3(256), 0, s, (int)n,
                       (const double *)P.lat, P.scale);
    VPMC_CHECK_LAUNCH();
    return VPMC_OK;
}

// out = M^-1 in, both in the field's cell order
static int lattice_pc_apply(const LatticePc &P, const double *in, double *out, hipStream_t s)
{
    const int n = P.nxl * P.nyl;
    const unsigned gb = (unsigned)((n + 255) / 256);
    hipLaunchKernelGGL(k_to_lattice, dim3(gb), dim3(256), 0, s, n, P.order, in, P.lat);
    VPMC_CHECK_LAUNCH();
    // [ix][iy] -> [k2][ix] -> [k1][k2], scale, -> [iy][k1] -> [ix][iy]
    VPMC_TRY(lattice_dft(P.nxl, P.nyl, true, false, -1.0, P.lat, P.c1, s));
    VPMC_TRY(lattice_dft(P.nyl, P.nxl, false, false, -1.0, P.c1, P.c2, s));
    hipLaunchKernelGGL(k_scale_spectrum, dim3(gb), dim3(256), 0, s, n,
                       (const double *)P.scale, P.c2);
    VPMC_CHECK_LAUNCH();
    VPMC_TRY(lattice_dft(P.nxl, P.nyl, false, false, 1.0, P.c2, P.c1, s));
    VPMC_TRY(lattice_dft(P.nyl, P.nxl, false, true, 1.0, P.c1, P.lat, s));
    hipLaunchKernelGGL(k_from_lattice, dim3(gb), dim3(256), 0, s, n, P.order,
                       (const double *)P.lat, out);
    VPMC_CHECK_LAUNCH();
    return VPMC_OK;
}

// MINRES (Paige & Saunders, SIAM J. Numer. Anal. 12, 1975) on A x = rhs from
// x = 0, FP64 vectors and FP64 matvec: the Gram matrix is symmetric but
// indefinite, where CG can meet p.Ap <= 0.  Stops when the residual estimate
// phibar < tol * ||rhs|| or after max_iter iterations.
static int minres_f64(const FieldD &f, float h_w, int n, const double *rhs, double *x,
                      CgBufs &B, double tol, int max_iter, hipStream_t s, int *it,
                      const LatticePc *pc = nullptr)
{
    unsigned gb = (unsigned)((n + 255) / 256);
    *it = 0;
    double *r1 = B.m[0], *r2 = B.m[1], *y = B.m[2], *v = B.m[3], *w = B.m[4], *w1 = B.m[5],
           *w2 = B.m[6];
    // z = M^-1 r2 (the preconditioned Lanczos vector); without M, z is r2
    double *z = pc ? pc->z : nullptr;
    hipLaunchKernelGGL(k_scale<double>, dim3(gb), dim3(256), 0, s, n, rhs, 0.0, x);
    hipLaunchKernelGGL(k_scale<double>, dim3(gb), dim3(256), 0, s, n, rhs, 0.0, w);
    hipLaunchKernelGGL(k_scale<double>, dim3(gb), dim3(256), 0, s, n, rhs, 0.0, w1);
    hipLaunchKernelGGL(k_scale<double>, dim3(gb), dim3(256), 0, s, n, rhs, 0.0, w2);
    hipLaunchKernelGGL(k_scale<double>, dim3(gb), dim3(256), 0, s, n, rhs, 1.0, r1);
    hipLaunchKernelGGL(k_scale<double>, dim3(gb), dim3(256), 0, s, n, rhs, 1.0, r2);
    VPMC_CHECK_LAUNCH();
    double b2;
    VPMC_TRY(dot_host(rhs, rhs, n, B, s, &b2));
    const double bnorm = sqrt(b2);
    if (bnorm == 0.0) return VPMC_OK;
    if (pc) {
        VPMC_TRY(lattice_pc_apply(*pc, rhs, z, s));
        VPMC_TRY(dot_host(rhs, (const double *)z, n, B, s, &b2));
        if (!(b2 > 0.0)) return VPMC_ERR_ARG;
    }
    const double beta1 = sqrt(b2);
    double oldb = 0.0, beta = beta1, dbar = 0.0, epsln = 0.0, phibar = beta1, cs = -1.0, sn = 0.0;
    // preconditioned: phibar estimates |r|_{M^-1}; the FP64 residual |b - A x|
    // is checked whenever the estimate passes the next decade below tol
    double check_at = tol * beta1;
    // r1, r2: the last two Lanczos residuals; y: free buffer the matvec fills
    for (int itn = 1; itn <= max_iter; ++itn) {
        hipLaunchKernelGGL(k_scale<double>, dim3(gb), dim3(256), 0, s, n,
                           (const double *)(pc ? z : r2), 1.0 / beta, v);
        hipLaunchKernelGGL(k_rbf_matvec_f64, dim3(gb), dim3(256), 0, s, f, h_w, n,
                           (const double *)v, y);
        if (itn >= 2)
            hipLaunchKernelGGL(k_axpy<double>, dim3(gb), dim3(256), 0, s, n, -beta / oldb,
                               (const double *)r1, y);
        VPMC_CHECK_LAUNCH();
        double alfa;
        VPMC_TRY(dot_host(v, y, n, B, s, &alfa));
        hipLaunchKernelGGL(k_axpy<double>, dim3(gb), dim3(256), 0, s, n, -alfa / beta,
                           (const double *)r2, y);
        VPMC_CHECK_LAUNCH();
        double *t = r1;
        r1 = r2;
        r2 = y;
        y = t;
        oldb = beta;
        double bb;
        if (pc) {
            VPMC_TRY(lattice_pc_apply(*pc, r2, z, s));
            VPMC_TRY(dot_host(r2, (const double *)z, n, B, s, &bb));
        } else {
            VPMC_TRY(dot_host(r2, r2, n, B, s, &bb));
        }
        beta = sqrt(fmax(bb, 0.0));
        double oldeps = epsln;
        double delta = cs * dbar + sn * alfa;
        double gbar = sn * dbar - cs * alfa;
        epsln = sn * beta;
        dbar = -cs * beta;
        double gamma = fmax(hypot(gbar, beta), 1e-300);
        cs = gbar / gamma;
        sn = beta / gamma;
        double phi = cs * phibar;
        phibar = sn * phibar;
        t = w1;
        w1 = w2;
        w2 = w;
        w = t;
        hipLaunchKernelGGL(k_minres_dir<double>, dim3(gb), dim3(256), 0, s, n,
                           (const double *)v, (const double *)w1, (const double *)w2, oldeps,
                           delta, 1.0 / gamma, phi, w, x);
        VPMC_CHECK_LAUNCH();
        *it = itn;
        if (beta == 0.0) break;
        if (!pc) {
            if (phibar < tol * beta1) break;
        } else if (phibar < check_at) {
            hipLaunchKernelGGL(k_rbf_matvec_f64, dim3(gb), dim3(256), 0, s, f, h_w, n,
                               (const double *)x, B.Aw64);
            hipLaunchKernelGGL(k_res64, dim3(gb), dim3(256), 0, s, n, rhs,
                               (const double *)B.Aw64, B.r64);
            VPMC_CHECK_LAUNCH();
            double r2n;
            VPMC_TRY(dot_host(B.r64, B.r64, n, B, s, &r2n));
            if (sqrt(r2n) < tol * bnorm) break;
            check_at *= 0.1;
        }
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
    return align256(16 * m) + align256(((size_t)nx * ny + 1) * 4) + align256(4 * m) +
           CG_F32_VECS * align256(4 * m) + CG_F64_VECS * align256(8 * m) +
           align256(8 * DOT_BLOCKS) + align256(32) +
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

__global__ void k_gather_sorted(int n, const int *__restrict__ order, const float *__restrict__ x,
                                float *__restrict__ xs)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) xs[t] = x[order[t]];
}
__global__ void k_scatter_sorted(int n, const int *__restrict__ order,
                                 const float *__restrict__ ys, float *__restrict__ y)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) y[order[t]] = ys[t];
}

int vpmc_rbf_matvec_particles(int n, const float *pos_xy, const float *x, float *y, float h_w,
                              float Lx, float Ly, int cells_per_hw, void *workspace,
                              size_t workspace_bytes, void *stream)
{
    if (n < 0 || !(h_w > 0.f)) return VPMC_ERR_ARG;
    if (n == 0) return VPMC_OK;
    if (!pos_xy || !x || !y) return VPMC_ERR_ARG;
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
    unsigned gb = (unsigned)((n + 255) / 256);
    hipLaunchKernelGGL(k_gather_sorted, dim3(gb), dim3(256), 0, s, n, (const int *)B.order, x, B.p);
    hipLaunchKernelGGL(k_rbf_matvec, dim3(gb), dim3(256), 0, s, f, 1.f / h_w,
                       40.f / (7.f * PI_F * h_w * h_w), n, (const float *)B.p, B.Ap);
    hipLaunchKernelGGL(k_scatter_sorted, dim3(gb), dim3(256), 0, s, n, (const int *)B.order,
                       (const float *)B.Ap, y);
    VPMC_CHECK_LAUNCH();
    return VPMC_OK;
}

}  // extern "C"

namespace vpmc {

static int cg_rbf_impl(int n, const float *pos_xy, const float *b, const double *b_f64,
                       float *w_out, double *w64_out, float h_w, float Lx, float Ly,
                       int cells_per_hw, double tol, int max_iter, int refine, int *iters,
                       double *rel_res, void *workspace, size_t workspace_bytes, void *stream,
                       int nx_l = 0, int ny_l = 0)
{
    if (n < 0 || !(h_w > 0.f) || max_iter < 0) return VPMC_ERR_ARG;
    if (iters) *iters = 0;
    if (rel_res) *rel_res = 0.0;
    if (n == 0) return VPMC_OK;
    if (!pos_xy || (!b && !b_f64) || !w_out) return VPMC_ERR_ARG;
    hipStream_t s = as_stream(stream);
    int nx, ny;
    VPMC_TRY(vpmc_field_grid(h_w, Lx, Ly, cells_per_hw, &nx, &ny));
    CgBufs B;
    const bool lattice = nx_l > 0 && refine;
    if (lattice && ((long long)nx_l * ny_l != n || !lattice_pc_ok(nx_l, ny_l)))
        return VPMC_ERR_ARG;
    const size_t pc_bytes = lattice ? lattice_pc_bytes(nx_l, ny_l) : 0;
    if (workspace_bytes < pc_bytes) return VPMC_ERR_WORKSPACE;
    // [0, pc_off): the solve's buffers; [pc_off, end): the preconditioner's
    const size_t pc_off = lattice ? (workspace_bytes - pc_bytes) / 256 * 256 : workspace_bytes;
    VPMC_TRY(carve_cg(n, nx, ny, workspace, pc_off, &B));
    VPMC_TRY(field_build_impl(nx, ny, Lx, Ly, n, nullptr, nullptr, pos_xy, nullptr, 0, B.pts,
                              B.cell_start, B.order, nullptr, B.fws, B.fws_bytes, s));
    vpmc_field fh{nx, ny, Lx, Ly, n, B.pts, B.cell_start, B.order};
    FieldGeom g;
    VPMC_TRY(field_geom(&fh, h_w, &g));
    FieldD f{(const float4 *)B.pts, B.cell_start, g};
    const float inv_hw = 1.f / h_w, sig2 = 40.f / (7.f * PI_F * h_w * h_w);
    const float diag = sig2;   // W(0)
    unsigned gb = (unsigned)((n + 255) / 256);
    if (b_f64)
        hipLaunchKernelGGL(k_permute_in64, dim3(gb), dim3(256), 0, s, n, (const int *)B.order,
                           b_f64, B.bs, B.b64);
    else
        hipLaunchKernelGGL(k_permute_in, dim3(gb), dim3(256), 0, s, n, (const int *)B.order, b,
                           B.bs, B.b64);
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
    } else if (lattice) {
        // the preconditioner's buffers sit at the end of the workspace
        LatticePc P;
        VPMC_TRY(lattice_pc_init(P, nx_l, ny_l, Lx, Ly, h_w, B.order,
                                 static_cast<char *>(workspace) + pc_off,
                                 workspace_bytes - pc_off, s));
        VPMC_TRY(minres_f64(f, h_w, n, (const double *)B.b64, B.w64, B, tol, max_iter, s,
                            &total_it, &P));
    } else {
        VPMC_TRY(minres_f64(f, h_w, n, (const double *)B.b64, B.w64, B, tol, max_iter, s,
                            &total_it));
    }
    // FP64 residual of the FP64 iterate (the strengths are returned rounded to FP32)
    hipLaunchKernelGGL(k_rbf_matvec_f64, dim3(gb), dim3(256), 0, s, f, h_w, n,
                       (const double *)B.w64, B.Aw64);
    hipLaunchKernelGGL(k_res64, dim3(gb), dim3(256), 0, s, n, (const double *)B.b64,
                       (const double *)B.Aw64, B.r64);
    hipLaunchKernelGGL(k_permute_out, dim3(gb), dim3(256), 0, s, n, (const int *)B.order,
                       (const double *)B.w64, w_out, w64_out);
    VPMC_CHECK_LAUNCH();
    double r2;
    VPMC_TRY(dot_host(B.r64, B.r64, n, B, s, &r2));
    rr = bn > 0 ? sqrt(r2) / bn : 0.0;
    if (iters) *iters = total_it;
    if (rel_res) *rel_res = rr;
    if (broke) return VPMC_ERR_CG_BREAKDOWN;
    if (!(rr < tol) && tol > 0.0) {
        // with refinement the FP64 residual is the criterion; without it the
        // FP32 recursion stops on its own residual, which the FP64 residual of
        // the returned strengths may exceed slightly -- only max_iter counts
        if (refine || total_it >= max_iter) return VPMC_ERR_NOT_CONVERGED;
    }
    return VPMC_OK;
}

}  // namespace vpmc

extern "C" {

int vpmc_cg_rbf(int n, const float *pos_xy, const float *b, float *w_out, float h_w, float Lx,
                float Ly, int cells_per_hw, double tol, int max_iter, int refine, int *iters,
                double *rel_res, void *workspace, size_t workspace_bytes, void *stream)
{
    return cg_rbf_impl(n, pos_xy, b, nullptr, w_out, nullptr, h_w, Lx, Ly, cells_per_hw, tol,
                       max_iter, refine, iters, rel_res, workspace, workspace_bytes, stream);
}

int vpmc_cg_rbf_f64(int n, const float *pos_xy, const double *b, float *w_out, double *w64_out,
                    float h_w, float Lx, float Ly, int cells_per_hw, double tol, int max_iter,
                    int *iters, double *rel_res, void *workspace, size_t workspace_bytes,
                    void *stream)
{
    return cg_rbf_impl(n, pos_xy, nullptr, b, w_out, w64_out, h_w, Lx, Ly, cells_per_hw, tol,
                       max_iter, 1, iters, rel_res, workspace, workspace_bytes, stream);
}

size_t vpmc_cg_lattice_workspace_bytes(int nx_l, int ny_l, float h_w, float Lx, float Ly,
                                       int cells_per_hw)
{
    int nx, ny;
    if (!lattice_pc_ok(nx_l, ny_l) || vpmc_field_grid(h_w, Lx, Ly, cells_per_hw, &nx, &ny))
        return 0;
    return vpmc_cg_workspace_bytes(nx_l * ny_l, nx, ny) + lattice_pc_bytes(nx_l, ny_l) + 256;
}

int vpmc_cg_rbf_lattice(int nx_l, int ny_l, const float *pos_xy, const double *b, float *w_out,
                        double *w64_out, float h_w, float Lx, float Ly, int cells_per_hw,
                        double tol, int max_iter, int *iters, double *rel_res, void *workspace,
                        size_t workspace_bytes, void *stream)
{
    if (!lattice_pc_ok(nx_l, ny_l)) return VPMC_ERR_ARG;
    return cg_rbf_impl(nx_l * ny_l, pos_xy, nullptr, b, w_out, w64_out, h_w, Lx, Ly,
                       cells_per_hw, tol, max_iter, 1, iters, rel_res, workspace,
                       workspace_bytes, stream, nx_l, ny_l);
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Reinitialisation (README.md:319-321, 416-420; the reference describes it
// and has no code for it): the current particles' signed field -- deformed
// kernels when F is given -- evaluated in FP64 at a fresh cell-centred
// lattice, then lattice strengths fitted to it by the FP64-RHS RBF solve.

namespace vpmc {

__global__ void k_reinit_rhs(FieldD f, double h_w, int nxl, int nyl, float Lx, float Ly,
                             float2 *__restrict__ pos_out, double *__restrict__ b)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nxl * nyl) return;
    const int ix = i / nyl, iy = i - ix * nyl;   // index i*Ny + j (x-major)
    const float px = ((float)ix + 0.5f) * (Lx / (float)nxl);
    const float py = ((float)iy + 0.5f) * (Ly / (float)nyl);
    pos_out[i] = make_float2(px, py);
    b[i] = field_sum_f64(f, nullptr, px, py, h_w);
}

struct ReinitBufs {
    float *pts, *gq;
    int *cell_start, *order;
    double *b;
    void *rest;
    size_t rest_bytes;
};

static int carve_reinit(int n, int nx_l, int ny_l, int nx, int ny, bool def, void *ws,
                        size_t bytes, ReinitBufs *B)
{
    Carve cv(ws, bytes);
    const size_t m = (size_t)(n > 0 ? n : 1);
    const int n_new = nx_l * ny_l;
    B->pts = cv.take<float>(4 * m);
    B->gq = def ? cv.take<float>(4 * m) : nullptr;
    B->cell_start = cv.take<int>((size_t)nx * ny + 1);
    B->order = cv.take<int>(m);
    B->b = cv.take<double>((size_t)(n_new > 0 ? n_new : 1));
    B->rest_bytes = std::max(vpmc_field_workspace_bytes(n, nx, ny),
                             vpmc_cg_workspace_bytes(n_new, nx, ny));
    if (lattice_pc_ok(nx_l, ny_l))
        B->rest_bytes = std::max(B->rest_bytes, vpmc_cg_workspace_bytes(n_new, nx, ny) +
                                                    lattice_pc_bytes(nx_l, ny_l) + 256);
    B->rest = cv.take<char>(B->rest_bytes);
    if (!B->pts || (def && !B->gq) || !B->cell_start || !B->order || !B->b || !B->rest)
        return VPMC_ERR_WORKSPACE;
    return VPMC_OK;
}

}  // namespace vpmc

extern "C" {

size_t vpmc_reinit_workspace_bytes(int n, int nx_l, int ny_l, float h_w, float Lx, float Ly,
                                   int cells_per_hw, int deformed)
{
    int nx, ny;
    if (nx_l <= 0 || ny_l <= 0 || vpmc_field_grid(h_w, Lx, Ly, cells_per_hw, &nx, &ny) != VPMC_OK)
        return 0;
    const int n_new = nx_l * ny_l;
    const size_t m = (size_t)(n > 0 ? n : 1), nn = (size_t)n_new;
    size_t rest = std::max(vpmc_field_workspace_bytes(n, nx, ny),
                           vpmc_cg_workspace_bytes(n_new, nx, ny));
    if (lattice_pc_ok(nx_l, ny_l))
        rest = std::max(rest, vpmc_cg_workspace_bytes(n_new, nx, ny) +
                                  lattice_pc_bytes(nx_l, ny_l) + 256);
    return align256(16 * m) * (deformed ? 2 : 1) + align256(((size_t)nx * ny + 1) * 4) +
           align256(4 * m) + align256(8 * nn) + align256(rest) + 4096;
}

int vpmc_reinit(int n, const float *pos_xy, const float *w, const float *jac, float h_w,
                float Lx, float Ly, int cells_per_hw, int nx_l, int ny_l, double tol,
                int max_iter, float *pos_out, float *w_out, int *iters, double *rel_res,
                void *workspace, size_t workspace_bytes, void *stream)
{
    if (n < 0 || nx_l <= 0 || ny_l <= 0 || !(h_w > 0.f) || !pos_out || !w_out ||
        (n > 0 && (!pos_xy || !w)))
        return VPMC_ERR_ARG;
    if ((long long)nx_l * ny_l >= (1LL << 31)) return VPMC_ERR_ARG;
    hipStream_t s = as_stream(stream);
    int nx, ny;
    VPMC_TRY(vpmc_field_grid(h_w, Lx, Ly, cells_per_hw, &nx, &ny));
    const int n_new = nx_l * ny_l;
    ReinitBufs B;
    VPMC_TRY(carve_reinit(n, nx_l, ny_l, nx, ny, jac != nullptr, workspace, workspace_bytes,
                          &B));
    // the field of every particle with its signed strength
    VPMC_TRY(field_build_impl(nx, ny, Lx, Ly, n, nullptr, nullptr, pos_xy, w, 0, B.pts,
                              B.cell_start, B.order, nullptr, B.rest, B.rest_bytes, s, jac,
                              B.gq, nullptr));
    vpmc_field fh{nx, ny, Lx, Ly, n, B.pts, B.cell_start, B.order, B.gq, 0.f};
    FieldGeom g;
    VPMC_TRY(field_geom(&fh, h_w, &g));
    FieldD f{(const float4 *)B.pts, B.cell_start, g, (const float4 *)B.gq};
    hipLaunchKernelGGL(k_reinit_rhs, dim3((n_new + 255) / 256), dim3(256), 0, s, f,
                       (double)h_w, nx_l, ny_l, Lx, Ly, (float2 *)pos_out, B.b);
    VPMC_CHECK_LAUNCH();
    // the lattice's circulant preconditioner where its lines fit the DFT kernel
    const bool pc = lattice_pc_ok(nx_l, ny_l);
    return cg_rbf_impl(n_new, pos_out, nullptr, B.b, w_out, nullptr, h_w, Lx, Ly, cells_per_hw,
                       tol, max_iter, 1, iters, rel_res, B.rest, B.rest_bytes, stream,
                       pc ? nx_l : 0, pc ? ny_l : 0);
}

}  // extern "C"
