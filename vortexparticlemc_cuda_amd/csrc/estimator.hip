// estimator.hip -- two-branch MIS Monte Carlo Biot-Savart estimator,
// advection and the plain time step.
//
// One workgroup (4 wave64) per (owner, sign); each lane takes samples
// k = lane, lane+256, ... of that owner, generates the sample point from a
// counter-based Philox stream (or injected uniforms), evaluates the sign's
// vorticity there with the fine-grid field sum, forms the MIS contribution
// and keeps a private running sum.  The workgroup reduces in a fixed order
// (wave butterflies, then the four wave totals in order), so the result does
// not depend on scheduling.
#include <math.h>

#include <mutex>

#include "field_query.hpp"

namespace vpmc {

constexpr int MC_THREADS = 256;
constexpr int KPER_TERMS_F32 = 16;

struct SetD {
    int n;
    const int *count_dev;
    const double *cdf;
    const double *total_dev;
    const int *guide;
    int glog2;
    const float2 *set_xy;
    FieldD field;
};

struct McD {
    int N1, N2;
    float h_w, Lx, Ly, inv_hw, sig2;
    float w1, w2, inv_tot;
    float R_ref, alpha_ref;   // reference disk: sqrt(Lx^2 + Ly^2)
    float Rc, alpha_c;        // consistent disk: circumradius of the min-image cell
    float inv_L, inv_2piL;
    uint32_t key0, key1, step, sim;
    const float *U;
    int b1, b2;
    int n_owner;
    const float2 *owner;
    SetD set[2];
    float2 *vel_sign;
    unsigned long long *eval_count;
    int nterms;
    float kc[KPER_TERMS_F32];
};

// Inverse of the cubic spline's radial mass fraction (FP64 bracketed
// Newton); the oracle runs the same iteration.  Inner piece solves
// q^2/2 - 1.5 q^4 + 1.2 q^5 = 7u/80 on [0, 1/2]; outer piece, in t = 1-q,
// t^4/2 - 0.4 t^5 = 7(1-u)/80 on [0, 1/2].
__device__ double mass_inverse(double u)
{
    if (u <= 0.0) return 0.0;
    if (u >= 1.0) return 1.0;
    const double u_half = 11.0 / 14.0;
    if (u <= u_half) {
        double a = 7.0 * u / 80.0, lo = 0.0, hi = 0.5;
        double x = sqrt(2.0 * a);
        if (x > hi) x = hi;
        for (int it = 0; it < 40; ++it) {
            double x2 = x * x;
            double f = 0.5 * x2 - 1.5 * x2 * x2 + 1.2 * x2 * x2 * x - a;
            if (f > 0.0) hi = x; else lo = x;
            double d = x - 6.0 * x2 * x + 6.0 * x2 * x2;
            double xn = (d > 0.0) ? x - f / d : 0.5 * (lo + hi);
            if (!(xn >= lo && xn <= hi)) xn = 0.5 * (lo + hi);
            if (xn == x) break;
            x = xn;
        }
        return x;
    }
    double b = 7.0 * (1.0 - u) / 80.0, lo = 0.0, hi = 0.5;
    double t = sqrt(sqrt(2.0 * b));
    if (t > hi) t = hi;
    for (int it = 0; it < 40; ++it) {
        double t2 = t * t;
        double f = 0.5 * t2 * t2 - 0.4 * t2 * t2 * t - b;
        if (f > 0.0) hi = t; else lo = t;
        double d = 2.0 * t2 * t - 2.0 * t2 * t2;
        double tn = (d > 0.0) ? t - f / d : 0.5 * (lo + hi);
        if (!(tn >= lo && tn <= hi)) tn = 0.5 * (lo + hi);
        if (tn == t) break;
        t = tn;
    }
    return 1.0 - t;
}

__device__ __forceinline__ void draws(const McD &d, int owner, int si, int branch, int k,
                                      float u[3])
{
    if (d.U) {
        size_t per = (size_t)d.N1 * d.b1 + (size_t)d.N2 * d.b2;
        const float *base = d.U + ((size_t)owner * 2 + si) * per;
        if (branch == 1) {
            const float *q = base + (size_t)k * d.b1;
            u[0] = q[0];
            u[1] = d.b1 > 1 ? q[1] : 0.f;
            u[2] = d.b1 > 2 ? q[2] : 0.f;
        } else {
            const float *q = base + (size_t)d.N1 * d.b1 + (size_t)k * d.b2;
            u[0] = q[0];
            u[1] = q[1];
            u[2] = 0.f;
        }
    } else {
        uint4 c = make_uint4((uint32_t)k, (uint32_t)owner, d.step,
                             (d.sim << 8) | ((uint32_t)si << 7) | ((uint32_t)branch << 4));
        uint4 x = philox10(c, make_uint2(d.key0, d.key1));
        u[0] = u01(x.x);
        u[1] = u01(x.y);
        u[2] = u01(x.z);
    }
}

__device__ __forceinline__ int sample_idx(float u, const SetD &S, int n)
{
    return guided_upper_bound(u, S.cdf, S.guide, S.glog2, n);
}

// smooth part of the doubly periodic kernel on the square lattice L(Z+iZ):
// u - i v = (i / (2 pi L)) (sum_m c_m s^(4m-1) + pi conj(s)),  s = r / L
__device__ __forceinline__ void kper_smooth(const McD &d, float rx, float ry, float &kx,
                                            float &ky)
{
    float sr = rx * d.inv_L, si = ry * d.inv_L;
    float s2r = sr * sr - si * si, s2i = 2.f * sr * si;
    float wr = s2r * s2r - s2i * s2i, wi = 2.f * s2r * s2i;
    float hr = d.kc[d.nterms - 1], hi = 0.f;
    for (int m = d.nterms - 2; m >= 0; --m) {
        float t = hr * wr - hi * wi + d.kc[m];
        hi = hr * wi + hi * wr;
        hr = t;
    }
    float s3r = s2r * sr - s2i * si, s3i = s2r * si + s2i * sr;
    float Ar = hr * s3r - hi * s3i + PI_F * sr;
    float Ai = hr * s3i + hi * s3r - PI_F * si;
    kx = -Ai * d.inv_2piL;
    ky = -Ar * d.inv_2piL;
}

template <bool COUNT, int MODE, bool PER>
__global__ void __launch_bounds__(MC_THREADS) k_mc_estimate(McD d)
{
    const int owner = blockIdx.x, si = blockIdx.y;
    const SetD &S = d.set[si];
    const float2 dst = d.owner[owner];
    const int n_s = S.count_dev ? *S.count_dev : S.n;
    float sx = 0.f, sy = 0.f;
    EvalCount ec;
    if (n_s > 0) {
        const float Om = (float)(*S.total_dev);
        float cum = Om;
        if (MODE == VPMC_MODE_REFERENCE && PER) cum *= 9.f;
        const int ns = d.N1 + d.N2;
        for (int k = threadIdx.x; k < ns; k += MC_THREADS) {
            const bool b1 = k < d.N1;
            float u[3];
            draws(d, owner, si, b1 ? 1 : 2, b1 ? k : k - d.N1, u);
            float spx, spy, rx = 0.f, ry = 0.f;
            bool live = true;
            if (MODE == VPMC_MODE_REFERENCE) {
                if (b1) {
                    int j = sample_idx(u[0], S, n_s);
                    float2 xj = S.set_xy[j];
                    spx = xj.x;
                    spy = xj.y;
                    if (PER) {   // include/monte_carlo_bs.hpp:92-102
                        float ddx = 0.f, ddy = 0.f;
                        if (u[1] < 1.f / 3.f) ddx = -1.f; else if (u[1] > 2.f / 3.f) ddx = 1.f;
                        if (u[2] < 1.f / 3.f) ddy = -1.f; else if (u[2] > 2.f / 3.f) ddy = 1.f;
                        spx = __fadd_rn(spx, __fmul_rn(ddx, d.Lx));
                        spy = __fadd_rn(spy, __fmul_rn(ddy, d.Ly));
                    }
                } else {   // src/cpu/monte_carlo_sampling.cpp:35-45
                    float phi = __fmul_rn(2.f * PI_F, u[0]);
                    float r = __fmul_rn(d.R_ref, sqrtf(u[1]));
                    float s, c;
                    sincosf(phi, &s, &c);
                    spx = __fadd_rn(dst.x, __fmul_rn(r, c));
                    spy = __fadd_rn(dst.y, __fmul_rn(r, s));
                }
            } else {
                if (b1) {
                    int j = sample_idx(u[0], S, n_s);
                    float2 xj = S.set_xy[j];
                    float rho = (float)mass_inverse((double)u[1]) * d.h_w;
                    float s, c;
                    sincosf(2.f * PI_F * u[2], &s, &c);
                    spx = xj.x + rho * c;
                    spy = xj.y + rho * s;
                    rx = dst.x - spx;
                    ry = dst.y - spy;
                    if (PER) {
                        if (rx > 0.5f * d.Lx) rx -= d.Lx; else if (rx < -0.5f * d.Lx) rx += d.Lx;
                        if (ry > 0.5f * d.Ly) ry -= d.Ly; else if (ry < -0.5f * d.Ly) ry += d.Ly;
                    }
                } else {
                    float rho = d.Rc * u[1];
                    float s, c;
                    sincosf(2.f * PI_F * u[0], &s, &c);
                    float cx = rho * c, cy = rho * s;
                    spx = dst.x + cx;
                    spy = dst.y + cy;
                    rx = -cx;
                    ry = -cy;
                    if (PER && (fabsf(rx) > 0.5f * d.Lx || fabsf(ry) > 0.5f * d.Ly)) live = false;
                }
            }
            if (!live) continue;
            float px = spx, py = spy;
            if (PER) {
                px = px - floorf(px / d.Lx) * d.Lx;
                py = py - floorf(py / d.Ly) * d.Ly;
            } else if (px < 0.f || px > d.Lx || py < 0.f || py > d.Ly) {
                continue;
            }
            float om = d.sig2 * field_sum<COUNT, PER>(S.field, px, py, d.inv_hw, ec);
            if (!(om > 0.f)) continue;
            if (MODE == VPMC_MODE_REFERENCE) {
                // mis_contrib, src/cpu/monte_carlo_bs.cpp:11-26
                float ax = dst.x - spx, ay = dst.y - spy;
                float r2 = ax * ax + ay * ay;
                float rn = fmaxf(sqrtf(r2), 1e-6f);
                float inv_r2 = 1.f / (r2 + 1e-8f);
                float kx = -ay * inv_r2 / (2.f * PI_F);
                float ky = ax * inv_r2 / (2.f * PI_F);
                float denom = d.w1 * om / cum + d.w2 * d.alpha_ref / rn;
                float g = om / denom;
                sx += kx * g;
                sy += ky * g;
            } else {
                float rn = sqrtf(rx * rx + ry * ry);
                if (!(rn > 0.f)) continue;
                // omega / (w1 p1 + w2 p2) with p1 = omega/Omega, p2 = alpha/r
                float g = om * rn / (d.w1 * om * rn / Om + d.w2 * d.alpha_c);
                float f = g / (2.f * PI_F * rn * rn);
                float kx = -ry * f, ky = rx * f;
                if (PER) {
                    float ksx, ksy;
                    kper_smooth(d, rx, ry, ksx, ksy);
                    kx += ksx * g;
                    ky += ksy * g;
                }
                sx += kx;
                sy += ky;
            }
        }
    }
    // fixed-order workgroup reduction
    __shared__ float red[2][MC_THREADS / WAVE];
    __shared__ unsigned redc[2][MC_THREADS / WAVE];
    const int lane = threadIdx.x & (WAVE - 1), wid = threadIdx.x / WAVE;
    sx = wave_sum(sx);
    sy = wave_sum(sy);
    unsigned ci = 0, cv = 0;
    if (COUNT) {
        ci = wave_sum(ec.in_support);
        cv = wave_sum(ec.visited);
    }
    if (lane == 0) {
        red[0][wid] = sx;
        red[1][wid] = sy;
        if (COUNT) { redc[0][wid] = ci; redc[1][wid] = cv; }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        float tx = red[0][0], ty = red[1][0];
        unsigned long long a = redc[0][0], b = redc[1][0];
#pragma unroll
        for (int w = 1; w < MC_THREADS / WAVE; ++w) {
            tx += red[0][w];
            ty += red[1][w];
            if (COUNT) { a += redc[0][w]; b += redc[1][w]; }
        }
        d.vel_sign[(size_t)si * d.n_owner + owner] = make_float2(tx * d.inv_tot, ty * d.inv_tot);
        if (COUNT) {
            atomicAdd(d.eval_count, a);
            atomicAdd(d.eval_count + 1, b);
        }
    }
}

__global__ void k_advect(int n, float2 *__restrict__ pos, const float2 *__restrict__ vs,
                         float2 *__restrict__ vel, float dt, float Lx, float Ly, int per)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float2 a = vs[i], b = vs[(size_t)n + i];
    float2 v = make_float2(__fsub_rn(a.x, b.x), __fsub_rn(a.y, b.y));
    if (vel) vel[i] = v;
    float2 p = pos[i];
    p.x = __fadd_rn(p.x, __fmul_rn(v.x, dt));
    p.y = __fadd_rn(p.y, __fmul_rn(v.y, dt));
    if (per) {   // src/cpu/simulation_step.cpp:57-59
        p.x = __fsub_rn(p.x, __fmul_rn(floorf(p.x / Lx), Lx));
        p.y = __fsub_rn(p.y, __fmul_rn(floorf(p.y / Ly), Ly));
    }
    pos[i] = p;
}

// c_m = G_{4m}(Z+iZ): G_4 from the Eisenstein q-series, the rest from the
// absolutely convergent lattice sum over one rotation orbit representative.
static void kper_coeffs_host(int nterms, double *c)
{
    double q = exp(-2.0 * PI_D), qn = 1.0, s = 0.0;
    for (int n = 1; n <= 12; ++n) {
        qn *= q;
        double sig = 0.0;
        for (int dd = 1; dd <= n; ++dd)
            if (n % dd == 0) sig += (double)dd * dd * dd;
        s += sig * qn;
    }
    double pi4 = PI_D * PI_D * PI_D * PI_D;
    c[0] = (pi4 / 45.0) * (1.0 + 240.0 * s);
    for (int m = 2; m <= nterms; ++m) {
        int R = (m == 2) ? 600 : 160;
        double acc = 0.0;
        for (int a = 1; a <= R; ++a)
            for (int b = 0; b <= R; ++b) {
                double re = a, im = -(double)b;
                double z2r = re * re - im * im, z2i = 2 * re * im;
                double z4r = z2r * z2r - z2i * z2i, z4i = 2 * z2r * z2i;
                double pr = 1.0, pim = 0.0;
                for (int k = 0; k < m; ++k) {
                    double t = pr * z4r - pim * z4i;
                    pim = pr * z4i + pim * z4r;
                    pr = t;
                }
                double mod2 = (double)a * a + (double)b * b;
                acc += pr / pow(mod2, 4.0 * m);
            }
        c[m - 1] = 4.0 * acc;
    }
}

static const float *kper_coeffs_f32()
{
    static float kc[KPER_TERMS_F32];
    static std::once_flag once;
    std::call_once(once, [] {
        double c[KPER_TERMS_F32];
        kper_coeffs_host(KPER_TERMS_F32, c);
        for (int i = 0; i < KPER_TERMS_F32; ++i) kc[i] = (float)c[i];
    });
    return kc;
}

static int make_mcd(const vpmc_mc_desc *in, McD *d)
{
    if (!in || in->N1 < 0 || in->N2 < 0 || in->N1 + in->N2 <= 0 || in->n_owner < 0 ||
        !(in->h_w > 0.f) || !(in->Lx > 0.f) || !(in->Ly > 0.f))
        return VPMC_ERR_ARG;
    if (in->mode != VPMC_MODE_REFERENCE && in->mode != VPMC_MODE_CONSISTENT) return VPMC_ERR_ARG;
    if (in->mode == VPMC_MODE_CONSISTENT && in->periodic && in->Lx != in->Ly)
        return VPMC_ERR_ARG;   // the periodic kernel is written for the square lattice
    if (in->n_owner > 0 && (!in->owner_xy || !in->vel_sign)) return VPMC_ERR_ARG;
    McD &m = *d;
    m.N1 = in->N1;
    m.N2 = in->N2;
    m.h_w = in->h_w;
    m.Lx = in->Lx;
    m.Ly = in->Ly;
    m.inv_hw = 1.f / in->h_w;
    m.sig2 = 40.f / (7.f * PI_F * in->h_w * in->h_w);
    m.w1 = (float)in->N1 / (float)(in->N1 + in->N2);
    m.w2 = (float)in->N2 / (float)(in->N1 + in->N2);
    m.inv_tot = 1.f / (float)(in->N1 + in->N2);
    {   // include/monte_carlo_bs.hpp:73-89 (the largest image offset)
        float maxR2 = 0.f;
        for (int dx = -1; dx <= 1; ++dx)
            for (int dy = -1; dy <= 1; ++dy) {
                float ix = dx * in->Lx, iy = dy * in->Ly;
                float r2 = ix * ix + iy * iy;
                if (r2 > maxR2) maxR2 = r2;
            }
        m.R_ref = sqrtf(maxR2);
        m.alpha_ref = 1.f / (2.f * PI_F * m.R_ref);
    }
    float diag = sqrtf(in->Lx * in->Lx + in->Ly * in->Ly);
    m.Rc = in->periodic ? 0.5f * diag : diag;
    m.alpha_c = 1.f / (2.f * PI_F * m.Rc);
    m.inv_L = 1.f / in->Lx;
    m.inv_2piL = 1.f / (2.f * PI_F * in->Lx);
    m.key0 = (uint32_t)(in->seed & 0xffffffffu);
    m.key1 = (uint32_t)(in->seed >> 32);
    m.step = in->step;
    m.sim = in->sim;
    m.U = in->uniforms;
    m.b1 = (in->mode == VPMC_MODE_REFERENCE && !in->periodic) ? 1 : 3;
    m.b2 = in->aux_draws ? 10 : 2;
    m.n_owner = in->n_owner;
    m.owner = (const float2 *)in->owner_xy;
    for (int s = 0; s < 2; ++s) {
        const vpmc_sign_set &S = in->set[s];
        SetD &D = m.set[s];
        D.n = S.n;
        D.count_dev = S.count_dev;
        D.cdf = S.cdf;
        D.total_dev = S.total_dev;
        D.guide = S.guide;
        D.glog2 = S.guide_log2;
        D.set_xy = (const float2 *)S.set_xy;
        if (S.n > 0 || S.count_dev) {
            if (!S.cdf || !S.total_dev || !S.guide || !S.set_xy || S.guide_log2 < 0 ||
                S.guide_log2 > 30)
                return VPMC_ERR_ARG;
        }
        if (!S.total_dev) return VPMC_ERR_ARG;
        VPMC_TRY(field_geom(&S.field, in->h_w, &D.field.g));
        D.field.pts = (const float4 *)S.field.pts;
        D.field.cell_start = S.field.cell_start;
        if (!D.field.cell_start) return VPMC_ERR_ARG;
    }
    m.vel_sign = (float2 *)in->vel_sign;
    m.eval_count = in->eval_count;
    m.nterms = KPER_TERMS_F32;
    const float *kc = kper_coeffs_f32();
    for (int i = 0; i < KPER_TERMS_F32; ++i) m.kc[i] = kc[i];
    return VPMC_OK;
}

template <bool COUNT>
static void launch_mc(const McD &m, int mode, int per, hipStream_t s)
{
    dim3 grid((unsigned)m.n_owner, 2), block(MC_THREADS);
    if (mode == VPMC_MODE_REFERENCE) {
        if (per) hipLaunchKernelGGL((k_mc_estimate<COUNT, VPMC_MODE_REFERENCE, true>), grid, block, 0, s, m);
        else hipLaunchKernelGGL((k_mc_estimate<COUNT, VPMC_MODE_REFERENCE, false>), grid, block, 0, s, m);
    } else {
        if (per) hipLaunchKernelGGL((k_mc_estimate<COUNT, VPMC_MODE_CONSISTENT, true>), grid, block, 0, s, m);
        else hipLaunchKernelGGL((k_mc_estimate<COUNT, VPMC_MODE_CONSISTENT, false>), grid, block, 0, s, m);
    }
}

int mc_estimate_impl(const vpmc_mc_desc *desc, hipStream_t s)
{
    McD m;
    VPMC_TRY(make_mcd(desc, &m));
    if (m.n_owner == 0) return VPMC_OK;
    if (desc->eval_count) launch_mc<true>(m, desc->mode, desc->periodic, s);
    else launch_mc<false>(m, desc->mode, desc->periodic, s);
    VPMC_CHECK_LAUNCH();
    return VPMC_OK;
}

}  // namespace vpmc

using namespace vpmc;

extern "C" {

int vpmc_mc_estimate(const vpmc_mc_desc *d, void *stream)
{
    return mc_estimate_impl(d, as_stream(stream));
}

int vpmc_advect(int n, float *pos_xy, const float *vel_sign, float *vel_xy, float dt, float Lx,
                float Ly, int periodic, void *stream)
{
    if (n < 0 || !(Lx > 0.f) || !(Ly > 0.f)) return VPMC_ERR_ARG;
    if (n == 0) return VPMC_OK;
    if (!pos_xy || !vel_sign) return VPMC_ERR_ARG;
    hipLaunchKernelGGL(k_advect, dim3((n + 255) / 256), dim3(256), 0, as_stream(stream), n,
                       (float2 *)pos_xy, (const float2 *)vel_sign, (float2 *)vel_xy, dt, Lx, Ly,
                       periodic);
    VPMC_CHECK_LAUNCH();
    return VPMC_OK;
}

size_t vpmc_step_workspace_bytes(int n_total, int n_pos, int n_neg, int nx, int ny)
{
    int m = n_pos > n_neg ? n_pos : n_neg;
    (void)n_total;
    return vpmc_field_workspace_bytes(m, nx, ny);
}

int vpmc_step(vpmc_mc_desc *d, int n_total, float *pos_xy, const float *omega,
              const int *sub_pos, const int *sub_neg, float *vel_xy, float dt, void *workspace,
              size_t workspace_bytes, void *stream)
{
    if (!d || n_total < 0 || d->n_owner != n_total || d->owner_xy != pos_xy) return VPMC_ERR_ARG;
    hipStream_t s = as_stream(stream);
    const int *subs[2] = {sub_pos, sub_neg};
    for (int k = 0; k < 2; ++k) {
        vpmc_sign_set &S = d->set[k];
        VPMC_TRY(field_build_impl(S.field.nx, S.field.ny, S.field.Lx, S.field.Ly, S.n, nullptr,
                                  subs[k], pos_xy, omega, 1, const_cast<float *>(S.field.pts),
                                  const_cast<int *>(S.field.cell_start),
                                  const_cast<int *>(S.field.order),
                                  const_cast<float *>(S.set_xy), workspace, workspace_bytes, s));
    }
    VPMC_TRY(mc_estimate_impl(d, s));
    return vpmc_advect(n_total, pos_xy, d->vel_sign, vel_xy, dt, d->Lx, d->Ly, d->periodic,
                       stream);
}

}  // extern "C"
