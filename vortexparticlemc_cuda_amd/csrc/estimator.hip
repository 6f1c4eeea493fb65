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

#include "mc_common.hpp"

namespace vpmc {

constexpr int MC_THREADS = 256;

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
            float qx, qy, rx, ry;
            if (!gen_sample<MODE, PER>(d, S, n_s, owner, si, k, dst, qx, qy, rx, ry)) continue;
            float om = d.sig2 * field_sum<COUNT, PER>(S.field, qx, qy, d.inv_hw, ec);
            mis_contrib<MODE, PER>(d, om, rx, ry, Om, cum, sx, sy);
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
        d.vel_sign[(size_t)si * d.vel_stride + owner] = make_float2(tx * d.inv_tot, ty * d.inv_tot);
        if (COUNT) {
            atomicAdd(d.eval_count, a);
            atomicAdd(d.eval_count + 1, b);
        }
    }
}

// update_jacobian_from5 + renorm_det1 (src/cuda/monte_carlo_bs.cu:274-288,
// include/monte_carlo_bs.cuh:140-144)
__global__ void k_update_jacobian(int n, const float2 *__restrict__ va, float c,
                                  float4 *__restrict__ F)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    F[i] = jacobian_from5(n, i, va, c);
}

__global__ void k_advect(int n, float2 *__restrict__ pos, const float2 *__restrict__ vs,
                         float2 *__restrict__ vel, float dt, float Lx, float Ly, int per)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    pos[i] = advect_one(vs[i], vs[(size_t)n + i], pos[i], vel ? vel + i : nullptr, dt, Lx, Ly, per);
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

// the first 24 coefficients in FP64 (FP64 forward of the frozen-sample estimator)
const double *kper_coeffs_f64_host(int nterms)
{
    static double kc[24];
    static std::once_flag once;
    std::call_once(once, [] { kper_coeffs_host(24, kc); });
    return nterms <= 24 ? kc : nullptr;
}

int make_mcd(const vpmc_mc_desc *in, McD *d)
{
    if (!in || in->N1 < 0 || in->N2 < 0 || in->N1 + in->N2 <= 0 || in->n_owner < 0 ||
        !(in->h_w > 0.f) || !(in->Lx > 0.f) || !(in->Ly > 0.f))
        return VPMC_ERR_ARG;
    if (in->mode != VPMC_MODE_REFERENCE && in->mode != VPMC_MODE_CONSISTENT) return VPMC_ERR_ARG;
    if (in->mode == VPMC_MODE_CONSISTENT && in->periodic && in->Lx != in->Ly)
        return VPMC_ERR_ARG;   // the periodic kernel is written for the square lattice
    if (in->n_owner > 0 && (!in->owner_xy || !in->vel_sign)) return VPMC_ERR_ARG;
    if (in->owner_base < 0 || (int64_t)in->owner_base + in->n_owner > INT32_MAX)
        return VPMC_ERR_ARG;
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
    m.pbase = in->owner_base;
    m.vel_stride = in->n_owner;
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
        D.field.gq = (const float4 *)S.field.gq;
        D.set_jac = (const float4 *)S.set_jac;
        if ((S.field.gq != nullptr) != (in->set[0].field.gq != nullptr)) return VPMC_ERR_ARG;
        if (!D.field.cell_start) return VPMC_ERR_ARG;
    }
    m.vel_sign = (float2 *)in->vel_sign;
    m.eval_count = in->eval_count;
    m.aux = in->aux ? 1 : 0;
    m.aux_dx = in->aux_dx;
    m.vel_aux = (float2 *)in->vel_aux;
    if (m.aux) {
        if (!(in->aux_dx > 0.f) || (in->n_owner > 0 && !in->vel_aux)) return VPMC_ERR_ARG;
        if (in->uniforms && !in->aux_draws) return VPMC_ERR_ARG;   // no aux draws to replay
        if ((size_t)in->N1 + 5 * (size_t)in->N2 > (size_t)1 << 30) return VPMC_ERR_ARG;
    }
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
    if (m.aux) return VPMC_ERR_ARG;   // auxiliary targets: the binned algorithm
    if (m.n_owner == 0) return VPMC_OK;
    VPMC_TRY(record_event(desc->eval_events[0], s));
    if (desc->eval_count) launch_mc<true>(m, desc->mode, desc->periodic, s);
    else launch_mc<false>(m, desc->mode, desc->periodic, s);
    VPMC_CHECK_LAUNCH();
    VPMC_TRY(record_event(desc->eval_events[1], s));
    return VPMC_OK;
}

}  // namespace vpmc

using namespace vpmc;

extern "C" {

int vpmc_mc_estimate(const vpmc_mc_desc *d, void *stream)
{
    return mc_estimate_impl(d, as_stream(stream));
}

int vpmc_update_jacobian(int n, const float *vel_sign, const float *vel_aux, float dt, float dx,
                         float *jac_out, void *stream)
{
    if (n < 0 || !(dx > 0.f)) return VPMC_ERR_ARG;
    if (n == 0) return VPMC_OK;
    if (!vel_aux || !jac_out) return VPMC_ERR_ARG;
    hipLaunchKernelGGL(k_update_jacobian, dim3((n + 255) / 256), dim3(256), 0, as_stream(stream),
                       n, (const float2 *)vel_aux, 0.5f * dt / dx, (float4 *)jac_out);
    VPMC_CHECK_LAUNCH();
    return VPMC_OK;
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

size_t vpmc_step_workspace_bytes(const vpmc_mc_desc *d)
{
    if (!d) return 0;
    int m = d->set[0].n > d->set[1].n ? d->set[0].n : d->set[1].n;
    size_t b = vpmc_field_workspace_bytes(m, d->set[0].field.nx, d->set[0].field.ny);
    if (d->algo == VPMC_ALGO_BINNED) {   // field build and estimate use it in turn
        size_t e = vpmc_mc_binned_workspace_bytes(d);
        if (e > b) b = e;
    }
    return b;
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
        // with a deformation: the fields' quadratic forms and the sets' F
        // (src/cuda/monte_carlo_bs.cu:210-212, include/vorticity_query.cuh:52-66)
        if (d->jac && (!S.field.gq || !S.set_jac)) return VPMC_ERR_ARG;
        if (!d->jac && (S.field.gq || S.set_jac)) return VPMC_ERR_ARG;
        VPMC_TRY(field_build_impl(S.field.nx, S.field.ny, S.field.Lx, S.field.Ly, S.n, nullptr,
                                  subs[k], pos_xy, omega, 1, const_cast<float *>(S.field.pts),
                                  const_cast<int *>(S.field.cell_start),
                                  const_cast<int *>(S.field.order),
                                  const_cast<float *>(S.set_xy), workspace, workspace_bytes, s,
                                  d->jac, const_cast<float *>(S.field.gq),
                                  const_cast<float *>(S.set_jac)));
    }
    if (d->algo == VPMC_ALGO_BINNED)
        VPMC_TRY(vpmc_mc_estimate_binned(d, workspace, workspace_bytes, stream));
    else
        VPMC_TRY(mc_estimate_impl(d, s));
    // step_cuda: F from this step's five estimates, before the move
    // (src/cuda/simulation_step.cu:68-72)
    if (d->jac && d->aux)
        VPMC_TRY(vpmc_update_jacobian(n_total, d->vel_sign, d->vel_aux, dt, d->aux_dx, d->jac,
                                      stream));
    return vpmc_advect(n_total, pos_xy, d->vel_sign, vel_xy, dt, d->Lx, d->Ly, d->periodic,
                       stream);
}

}  // extern "C"
