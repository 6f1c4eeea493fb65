// mc_common.hpp -- sample generation and MIS contribution shared by the
// direct and the binned estimator (device code), and the kernel-argument
// block built from vpmc_mc_desc.
#pragma once

#include "field_query.hpp"

namespace vpmc {

constexpr int KPER_TERMS_F32 = 16;
constexpr int MAX_COL_EXT = 129;   // columns of the binned neighbour stencil

struct SetD {
    int n;
    const int *count_dev;
    const double *cdf;
    const double *total_dev;
    const int *guide;
    int glog2;
    const float2 *set_xy;
    FieldD field;
    const float4 *set_jac;    // deformation F by local index, or NULL
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
    int pbase;                // global index of owner 0 (Philox counter, replay stream)
    int vel_stride;           // owners per sign row of vel_sign / vel_aux
    SetD set[2];
    float2 *vel_sign;
    unsigned long long *eval_count;
    int nterms;
    float kc[KPER_TERMS_F32];
    int aux;                  // 1: four auxiliary targets (N2 disk samples each)
    float aux_dx;
    float2 *vel_aux;
};

// offset of auxiliary target t (1..4: L, R, T, B) from the owner
__device__ __forceinline__ float2 aux_offset(int t, float dx)
{
    return make_float2(t == 1 ? -dx : (t == 2 ? dx : 0.f), t == 3 ? dx : (t == 4 ? -dx : 0.f));
}

// samples per (owner, sign): N1 + N2 (1 + 4 aux)
__device__ __host__ __forceinline__ int samples_per_pair(int N1, int N2, int aux)
{
    return N1 + N2 * (aux ? 5 : 1);
}

int make_mcd(const vpmc_mc_desc *in, McD *d);

// Inverse of the cubic spline's radial mass fraction, FP32: four Newton
// steps from an upper bound of the root, unrolled (no data-dependent loop).
// Inner piece q^2/2 - 1.5 q^4 + 1.2 q^5 = 7u/80 on [0, 1/2] from sqrt(2a);
// outer piece, in t = 1-q, t^4/2 - 0.4 t^5 = 7(1-u)/80 on [0, 1/2] from
// (2b)^(1/4).  Both functions are increasing there, and the dropped
// higher-order terms are negative, so the start lies at or right of the
// root.  Both pieces are Newton on c2 y^2 + c4 y^4 + c5 y^5 = rhs with the
// piece's coefficients chosen per lane, so a wave whose draws fall in both
// pieces runs one sequence of steps instead of both.  Agrees with the
// oracle's FP64 bracketed Newton to 2.3e-7 relative.
__device__ __forceinline__ float mass_inverse(float u)
{
    const bool inner = u <= 11.f / 14.f;
    const float rhs = (7.f / 80.f) * (inner ? u : 1.f - u);
    const float c2 = inner ? 0.5f : 0.f, c4 = inner ? -1.5f : 0.5f, c5 = inner ? 1.2f : -0.4f;
    const float e2 = inner ? 1.f : 0.f, e4 = inner ? -6.f : 2.f, e5 = inner ? 6.f : -2.f;
    const float s = __builtin_amdgcn_sqrtf(2.f * rhs);
    float y = fminf(inner ? s : __builtin_amdgcn_sqrtf(s), 0.5f);
#pragma unroll
    for (int it = 0; it < 4; ++it) {
        const float y2 = y * y;
        const float g = fmaf(y2, fmaf(y2, fmaf(c5, y, c4), c2), -rhs);
        const float d = y * fmaf(y2, fmaf(e5, y, e4), e2);   // dg/dy
        const float yn = fminf(fmaxf(y - g * __builtin_amdgcn_rcpf(d), 0.f), 0.5f);
        y = d > 0.f ? yn : y;
    }
    const float x = inner ? y : 1.f - y;
    return u > 0.f ? (u >= 1.f ? 1.f : x) : 0.f;
}

// sin and cos of 2 pi u for a draw u in [0, 1]: v_sin_f32 and v_cos_f32 take
// their argument in revolutions (max abs error 2e-7 over every 24-bit draw
// against FP64, as sincosf(2 pi u))
__device__ __forceinline__ void sincos_rev(float u, float &s, float &c)
{
    s = __builtin_amdgcn_sinf(u);
    c = __builtin_amdgcn_cosf(u);
}

__device__ __forceinline__ void draws(const McD &d, int owner, int si, int branch, int k,
                                      float u[3])
{
    const uint32_t og = (uint32_t)(d.pbase + owner);   // global owner index
    if (d.U) {
        size_t per = (size_t)d.N1 * d.b1 + (size_t)d.N2 * d.b2;
        const float *base = d.U + ((size_t)og * 2 + si) * per;
        if (branch == 1) {
            const float *q = base + (size_t)k * d.b1;
            u[0] = q[0];
            u[1] = d.b1 > 1 ? q[1] : 0.f;
            u[2] = d.b1 > 2 ? q[2] : 0.f;
        } else {   // branch 2 + t for the disk around target t: draws 2t, 2t + 1
            const float *q = base + (size_t)d.N1 * d.b1 + (size_t)k * d.b2 + 2 * (branch - 2);
            u[0] = q[0];
            u[1] = q[1];
            u[2] = 0.f;
        }
    } else {
        uint4 c = make_uint4((uint32_t)k, og, d.step,
                             (d.sim << 8) | ((uint32_t)si << 7) | ((uint32_t)branch << 4));
        uint4 x = philox10(c, make_uint2(d.key0, d.key1));
        u[0] = u01(x.x);
        u[1] = u01(x.y);
        u[2] = u01(x.z);
    }
}

// smooth part of the doubly periodic kernel on the square lattice L(Z+iZ):
// u - i v = (i / (2 pi L)) (sum_m c_m s^(4m-1) + pi conj(s)),  s = r / L
__device__ __forceinline__ void kper_smooth(const McD &d, float rx, float ry, float &kx,
                                            float &ky)
{
    float sr = rx * d.inv_L, si = ry * d.inv_L;
    float s2r = sr * sr - si * si, s2i = 2.f * sr * si;
    float wr = s2r * s2r - s2i * s2i, wi = 2.f * s2r * s2i;
    // KPER_TERMS_F32 terms (make_mcd sets nterms to it), unrolled so the
    // coefficients load once, ahead of the chain
    float hr = d.kc[KPER_TERMS_F32 - 1], hi = 0.f;
#pragma unroll
    for (int m = KPER_TERMS_F32 - 2; m >= 0; --m) {
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

// One sample's generation in three phases, so that a thread generating two
// samples can issue both samples' dependent loads together (gen_two):
//   gen_begin   draws (Philox or the replay stream), the target, and for a
//               branch-1 sample the CDF search bracket from the guide table;
//   the search  std::upper_bound inside the bracket (gen_search);
//   gen_finish  the query point and the MIS displacement from the drawn
//               particle's position (and F).
struct GenState {
    float u[3];
    float2 dst;   // the target: the owner or an auxiliary target
    bool b1;      // branch 1 (a particle drawn from |w|)
    int lo, hi;   // branch 1: the search bracket, then lo = the drawn index
};

template <int MODE, bool PER>
__device__ __forceinline__ void gen_begin(const McD &d, const SetD &S, int owner, int si, int k,
                                          float2 dst, GenState &st)
{
    const bool b1 = k < d.N1;
    int t = 0, kb = b1 ? k : k - d.N1;   // t: target of a disk sample (0 = owner)
    if (!b1 && kb >= d.N2) {
        t = 1 + (kb - d.N2) / d.N2;
        kb = (kb - d.N2) - (t - 1) * d.N2;
    }
    draws(d, owner, si, b1 ? 1 : 2 + t, kb, st.u);
    if (t) {   // the auxiliary target, as the reference forms it (dst.x - dx, ...)
        const float2 o = aux_offset(t, d.aux_dx);
        dst = make_float2(__fadd_rn(dst.x, o.x), __fadd_rn(dst.y, o.y));
    }
    st.dst = dst;
    st.b1 = b1;
    st.lo = st.hi = 0;
    if (b1) {
        const int g = guide_bucket(st.u[0], S.glog2);
        st.lo = S.guide[g];
        st.hi = S.guide[g + 1];
    }
}

// upper_bound of the branch-1 draw inside its bracket (guided_upper_bound)
__device__ __forceinline__ int gen_search(const SetD &S, int n_s, const GenState &st)
{
    int lo = st.lo, hi = st.hi;
    const double ud = (double)st.u[0];
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (ud < S.cdf[mid]) hi = mid; else lo = mid + 1;
    }
    return lo >= n_s ? n_s - 1 : lo;
}

// returns false for a sample that contributes nothing without a query (a
// branch-2 point outside the min-image cell, or a non-finite point).
// (qx, qy): the query point inside the box; (rx, ry): the displacement the
// MIS contribution uses (reference: dst - sp unwrapped; consistent: min image).
// j, xj, F: a branch-1 sample's drawn set-local index, its position and
// (deformed sets) its F.  src (optional) receives j, -1 for a branch-2
// sample; jit (optional) the consistent branch-1 jitter before F is applied
// (the frozen offset the adjoint differentiates F through), (0, 0) otherwise.
template <int MODE, bool PER>
__device__ __forceinline__ bool gen_finish(const McD &d, const GenState &st, int j, float2 xj,
                                           float4 F, bool def, float &qx, float &qy, float &rx,
                                           float &ry, int *src, float2 *jit)
{
    if (jit) *jit = make_float2(0.f, 0.f);
    const float *u = st.u;
    const float2 dst = st.dst;
    const bool b1 = st.b1;
    float spx, spy;
    if (MODE == VPMC_MODE_REFERENCE) {
        if (b1) {   // src/cpu/monte_carlo_bs.cpp:44-58
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
        rx = dst.x - spx;
        ry = dst.y - spy;
    } else {
        if (b1) {   // particle from |w|/Omega, jitter from W (exact radial inverse)
            if (src) *src = j;
            float rho = mass_inverse(u[1]) * d.h_w;
            float s, c;
            sincos_rev(u[2], s, c);
            float jx = rho * c, jy = rho * s;
            if (jit) *jit = make_float2(jx, jy);
            if (def) {   // the deformed blob: F_j u (src/cuda/monte_carlo_bs.cu:210-212)
                const float ex = fmaf(F.x, jx, F.y * jy), ey = fmaf(F.z, jx, F.w * jy);
                jx = ex;
                jy = ey;
            }
            spx = xj.x + jx;
            spy = xj.y + jy;
            rx = dst.x - spx;
            ry = dst.y - spy;
            if (PER) {
                if (rx > 0.5f * d.Lx) rx -= d.Lx; else if (rx < -0.5f * d.Lx) rx += d.Lx;
                if (ry > 0.5f * d.Ly) ry -= d.Ly; else if (ry < -0.5f * d.Ly) ry += d.Ly;
            }
        } else {   // disk radius uniform in [0, Rc): density alpha / r
            if (src) *src = -1;
            float rho = d.Rc * u[1];
            float s, c;
            sincos_rev(u[0], s, c);
            float cx = rho * c, cy = rho * s;
            spx = dst.x + cx;
            spy = dst.y + cy;
            rx = -cx;
            ry = -cy;
            if (PER && (fabsf(rx) > 0.5f * d.Lx || fabsf(ry) > 0.5f * d.Ly)) return false;
        }
    }
    if (!(isfinite(spx) && isfinite(spy))) return false;
    if (PER) {
        if (d.Lx == 1.f && d.Ly == 1.f) {   // the unit box: no division
            qx = spx - floorf(spx);
            qy = spy - floorf(spy);
        } else {
            qx = spx - floorf(spx / d.Lx) * d.Lx;
            qy = spy - floorf(spy / d.Ly) * d.Ly;
        }
    } else {
        if (spx < 0.f || spx > d.Lx || spy < 0.f || spy > d.Ly) return false;
        qx = spx;
        qy = spy;
    }
    return true;
}

// One sample of owner `owner`, sign set si, index k in [0, N1+N2) (the three
// phases above in turn; outputs as gen_finish's)
template <int MODE, bool PER>
__device__ __forceinline__ bool gen_sample(const McD &d, const SetD &S, int n_s, int owner,
                                           int si, int k, float2 dst, float &qx, float &qy,
                                           float &rx, float &ry, int *src = nullptr,
                                           float2 *jit = nullptr)
{
    GenState st;
    gen_begin<MODE, PER>(d, S, owner, si, k, dst, st);
    int j = 0;
    float2 xj = make_float2(0.f, 0.f);
    float4 F = make_float4(1.f, 0.f, 0.f, 1.f);
    if (st.b1) {
        j = gen_search(S, n_s, st);
        xj = S.set_xy[j];
        if (S.set_jac) F = S.set_jac[j];
    }
    return gen_finish<MODE, PER>(d, st, j, xj, F, S.set_jac != nullptr, qx, qy, rx, ry, src,
                                 jit);
}

// Two samples of one thread: the draws and every dependent load of both
// (guide brackets, the interleaved upper_bound searches, positions and F)
// issued together, so that a wave keeps two dependent-load chains in flight
// where one sample would keep one.  gen_finish then runs once per sample
// from the outputs (the caller loops over them without unrolling, so that
// both samples go through one copy of its arithmetic).
struct GenTwo {
    GenState st[2];
    int j[2];
    float2 xj[2];
    float4 F[2];
};

template <int MODE, bool PER>
__device__ __forceinline__ void gen_two(const McD &d, const SetD &Sa, const SetD &Sb, int n_a,
                                        int n_b, int oa, int ob, int sa, int sb, int ka, int kb,
                                        float2 da, float2 db, GenTwo &g2)
{
    gen_begin<MODE, PER>(d, Sa, oa, sa, ka, da, g2.st[0]);
    gen_begin<MODE, PER>(d, Sb, ob, sb, kb, db, g2.st[1]);
    int lo[2] = {g2.st[0].lo, g2.st[1].lo}, hi[2] = {g2.st[0].hi, g2.st[1].hi};
    const double ud[2] = {(double)g2.st[0].u[0], (double)g2.st[1].u[0]};
    const SetD *S[2] = {&Sa, &Sb};
    while (lo[0] < hi[0] || lo[1] < hi[1]) {   // both upper_bound searches, interleaved
        double c[2];
        int mid[2];
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            mid[q] = (lo[q] + hi[q]) >> 1;
            c[q] = lo[q] < hi[q] ? S[q]->cdf[mid[q]] : 0.0;
        }
#pragma unroll
        for (int q = 0; q < 2; ++q)
            if (lo[q] < hi[q]) {
                if (ud[q] < c[q]) hi[q] = mid[q]; else lo[q] = mid[q] + 1;
            }
    }
    const int n[2] = {n_a, n_b};
#pragma unroll
    for (int q = 0; q < 2; ++q) {
        g2.j[q] = lo[q] >= n[q] ? n[q] - 1 : lo[q];
        g2.xj[q] = make_float2(0.f, 0.f);
        g2.F[q] = make_float4(1.f, 0.f, 0.f, 1.f);
        if (g2.st[q].b1 && n[q] > 0) {
            g2.xj[q] = S[q]->set_xy[g2.j[q]];
            if (S[q]->set_jac) g2.F[q] = S[q]->set_jac[g2.j[q]];
        }
    }
}

// MIS contribution of one sample with vorticity om at displacement r
template <int MODE, bool PER>
__device__ __forceinline__ void mis_contrib(const McD &d, float om, float rx, float ry, float Om,
                                            float cum, float &sx, float &sy)
{
    if (!(om > 0.f)) return;
    if (MODE == VPMC_MODE_REFERENCE) {
        // mis_contrib, src/cpu/monte_carlo_bs.cpp:11-26
        float r2 = rx * rx + ry * ry;
        float rn = fmaxf(sqrtf(r2), 1e-6f);
        float inv_r2 = 1.f / (r2 + 1e-8f);
        float kx = -ry * inv_r2 / (2.f * PI_F);
        float ky = rx * inv_r2 / (2.f * PI_F);
        float denom = d.w1 * om / cum + d.w2 * d.alpha_ref / rn;
        float g = om / denom;
        sx += kx * g;
        sy += ky * g;
    } else {
        float rn = sqrtf(rx * rx + ry * ry);
        if (!(rn > 0.f)) return;
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

// update_jacobian_from5 + renorm_det1 (src/cuda/monte_carlo_bs.cu:274-288,
// include/monte_carlo_bs.cuh:140-144) for owner j of n: vs/va are the sign
// estimates ([sign][owner]) and the aux ones ([sign][L, R, T, B][owner]);
// c = dt / (2 dx)
__device__ __forceinline__ float4 jacobian_from5(int n, int j, const float2 *__restrict__ va,
                                                 float c)
{
    float2 u[4];   // L, R, T, B: v+ - v-
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        const float2 p = va[(size_t)t * n + j], q = va[(size_t)(4 + t) * n + j];
        u[t] = make_float2(p.x - q.x, p.y - q.y);
    }
    float4 f;   // explicit fmaf: the same rounding in every kernel inlining this
    f.x = fmaf(u[1].x - u[0].x, c, 1.f);
    f.y = (u[2].x - u[3].x) * c;
    f.z = (u[1].y - u[0].y) * c;
    f.w = fmaf(u[2].y - u[3].y, c, 1.f);
    const float det = fmaf(f.x, f.w, -(f.y * f.z));
    const float sc = 1.f / sqrtf(fmaxf(det, 1e-12f));   // det F -> 1 (incompressible)
    return make_float4(f.x * sc, f.y * sc, f.z * sc, f.w * sc);
}

// u = v+ - v-, x += u dt, periodic wrap (src/cpu/simulation_step.cpp:46-71)
__device__ __forceinline__ float2 advect_one(float2 a, float2 b, float2 p, float2 *vel_out,
                                             float dt, float Lx, float Ly, int per)
{
    float2 v = make_float2(__fsub_rn(a.x, b.x), __fsub_rn(a.y, b.y));
    if (vel_out) *vel_out = v;
    p.x = __fadd_rn(p.x, __fmul_rn(v.x, dt));
    p.y = __fadd_rn(p.y, __fmul_rn(v.y, dt));
    if (per) {   // src/cpu/simulation_step.cpp:57-59
        p.x = __fsub_rn(p.x, __fmul_rn(floorf(p.x / Lx), Lx));
        p.y = __fsub_rn(p.y, __fmul_rn(floorf(p.y / Ly), Ly));
    }
    return p;
}

}  // namespace vpmc
