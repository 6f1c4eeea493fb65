// adjoint.hip -- reverse-mode derivative of the Monte Carlo velocity and of
// the advection step with respect to particle positions and strengths, under
// common random numbers; the grid vorticity used as a loss, and its adjoint.
//
// The consistent estimator's velocity of owner i from sign set s is
//   v_i = (1/NS) sum_k K(r_k) om(y_k) / D_k,
//   D_k = w1 om0(y_k) / Omega0 + w2 alpha / |r_k|,
// with om(y) = sum_{m in s} a_m W(|y - x_m|), a_m = |w_m|, K the doubly
// periodic kernel, y_k the sample point and r_k the min-image displacement
// x_i - y_k.  Samples are frozen: the Philox draws, the particle a branch-1
// sample was drawn from and the jitter / disk offsets stay fixed, so y_k
// moves with its source x_j (branch 1, y = x_j + jitter) or with the owner
// x_i (branch 2, y = x_i + offset).  D_k is the density the samples were
// drawn from: it follows the positions (om0 is om at the strengths the
// sampling tables were built from) but not the strengths, because the frozen
// indices do not follow them.  With the density held, the estimate remains
// an unbiased estimate of the exact velocity for every strength vector, and
// so does its derivative.  At the nominal strengths om0 = om and the forward
// value is the product estimator's.
//
// Per sample, with a the adjoint of v_i (scaled by 1/NS), s = (a.K) om / D:
//   beta   = ds/d om   = (a.K) / D
//   beta0  = ds/d om0  = -(a.K) om w1 / (Omega0 D^2)
//   ds/dr  = (om/D) grad_r(a.K) + (a.K) om w2 alpha r / (D^2 |r|^3)
//   ds/dy  = (beta + beta0) grad om(y)
// branch 1: x_i += ds/dr, x_j += ds/dy - ds/dr;  branch 2: x_i += ds/dy;
// every particle m of the set within h_w of y: w_m += sign_m beta W,
// x_m -= a_m (beta + beta0) grad_y W(y - x_m).
//
// Passes (one launch each; samples binned by fine cell as in the forward):
//   gen      the forward's draws, plus the global index of the branch-1 source
//   sort     stable radix sort by (sign, fine cell)
//   eval     one workgroup per tile of samples (the forward's LDS staging):
//            om(y) and grad om(y)
//   final    one wave per (owner, sign): adjoint weights per sample, the
//            owner's gradient (wave butterfly in sample order) and each
//            branch-1 sample's source gradient
//   recs     samples in cell order as (y, beta, beta + beta0)
//   field    one workgroup per (sign, fine cell): the cell's particles
//            against the samples in the cell's stencil -> d/dw_m, d/dx_m
//   sources  stable radix sort of the source gradients by source particle,
//            then one thread per particle adds its run in sample order
// Every sum has a fixed order, so the gradients are deterministic.
//
// The deformed step (src/cuda/simulation_step.cu:51-73: four auxiliary
// targets at x_i -/+ dx e_x, x_i +/- dx e_y and F per particle) has the
// state (x, F): particle m's blob is W(|F_m^-1 (y - x_m)|) through its
// quadratic form G_m = F_m^-T F_m^-1, a branch-1 sample is y = x_j + F_j u
// with u frozen, and the step returns F' = renorm(I + dt/(2 dx)
// [u_R - u_L | u_T - u_B]) beside x'.  Its adjoint (vpmc_step_adjoint_deformed)
// runs the step's forward estimate, turns dL/dF' into the four aux targets'
// adjoint weights (k_jac_adj), and then the passes above with: every
// branch-1 sample weighted at the five targets, each disk sample at its own
// (k_adj_final); om and grad om = sum a_m W'(q)/(q h_w^2) G_m (y - x_m) per
// sample (k_adj_eval_def); dL/dG_m = a_m (beta + beta0) W'(q)/(2 q h_w^2)
// (dx^2, dx dy, dy^2) for (G11, 2 G12, G22), mapped to dL/dF_m
// (k_adj_field<true>, quad_form_vjp); and dL/dF_j += (dL/dy) u^T from the
// source records (k_adj_sources).
#include <math.h>

#include <mutex>

#include "binned_core.hpp"

namespace vpmc {

constexpr int ADJ_THREADS = 256;

// ---- device helpers ------------------------------------------------------------

// W(r)/sig2 and (dW/dr)/(r sig2) of the cubic spline (include/utilities.hpp:7-17),
// branch-free: W/sig2 = 2 t^3 - 8 u^3 with t = max(1-q, 0), u = max(1/2-q, 0);
// (dW/dq)/q = 18 q - 12 for q <= 1/2 and -6 t^2 / q beyond (0 past the support)
__device__ __forceinline__ void spline_wg(float r2, float inv_hw, float &w, float &gr)
{
    const float q = __builtin_amdgcn_sqrtf(r2) * inv_hw;
    const float t = fmaxf(1.f - q, 0.f), u = fmaxf(0.5f - q, 0.f);
    w = fmaf(2.f * t * t, t, -8.f * u * u * u);
    const float gin = fmaf(18.f, q, -12.f);
    const float gout = -6.f * t * t * __builtin_amdgcn_rcpf(q);
    gr = (q <= 0.5f ? gin : gout) * (inv_hw * inv_hw);
}

__device__ __forceinline__ void spline_wg64(double r2, double inv_hw, double &w, double &gr)
{
    const double q = sqrt(r2) * inv_hw;
    w = 0.0;
    gr = 0.0;
    if (!(q < 1.0)) return;
    if (q <= 0.5) {
        w = 6.0 * q * q * (q - 1.0) + 1.0;
        gr = (18.0 * q - 12.0) * inv_hw * inv_hw;
    } else {
        const double t = 1.0 - q;
        w = 2.0 * t * t * t;
        gr = -6.0 * t * t * inv_hw * inv_hw / q;
    }
}

// Visits every record of field f within the search radius of p (periodic
// images by column / row wrap, as field_sum): fn(record, dx, dy, k) with
// (dx, dy) = record position - p in the image frame and k the record's index.
template <class F>
__device__ __forceinline__ void field_visit(const FieldD &f, float px, float py, F &fn)
{
    const FieldGeom &g = f.g;
    const int cx = min((int)(px * g.inv_cx), g.nx - 1);
    for (int ox = -g.reach; ox <= g.reach; ++ox) {
        const int col = cx + ox;
        float gx = 0.f;
        if (ox < 0) gx = px - (float)(col + 1) * g.cw;
        else if (ox > 0) gx = (float)col * g.cw - px;
        gx = fmaxf(gx, 0.f);
        const float rem = g.rs2 - gx * gx;
        if (rem < 0.f) continue;
        const float ry = sqrtf(rem);
        int r0 = (int)floorf((py - ry) * g.inv_cy);
        int r1 = (int)floorf((py + ry) * g.inv_cy);
        float qx = px;
        int colw = col;
        if (colw < 0) { colw += g.nx; qx = px + g.Lx; }
        else if (colw >= g.nx) { colw -= g.nx; qx = px - g.Lx; }
        const int *cs = f.cell_start + (size_t)colw * g.ny;
        auto run = [&](int b, int e, float sx, float sy) {
            for (int k = b; k < e; ++k) {
                const float4 c = f.pts[k];
                fn(c, c.x - sx, c.y - sy, k);
            }
        };
        if (r0 < 0) { run(cs[g.ny + r0], cs[g.ny], qx, py + g.Ly); r0 = 0; }
        if (r1 >= g.ny) { run(cs[0], cs[r1 - g.ny + 1], qx, py - g.Ly); r1 = g.ny - 1; }
        run(cs[r0], cs[r1 + 1], qx, py);
    }
}

// The same traversal over FP64 records (x, y, a, a0) kept in the float
// field's cell order; cell ranges from the FP32 geometry, distances in FP64.
template <class F>
__device__ __forceinline__ void field_visit64(const FieldGeom &g, const int *__restrict__ cell_start,
                                              const double4 *__restrict__ recs, double px,
                                              double py, F &fn)
{
    const float pxf = (float)px, pyf = (float)py;
    const int cx = min((int)(pxf * g.inv_cx), g.nx - 1);
    for (int ox = -g.reach; ox <= g.reach; ++ox) {
        const int col = cx + ox;
        float gx = 0.f;
        if (ox < 0) gx = pxf - (float)(col + 1) * g.cw;
        else if (ox > 0) gx = (float)col * g.cw - pxf;
        gx = fmaxf(gx, 0.f);
        const float rem = g.rs2 - gx * gx;
        if (rem < 0.f) continue;
        const float ry = sqrtf(rem);
        int r0 = (int)floorf((pyf - ry) * g.inv_cy);
        int r1 = (int)floorf((pyf + ry) * g.inv_cy);
        double qx = px;
        int colw = col;
        if (colw < 0) { colw += g.nx; qx = px + (double)g.Lx; }
        else if (colw >= g.nx) { colw -= g.nx; qx = px - (double)g.Lx; }
        const int *cs = cell_start + (size_t)colw * g.ny;
        auto run = [&](int b, int e, double sx, double sy) {
            for (int k = b; k < e; ++k) {
                const double4 c = recs[k];
                fn(c, c.x - sx, c.y - sy, k);
            }
        };
        if (r0 < 0) { run(cs[g.ny + r0], cs[g.ny], qx, py + (double)g.Ly); r0 = 0; }
        if (r1 >= g.ny) { run(cs[0], cs[r1 - g.ny + 1], qx, py - (double)g.Ly); r1 = g.ny - 1; }
        run(cs[r0], cs[r1 + 1], qx, py);
    }
}

// gradient of a . K_s(r) with respect to r, K_s the smooth part of the
// periodic kernel (kper_smooth): with A = P(s) + pi conj(s), s = r/L,
// P(s) = sum_m c_m s^(4m+3) and Q = P'(s):
//   d/drx = -(ax Im Q + ay (Re Q + pi)) / (2 pi L^2)
//   d/dry = -(ax (Re Q - pi) - ay Im Q) / (2 pi L^2)
__device__ __forceinline__ void kper_smooth_grad(const McD &d, float rx, float ry, float ax,
                                                 float ay, float &gx, float &gy)
{
    const float sr = rx * d.inv_L, si = ry * d.inv_L;
    const float s2r = sr * sr - si * si, s2i = 2.f * sr * si;
    const float wr = s2r * s2r - s2i * s2i, wi = 2.f * s2r * s2i;
    constexpr int nt = KPER_TERMS_F32;   // make_mcd's term count; unrolled
    float hr = d.kc[nt - 1] * (float)(4 * (nt - 1) + 3), hi = 0.f;
#pragma unroll
    for (int m = nt - 2; m >= 0; --m) {
        const float t = hr * wr - hi * wi + d.kc[m] * (float)(4 * m + 3);
        hi = hr * wi + hi * wr;
        hr = t;
    }
    const float Qr = hr * s2r - hi * s2i, Qi = hr * s2i + hi * s2r;
    const float c = d.inv_2piL * d.inv_L;
    gx = -(ax * Qi + ay * (Qr + PI_F)) * c;
    gy = -(ax * (Qr - PI_F) - ay * Qi) * c;
}

__device__ __forceinline__ void kper_smooth_f64(const double *kc, int nt, double L, double rx,
                                                double ry, double &kx, double &ky)
{
    const double sr = rx / L, si = ry / L;
    const double s2r = sr * sr - si * si, s2i = 2.0 * sr * si;
    const double wr = s2r * s2r - s2i * s2i, wi = 2.0 * s2r * s2i;
    double hr = kc[nt - 1], hi = 0.0;
    for (int m = nt - 2; m >= 0; --m) {
        const double t = hr * wr - hi * wi + kc[m];
        hi = hr * wi + hi * wr;
        hr = t;
    }
    const double s3r = s2r * sr - s2i * si, s3i = s2r * si + s2i * sr;
    const double Ar = hr * s3r - hi * s3i + PI_D * sr;
    const double Ai = hr * s3i + hi * s3r - PI_D * si;
    const double c = 1.0 / (2.0 * PI_D * L);
    kx = -Ai * c;
    ky = -Ar * c;
}

// FP64 inverse of the cubic spline's radial mass fraction (the FP32
// mass_inverse's equation, bracketed Newton to convergence)
__device__ __forceinline__ double mass_inverse_f64(double u)
{
    if (!(u > 0.0)) return 0.0;
    if (u >= 1.0) return 1.0;
    const double target = (7.0 / 80.0) * u;
    double lo = 0.0, hi = 1.0, q = 0.5;
    for (int it = 0; it < 60; ++it) {
        double f, df;
        const double q2 = q * q;
        if (q <= 0.5) {
            f = 0.5 * q2 - 1.5 * q2 * q2 + 1.2 * q2 * q2 * q;
            df = q * (1.0 - 6.0 * q2 + 6.0 * q2 * q);
        } else {
            const double t = 1.0 - q, t2 = t * t;
            f = 7.0 / 80.0 - (0.5 * t2 * t2 - 0.4 * t2 * t2 * t);
            df = 2.0 * t2 * t * (1.0 - t);
        }
        f -= target;
        if (f > 0.0) hi = q; else lo = q;
        double qn = (df > 0.0) ? q - f / df : 0.5 * (lo + hi);
        if (!(qn > lo && qn < hi)) qn = 0.5 * (lo + hi);
        if (fabs(qn - q) < 1e-16) { q = qn; break; }
        q = qn;
    }
    return q;
}

// ---- the estimator adjoint -------------------------------------------------------

__global__ void __launch_bounds__(BIN_THREADS)
    k_adj_gen(McD d, BinGeom g, const int *__restrict__ sub0, const int *__restrict__ sub1,
              uint32_t *__restrict__ key, float2 *__restrict__ q, float2 *__restrict__ r,
              int *__restrict__ src, float2 *__restrict__ jit)
{
    const size_t sid = (size_t)blockIdx.x * BIN_THREADS + threadIdx.x;
    if (sid >= g.S) return;
    const uint32_t pair = div_ns(g, (uint32_t)sid);   // sid < 2^31 (bin_geom)
    const int k = (int)((uint32_t)sid - pair * g.NS);
    const int owner = (int)(pair >> 1), si = (int)(pair & 1);
    const SetD &S = d.set[si];
    const int n_s = S.count_dev ? *S.count_dev : S.n;
    float qx = 0.f, qy = 0.f, rx = 0.f, ry = 0.f;
    int j = -1;
    float2 ju = make_float2(0.f, 0.f);
    const bool live = n_s > 0 && gen_sample<VPMC_MODE_CONSISTENT, true>(
                                     d, S, n_s, owner, si, k, d.owner[owner], qx, qy, rx, ry, &j,
                                     jit ? &ju : nullptr);
    if (jit) jit[sid] = ju;
    r[sid] = make_float2(rx, ry);
    q[sid] = make_float2(qx, qy);
    const int *sub = si ? sub1 : sub0;
    src[sid] = (live && j >= 0) ? sub[j] : -1;
    if (!live) {
        key[sid] = (uint32_t)g.ntile;
        return;
    }
    const int cx = min(max((int)(qx * g.inv_cx), 0), g.nx - 1);
    const int cy = min(max((int)(qy * g.inv_cy), 0), g.ny - 1);
    key[sid] = (uint32_t)(si * g.ncell + cx * g.ny + cy);
}

// om(y) and grad om(y) for every live sample: one workgroup per tile (sign,
// fine cell) of samples, the cell's particle stencil staged in LDS as in the
// forward's evaluation (tile_stage, in world units: scale 1), one sample per
// thread.
__global__ void __launch_bounds__(ADJ_THREADS)
    k_adj_eval(McD m, BinGeom g, const int *__restrict__ start, const float4 *__restrict__ srec,
               float *__restrict__ om, float2 *__restrict__ G)
{
    constexpr int CAP = TileCfg<ADJ_THREADS>::CAP;
    const int tile = (int)(blockIdx.x & 7) * g.per_xcd + (int)(blockIdx.x >> 3);
    if (tile >= g.ntile) return;
    const int beg = start[tile], end = start[tile + 1];
    if (beg >= end) return;   // uniform over the workgroup
    __shared__ TileLds<ADJ_THREADS> L;
    const int si = tile >= g.ncell ? 1 : 0;
    const int c = tile - si * g.ncell;
    const int cx = c / g.ny, cy = c - cx * g.ny;
    const float ox = ((float)cx + 0.5f) * g.cw, oy = ((float)cy + 0.5f) * g.ch;
    const float4 *__restrict__ pts = m.set[si].field.pts;
    const int total = tile_pieces<ADJ_THREADS>(g, m.set[si].field.cell_start, cx, cy, m.Lx, m.Ly,
                                                L);
    const int npieces = 3 * (2 * g.cols + 1);
    const int npass = (total + CAP - 1) / CAP;
    int n = 0;
    if (npass == 1)
        n = tile_stage<ADJ_THREADS>(g, pts, npieces, 0, total, cx, ox, oy, m.Lx, m.Ly, 1.f, L);
    const int wid = (int)threadIdx.x / WAVE;
    for (int bb = beg; bb < end; bb += ADJ_THREADS) {   // uniform trip count
        if (npass == 1 && bb + wid * WAVE >= end) break;   // no sample left for this wave
        const int idx = bb + (int)threadIdx.x;
        const bool live = idx < end;
        const float4 p = live ? srec[idx] : make_float4(ox, oy, 0.f, 0.f);
        const float px = p.x - ox, py = p.y - oy;
        float acc = 0.f, gx = 0.f, gy = 0.f;
        for (int pass = 0; pass < npass; ++pass) {
            if (npass > 1)
                n = tile_stage<ADJ_THREADS>(g, pts, npieces, pass * CAP,
                                            min(total, (pass + 1) * CAP), cx, ox, oy, m.Lx, m.Ly,
                                            1.f, L);
            for (int k = 0; k < n; ++k) {
                // staged record (-2 c', |c'|^2, w), c' = particle - cell centre
                const float4 r = L.cand[k];
                const float dx = fmaf(-0.5f, r.x, -px), dy = fmaf(-0.5f, r.y, -py);
                float w, gr;
                spline_wg(fmaf(dx, dx, dy * dy), m.inv_hw, w, gr);
                acc = fmaf(r.w, w, acc);
                // grad_y W(|y - x|) = W'(r)/r (y - x), (dx, dy) = x - y
                gx = fmaf(-r.w * gr, dx, gx);
                gy = fmaf(-r.w * gr, dy, gy);
            }
        }
        if (live) {
            const uint32_t sid = __float_as_uint(p.z);
            om[sid] = m.sig2 * acc;
            G[sid] = make_float2(m.sig2 * gx, m.sig2 * gy);
        }
    }
}

// Deformed fields: om(y) and grad om(y) per live sample, one thread per
// sample in tile order (the lanes of a wave query points of one cell and
// walk the same records); record k's kernel is W(|F^-1 d|) through its
// quadratic form G (field.gq), so grad_y W = W'(q)/(q h_w^2) G (y - x).
__global__ void __launch_bounds__(ADJ_THREADS)
    k_adj_eval_def(McD m, BinGeom g, const int *__restrict__ start,
                   const float4 *__restrict__ srec, const uint32_t *__restrict__ keys_sorted,
                   float *__restrict__ om, float2 *__restrict__ G)
{
    const int p = blockIdx.x * ADJ_THREADS + threadIdx.x;
    if (p >= start[g.ntile]) return;   // dead samples sort last
    const float4 sr = srec[p];
    const int si = keys_sorted[p] >= (uint32_t)g.ncell ? 1 : 0;
    const FieldD &f = m.set[si].field;
    float acc = 0.f, gx = 0.f, gy = 0.f;
    auto fn = [&](const float4 &c, float dx, float dy, int k) {   // (dx, dy) = x - y
        const float4 Q = f.gq[k];   // (G11, 2 G12, G22, sigma_max^2)
        // the search covers the deformation bound; this record reaches only
        // sigma_max h_w (beyond it its W and W' are exactly zero)
        if (fmaf(dx, dx, dy * dy) > Q.w * g.rs2_1) return;
        const float r2 = fmaxf(fmaf(dx, fmaf(Q.y, dy, Q.x * dx), Q.z * dy * dy), 0.f);
        float w, gr;
        spline_wg(r2, m.inv_hw, w, gr);
        acc = fmaf(c.z, w, acc);
        const float hx = fmaf(0.5f * Q.y, dy, Q.x * dx), hy = fmaf(0.5f * Q.y, dx, Q.z * dy);
        gx = fmaf(-c.z * gr, hx, gx);   // G (y - x) = -(hx, hy)
        gy = fmaf(-c.z * gr, hy, gy);
    };
    field_visit(f, sr.x, sr.y, fn);
    const uint32_t sid = __float_as_uint(sr.z);
    om[sid] = m.sig2 * acc;
    G[sid] = make_float2(m.sig2 * gx, m.sig2 * gy);
}

// One (sample, target) term: (ax, ay) the target's adjoint weight (sign and
// 1/(N1+N2) applied), o = om(y), (rx, ry) = target - y.  Adds ds/d om to b
// and ds/d om + ds/d om0 to bt, and returns ds/dr in (drx, dry); a term the
// forward does not add (om <= 0, r = 0) adds nothing.
__device__ __forceinline__ void adj_term(const McD &d, float ax, float ay, float o, float Om,
                                         float rx, float ry, float &b, float &bt, float &drx,
                                         float &dry)
{
    drx = dry = 0.f;
    const float r2 = rx * rx + ry * ry, rn = sqrtf(r2);
    if (!(o > 0.f && rn > 0.f)) return;
    const float f = 1.f / (2.f * PI_F * r2);
    float ksx, ksy;
    kper_smooth(d, rx, ry, ksx, ksy);
    const float Kx = -ry * f + ksx, Ky = rx * f + ksy;
    const float aK = ax * Kx + ay * Ky;
    const float D = d.w1 * o / Om + d.w2 * d.alpha_c / rn;
    const float iD = 1.f / D;
    const float bb = aK * iD;
    b += bb;
    bt += bb - aK * o * d.w1 * iD * iD / Om;
    // grad_r (a . K): free part, then the smooth part
    const float P = -ax * ry + ay * rx;
    const float gfx = (ay - 2.f * P * rx / r2) * f;
    const float gfy = (-ax - 2.f * P * ry / r2) * f;
    float gsx, gsy;
    kper_smooth_grad(d, rx, ry, ax, ay, gsx, gsy);
    const float c1 = o * iD, c2 = aK * o * d.w2 * d.alpha_c * iD * iD / (r2 * rn);
    drx = c1 * (gfx + gsx) + c2 * rx;
    dry = c1 * (gfy + gsy) + c2 * ry;
}

// one wave per (owner, sign).  With auxiliary targets (d.aux), a branch-1
// sample adds its term at each of the five targets (displacement r + offset,
// min image, as the forward's final pass) with that target's adjoint weight
// adj_aux[(t-1) n + owner], and a disk sample of target t at t only; the
// targets move with the owner.  jit (deformed sets): the sample's jitter
// before F, carried to the source record for dL/dF of the source.
__global__ void __launch_bounds__(BIN_THREADS)
    k_adj_final(McD d, BinGeom g, const uint32_t *__restrict__ key, const float2 *__restrict__ r,
                const float *__restrict__ om, const float2 *__restrict__ G,
                const int *__restrict__ src, const float2 *__restrict__ adj_vel,
                const float2 *__restrict__ adj_aux, const float2 *__restrict__ jit,
                float2 *__restrict__ own_grad, float2 *__restrict__ beta,
                float4 *__restrict__ srcrec, uint32_t *__restrict__ srckey, uint32_t dead_src)
{
    const int pair = blockIdx.x * (BIN_THREADS / WAVE) + threadIdx.x / WAVE;
    if (pair >= 2 * d.n_owner) return;
    const int lane = threadIdx.x & (WAVE - 1);
    const int owner = pair >> 1, si = pair & 1;
    const SetD &S = d.set[si];
    const int n_s = S.count_dev ? *S.count_dev : S.n;
    const float sg = si ? -d.inv_tot : d.inv_tot;   // u = v+ - v-, v = sum / (N1 + N2)
    float ax[5], ay[5];
    {
        const float2 av = adj_vel[owner];
        ax[0] = sg * av.x;
        ay[0] = sg * av.y;
    }
#pragma unroll
    for (int t = 1; t < 5; ++t) {
        const float2 av = d.aux ? adj_aux[(size_t)(t - 1) * d.n_owner + owner]
                                : make_float2(0.f, 0.f);
        ax[t] = sg * av.x;
        ay[t] = sg * av.y;
    }
    const float Om = n_s > 0 ? (float)(*S.total_dev) : 1.f;
    float gxs = 0.f, gys = 0.f;
    const size_t base = (size_t)pair * g.NS;
    for (uint32_t k = lane; k < g.NS; k += WAVE) {
        const size_t sid = base + k;
        float b = 0.f, bt = 0.f, ex = 0.f, ey = 0.f;
        float2 ju = make_float2(0.f, 0.f);
        uint32_t sk = dead_src;
        if (n_s > 0 && key[sid] < (uint32_t)g.ntile) {
            const float o = om[sid];
            const float2 rr = r[sid];
            const float2 Gk = G[sid];
            const int j = src[sid];
            if (j >= 0) {   // branch 1: r = target - (x_j + jitter)
                float drx, dry;
                adj_term(d, ax[0], ay[0], o, Om, rr.x, rr.y, b, bt, drx, dry);
                float sx = drx, sy = dry;
                if (d.aux) {
#pragma unroll
                    for (int t = 1; t < 5; ++t) {
                        const float2 of = aux_offset(t, d.aux_dx);
                        float rx = rr.x + of.x, ry = rr.y + of.y;
                        if (rx > 0.5f * d.Lx) rx -= d.Lx; else if (rx < -0.5f * d.Lx) rx += d.Lx;
                        if (ry > 0.5f * d.Ly) ry -= d.Ly; else if (ry < -0.5f * d.Ly) ry += d.Ly;
                        adj_term(d, ax[t], ay[t], o, Om, rx, ry, b, bt, drx, dry);
                        sx += drx;
                        sy += dry;
                    }
                }
                gxs += sx;
                gys += sy;
                ex = bt * Gk.x - sx;
                ey = bt * Gk.y - sy;
                sk = (uint32_t)j;
                if (jit) ju = jit[sid];
            } else {   // branch 2: y = target + offset, r fixed
                const int kb = (int)k - d.N1;
                const int t = kb < d.N2 ? 0 : 1 + (kb - d.N2) / d.N2;
                float drx, dry, bx = ax[0], by = ay[0];
#pragma unroll
                for (int tt = 1; tt < 5; ++tt)
                    if (tt == t) {
                        bx = ax[tt];
                        by = ay[tt];
                    }
                adj_term(d, bx, by, o, Om, rr.x, rr.y, b, bt, drx, dry);
                gxs += bt * Gk.x;
                gys += bt * Gk.y;
            }
        }
        beta[sid] = make_float2(b, bt);
        srcrec[sid] = make_float4(ex, ey, ju.x, ju.y);
        srckey[sid] = sk;
    }
    gxs = wave_sum(gxs);
    gys = wave_sum(gys);
    if (lane == 0) own_grad[pair] = make_float2(gxs, gys);
}

__global__ void k_adj_recs(const int *__restrict__ start, int ntile, const float4 *__restrict__ srec,
                           const float2 *__restrict__ beta, float4 *__restrict__ rec2)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= start[ntile]) return;
    const float4 s = srec[p];
    const float2 b = beta[__float_as_uint(s.z)];
    rec2[p] = make_float4(s.x, s.y, b.x, b.y);
}

// dL/dF of F = (a b; c d) from dL/d(G11, 2 G12, G22) of its quadratic form
// G = F^-T F^-1 = (q, -r; -r, p) / det^2 with p = a^2 + b^2, q = c^2 + d^2,
// r = ac + bd, det = ad - bc (quad_form's first three outputs)
__device__ __forceinline__ float4 quad_form_vjp(float4 F, float g1, float g2, float g3)
{
    const float a = F.x, b = F.y, c = F.z, d = F.w;
    const float det = a * d - b * c;
    const float i2 = 1.f / (det * det), gi = -2.f * i2 / det;   // d(det^-2)/d det
    const float p = a * a + b * b, q = c * c + d * d, r = a * c + b * d;
    const float S = (g1 * q - 2.f * g2 * r + g3 * p) * gi;   // coefficient of d det
    return make_float4(i2 * (2.f * g3 * a - 2.f * g2 * c) + S * d,
                       i2 * (2.f * g3 * b - 2.f * g2 * d) - S * c,
                       i2 * (2.f * g1 * c - 2.f * g2 * a) - S * b,
                       i2 * (2.f * g1 * d - 2.f * g2 * b) + S * a);
}

// One workgroup per (sign, fine cell): the particles of the cell, up to
// ADJ_PCHUNK at a time, against every sample of the same sign in the cell's
// stencil (the forward's tile stencil over the sample array), threads
// striding over the samples; fixed-order block reduction per particle.
// DEF: the particle's kernel is W(|F^-1 d|) (its quadratic form gq in field
// order, F = jac by global index): grad_y W = W'(q)/(q h_w^2) G d, and
// dW/dG = W'(q)/(2 q h_w^2) (dx^2, dx dy, dy^2) for (G11, 2 G12, G22), which
// quad_form_vjp turns into dL/dF (grad_jac).
constexpr int ADJ_PCHUNK = 8;

template <bool DEF>
__global__ void __launch_bounds__(ADJ_THREADS)
    k_adj_field(McD d, BinGeom g, const int *__restrict__ start, const float4 *__restrict__ rec2,
                const int *__restrict__ order0, const int *__restrict__ order1,
                const int *__restrict__ sub0, const int *__restrict__ sub1,
                const float4 *__restrict__ jac, float2 *__restrict__ grad_pos,
                float *__restrict__ grad_w, float4 *__restrict__ grad_jac)
{
    constexpr int NR = DEF ? 6 : 3;   // reduced values per particle
    const int tile = (int)blockIdx.x;
    const int si = tile >= g.ncell ? 1 : 0;
    const int c = tile - si * g.ncell;
    const SetD &S = d.set[si];
    const int pb = S.field.cell_start[c], pe = S.field.cell_start[c + 1];
    if (pb >= pe) return;   // uniform: no particle of this sign in the cell
    const int cx = c / g.ny, cy = c - cx * g.ny;
    __shared__ TileLds<ADJ_THREADS> L;
    __shared__ float red[ADJ_THREADS / WAVE][NR * ADJ_PCHUNK];
    const int total = tile_pieces<ADJ_THREADS>(g, start + (size_t)si * g.ncell, cx, cy, d.Lx,
                                                d.Ly, L);
    const int npieces = 3 * (2 * g.cols + 1);
    const int lane = threadIdx.x & (WAVE - 1), wid = threadIdx.x / WAVE;
    const int *order = si ? order1 : order0;
    const int *sub = si ? sub1 : sub0;
    const float sg = si ? -1.f : 1.f;   // d|w|/dw on the set
    for (int p0 = pb; p0 < pe; p0 += ADJ_PCHUNK) {
        const int np = min(ADJ_PCHUNK, pe - p0);
        float px[ADJ_PCHUNK], py[ADJ_PCHUNK], acc[NR][ADJ_PCHUNK];
        float4 Q[DEF ? ADJ_PCHUNK : 1];
#pragma unroll
        for (int j = 0; j < ADJ_PCHUNK; ++j) {
            const float4 q = S.field.pts[p0 + min(j, np - 1)];
            px[j] = q.x;
            py[j] = q.y;
            if constexpr (DEF) Q[j] = S.field.gq[p0 + min(j, np - 1)];
#pragma unroll
            for (int v = 0; v < NR; ++v) acc[v][j] = 0.f;
        }
        int lo = 0;   // last piece with p_off <= i (i only grows)
        for (int i = threadIdx.x; i < total; i += ADJ_THREADS) {
            while (lo + 1 < npieces && L.p_off[lo + 1] <= i) ++lo;
            const float4 sm = rec2[L.p_beg[lo] + (i - L.p_off[lo])];
            const float2 sh = piece_shift(g, cx, lo, d.Lx, d.Ly);
            const float sx = sm.x + sh.x, sy = sm.y + sh.y;
#pragma unroll
            for (int j = 0; j < ADJ_PCHUNK; ++j) {
                if (j >= np) break;   // uniform
                // (dx, dy) = y - x_m; grad_y W = W'(r)/r (y - x_m)
                const float dx = sx - px[j], dy = sy - py[j];
                float w, gr;
                if constexpr (DEF) {
                    const float4 G = Q[j];
                    // the stencil covers the deformation bound, the particle
                    // only its own reach sigma_max h_w (zero terms beyond)
                    if (fmaf(dx, dx, dy * dy) > G.w * g.rs2_1) continue;
                    const float r2 =
                        fmaxf(fmaf(dx, fmaf(G.y, dy, G.x * dx), G.z * dy * dy), 0.f);
                    spline_wg(r2, d.inv_hw, w, gr);
                    const float sg2 = sm.w * gr;
                    acc[0][j] = fmaf(sm.z, w, acc[0][j]);
                    acc[1][j] = fmaf(sg2, fmaf(0.5f * G.y, dy, G.x * dx), acc[1][j]);
                    acc[2][j] = fmaf(sg2, fmaf(0.5f * G.y, dx, G.z * dy), acc[2][j]);
                    acc[3][j] = fmaf(sg2 * dx, dx, acc[3][j]);
                    acc[4][j] = fmaf(sg2 * dx, dy, acc[4][j]);
                    acc[5][j] = fmaf(sg2 * dy, dy, acc[5][j]);
                } else {
                    spline_wg(fmaf(dx, dx, dy * dy), d.inv_hw, w, gr);
                    acc[0][j] = fmaf(sm.z, w, acc[0][j]);
                    acc[1][j] = fmaf(sm.w * gr, dx, acc[1][j]);
                    acc[2][j] = fmaf(sm.w * gr, dy, acc[2][j]);
                }
            }
        }
#pragma unroll
        for (int j = 0; j < ADJ_PCHUNK; ++j) {
#pragma unroll
            for (int v = 0; v < NR; ++v) {
                const float a = wave_sum(acc[v][j]);
                if (lane == 0) red[wid][NR * j + v] = a;
            }
        }
        __syncthreads();
        if ((int)threadIdx.x < np) {
            const int j = threadIdx.x;
            float tot[NR];
#pragma unroll
            for (int v = 0; v < NR; ++v) tot[v] = 0.f;
#pragma unroll
            for (int w = 0; w < ADJ_THREADS / WAVE; ++w)
#pragma unroll
                for (int v = 0; v < NR; ++v) tot[v] += red[w][NR * j + v];
            const int t = p0 + j;
            const float aw = S.field.pts[t].z;
            const int gm = sub[order[t]];
            grad_w[gm] += sg * d.sig2 * tot[0];
            float2 gp = grad_pos[gm];
            gp.x -= aw * d.sig2 * tot[1];
            gp.y -= aw * d.sig2 * tot[2];
            grad_pos[gm] = gp;
            if constexpr (DEF) {
                const float h = 0.5f * aw * d.sig2;
                const float4 fb = quad_form_vjp(jac[gm], h * tot[3], h * tot[4], h * tot[5]);
                float4 gj = grad_jac[gm];
                gj.x += fb.x;
                gj.y += fb.y;
                gj.z += fb.z;
                gj.w += fb.w;
                grad_jac[gm] = gj;
            }
        }
        __syncthreads();   // red is rewritten by the next chunk
    }
}

// owner terms (both signs) and the branch-1 source terms, per particle; the
// source records are (dL/dy, jitter before F): y = x_j + F_j u gives
// dL/dF_j += dL/dy u^T (grad_jac, deformed sets)
__global__ void k_adj_sources(int n, const int *__restrict__ src_start,
                              const float4 *__restrict__ srcrec_sorted,
                              const float2 *__restrict__ own_grad, float2 *__restrict__ grad_pos,
                              float4 *__restrict__ grad_jac)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float2 a = own_grad[2 * (size_t)i], b = own_grad[2 * (size_t)i + 1];
    float sx = a.x + b.x, sy = a.y + b.y;
    float fa = 0.f, fb = 0.f, fc = 0.f, fd = 0.f;
    for (int p = src_start[i]; p < src_start[i + 1]; ++p) {
        const float4 e = srcrec_sorted[p];
        sx += e.x;
        sy += e.y;
        fa = fmaf(e.x, e.z, fa);
        fb = fmaf(e.x, e.w, fb);
        fc = fmaf(e.y, e.z, fc);
        fd = fmaf(e.y, e.w, fd);
    }
    float2 gp = grad_pos[i];
    gp.x += sx;
    gp.y += sy;
    grad_pos[i] = gp;
    if (grad_jac) {
        float4 gj = grad_jac[i];
        gj.x += fa;
        gj.y += fb;
        gj.z += fc;
        gj.w += fd;
        grad_jac[i] = gj;
    }
}

// The F update's adjoint, per owner: F' = A / sqrt(det A) with
// A = I + c [u_R - u_L | u_T - u_B] (jacobian_from5; c = dt / (2 dx), u = v+ - v-
// of this step's aux estimates va[sign][L, R, T, B][owner]).  jbar holds
// dL/dF' on entry; the adjoint weights of the four targets go to
// adj_aux[t][owner] (L, R, T, B) and jbar is cleared for the step's dL/dF.
__global__ void k_jac_adj(int n, const float2 *__restrict__ va, float c,
                          float4 *__restrict__ jbar, float2 *__restrict__ adj_aux)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    float2 u[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        const float2 p = va[(size_t)t * n + j], q = va[(size_t)(4 + t) * n + j];
        u[t] = make_float2(p.x - q.x, p.y - q.y);
    }
    float4 A;
    A.x = fmaf(u[1].x - u[0].x, c, 1.f);
    A.y = (u[2].x - u[3].x) * c;
    A.z = (u[1].y - u[0].y) * c;
    A.w = fmaf(u[2].y - u[3].y, c, 1.f);
    const float det = fmaf(A.x, A.w, -(A.y * A.z));
    const bool clamp = !(det > 1e-12f);
    const float sc = 1.f / sqrtf(clamp ? 1e-12f : det);
    const float4 Fb = jbar[j];
    // dsc/dA = -sc^3 / 2 d det/dA, d det/dA = (A.w, -A.z, -A.y, A.x)
    const float k = clamp ? 0.f : -0.5f * sc * sc * sc *
                                      (Fb.x * A.x + Fb.y * A.y + Fb.z * A.z + Fb.w * A.w);
    const float4 Ab = make_float4(Fb.x * sc + k * A.w, Fb.y * sc - k * A.z, Fb.z * sc - k * A.y,
                                  Fb.w * sc + k * A.x);
    const float2 aR = make_float2(c * Ab.x, c * Ab.z), aT = make_float2(c * Ab.y, c * Ab.w);
    adj_aux[j] = make_float2(-aR.x, -aR.y);                  // L
    adj_aux[(size_t)n + j] = aR;                             // R
    adj_aux[2 * (size_t)n + j] = aT;                         // T
    adj_aux[3 * (size_t)n + j] = make_float2(-aT.x, -aT.y);  // B
    jbar[j] = make_float4(0.f, 0.f, 0.f, 0.f);
}

__global__ void k_scale_xy(int n, const float2 *__restrict__ x, float a, float2 *__restrict__ y)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = make_float2(a * x[i].x, a * x[i].y);
}

struct AdjBufs {
    uint32_t *key, *keys_sorted, *srckey, *srckey_sorted;
    float2 *q, *r, *G, *beta, *own, *jit;
    float *om;
    int *src, *start, *src_start;
    float4 *srec, *rec2, *srcrec, *srcrec_sorted;
    void *sort_ws;
    size_t sort_bytes;
};

// def: deformed sets (the branch-1 jitter per sample is kept)
static size_t adj_ws_bytes(const BinGeom &g, int n, bool def)
{
    const size_t S = g.S > 0 ? g.S : 1, T = (size_t)g.ntile, N = (size_t)(n > 0 ? n : 1);
    return 4 * align256(S * 4) + (def ? 6 : 5) * align256(S * 8) + align256(S * 4) +
           align256(S * 4) + align256((T + 1) * 4) + align256((N + 1) * 4) +
           4 * align256(S * 16) + align256(2 * N * 8) + radix_workspace_bytes((int)S) + 8192;
}

static int carve_adj(const BinGeom &g, int n, bool def, void *ws, size_t bytes, AdjBufs *B)
{
    Carve cv(ws, bytes);
    const size_t S = g.S > 0 ? g.S : 1, T = (size_t)g.ntile, N = (size_t)(n > 0 ? n : 1);
    B->key = cv.take<uint32_t>(S);
    B->keys_sorted = cv.take<uint32_t>(S);
    B->srckey = cv.take<uint32_t>(S);
    B->srckey_sorted = cv.take<uint32_t>(S);
    B->q = cv.take<float2>(S);
    B->r = cv.take<float2>(S);
    B->G = cv.take<float2>(S);
    B->beta = cv.take<float2>(S);
    B->jit = def ? cv.take<float2>(S) : nullptr;
    B->own = cv.take<float2>(2 * N);
    B->om = cv.take<float>(S);
    B->src = cv.take<int>(S);
    B->start = cv.take<int>(T + 1);
    B->src_start = cv.take<int>(N + 1);
    B->srec = cv.take<float4>(S);
    B->rec2 = cv.take<float4>(S);
    B->srcrec = cv.take<float4>(S);
    B->srcrec_sorted = cv.take<float4>(S);
    B->sort_bytes = radix_workspace_bytes((int)S);
    B->sort_ws = cv.take<char>(B->sort_bytes);
    if (!B->key || !B->keys_sorted || !B->srckey || !B->srckey_sorted || !B->q || !B->r ||
        !B->G || !B->beta || (def && !B->jit) || !B->own || !B->om || !B->src || !B->start ||
        !B->src_start || !B->srec || !B->rec2 || !B->srcrec || !B->srcrec_sorted || !B->sort_ws)
        return VPMC_ERR_WORKSPACE;
    return VPMC_OK;
}

// def = false: undeformed kernels only.  def = true: the deformed step (the
// sets' F and quadratic forms, the four auxiliary targets).
static int adj_check(const vpmc_mc_desc *d, bool def, McD *m, BinGeom *g)
{
    if (!d) return VPMC_ERR_ARG;
    if (d->mode != VPMC_MODE_CONSISTENT || !d->periodic || d->uniforms) return VPMC_ERR_ARG;
    if (def) {
        if (!d->jac || !d->aux || !d->vel_aux || !(d->aux_dx > 0.f)) return VPMC_ERR_ARG;
        for (int s = 0; s < 2; ++s)
            if (!d->set[s].field.gq || !d->set[s].set_jac) return VPMC_ERR_ARG;
    } else if (d->aux || d->jac || d->set[0].field.gq || d->set[1].field.gq) {
        return VPMC_ERR_ARG;   // deformed steps: vpmc_step_adjoint_deformed
    }
    VPMC_TRY(make_mcd(d, m));
    VPMC_TRY(bin_geom(*m, d->N1, d->N2, d->n_owner, g));
    if (!tile_threads_ok(*g, ADJ_THREADS)) return VPMC_ERR_GRID;
    for (int s = 0; s < 2; ++s)
        if ((d->set[s].n > 0 || d->set[s].count_dev) && !d->set[s].field.order)
            return VPMC_ERR_ARG;
    return VPMC_OK;
}

// adj_aux (4 n float2, targets L, R, T, B), jac (F by global index) and
// grad_jac (dL/dF, accumulated) are given for a deformed descriptor only
int mc_adjoint_impl(const vpmc_mc_desc *d, bool def, const int *sub_pos, const int *sub_neg,
                    const float *adj_vel, const float *adj_aux, float *grad_pos, float *grad_w,
                    float *grad_jac, void *ws, size_t ws_bytes, hipStream_t s)
{
    McD m;
    BinGeom g;
    VPMC_TRY(adj_check(d, def, &m, &g));
    const int n = d->n_owner;
    if (n == 0) return VPMC_OK;
    if (!adj_vel || !grad_pos || !grad_w || !sub_pos || !sub_neg) return VPMC_ERR_ARG;
    if (def && (!adj_aux || !grad_jac)) return VPMC_ERR_ARG;
    AdjBufs B;
    VPMC_TRY(carve_adj(g, n, def, ws, ws_bytes, &B));
    const unsigned gs = (unsigned)((g.S + BIN_THREADS - 1) / BIN_THREADS);
    hipLaunchKernelGGL(k_adj_gen, dim3(gs), dim3(BIN_THREADS), 0, s, m, g, sub_pos, sub_neg,
                       B.key, B.q, B.r, B.src, B.jit);
    VPMC_CHECK_LAUNCH();
    VPMC_TRY(radix_sort_to_records(B.key, (int)g.S, key_bits_for((uint32_t)g.ntile),
                                   (const float2 *)B.q, nullptr, B.srec, B.keys_sorted, g.ntile,
                                   B.start, B.sort_ws, B.sort_bytes, s));
    const unsigned ga = (unsigned)((g.S + ADJ_THREADS - 1) / ADJ_THREADS);
    if (def)
        hipLaunchKernelGGL(k_adj_eval_def, dim3(ga), dim3(ADJ_THREADS), 0, s, m, g,
                           (const int *)B.start, (const float4 *)B.srec,
                           (const uint32_t *)B.keys_sorted, B.om, B.G);
    else
        hipLaunchKernelGGL(k_adj_eval, dim3((unsigned)(8 * g.per_xcd)), dim3(ADJ_THREADS), 0, s,
                           m, g, (const int *)B.start, (const float4 *)B.srec, B.om, B.G);
    const unsigned gf = (unsigned)((2 * (size_t)n + 3) / 4);
    hipLaunchKernelGGL(k_adj_final, dim3(gf), dim3(BIN_THREADS), 0, s, m, g,
                       (const uint32_t *)B.key, (const float2 *)B.r, (const float *)B.om,
                       (const float2 *)B.G, (const int *)B.src, (const float2 *)adj_vel,
                       (const float2 *)adj_aux, (const float2 *)B.jit, B.own, B.beta, B.srcrec,
                       B.srckey, (uint32_t)n);
    hipLaunchKernelGGL(k_adj_recs, dim3(ga), dim3(ADJ_THREADS), 0, s, (const int *)B.start,
                       g.ntile, (const float4 *)B.srec, (const float2 *)B.beta, B.rec2);
    VPMC_CHECK_LAUNCH();
    if (def)
        hipLaunchKernelGGL(k_adj_field<true>, dim3((unsigned)g.ntile), dim3(ADJ_THREADS), 0, s,
                           m, g, (const int *)B.start, (const float4 *)B.rec2,
                           d->set[0].field.order, d->set[1].field.order, sub_pos, sub_neg,
                           (const float4 *)d->jac, (float2 *)grad_pos, grad_w,
                           (float4 *)grad_jac);
    else
        hipLaunchKernelGGL(k_adj_field<false>, dim3((unsigned)g.ntile), dim3(ADJ_THREADS), 0, s,
                           m, g, (const int *)B.start, (const float4 *)B.rec2,
                           d->set[0].field.order, d->set[1].field.order, sub_pos, sub_neg,
                           (const float4 *)nullptr, (float2 *)grad_pos, grad_w,
                           (float4 *)nullptr);
    VPMC_CHECK_LAUNCH();
    VPMC_TRY(radix_sort_to_records(B.srckey, (int)g.S, key_bits_for((uint32_t)n), nullptr,
                                   (const float4 *)B.srcrec, B.srcrec_sorted, B.srckey_sorted, n,
                                   B.src_start, B.sort_ws, B.sort_bytes, s));
    hipLaunchKernelGGL(k_adj_sources, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n,
                       (const int *)B.src_start, (const float4 *)B.srcrec_sorted,
                       (const float2 *)B.own, (float2 *)grad_pos,
                       def ? (float4 *)grad_jac : (float4 *)nullptr);
    VPMC_CHECK_LAUNCH();
    return VPMC_OK;
}

// ---- FP64 forward of the same frozen-sample estimator (checks) -------------------

constexpr int KPER_TERMS_ADJ64 = 24;
struct Kper64 {
    double c[KPER_TERMS_ADJ64];
};
const double *kper_coeffs_f64_host(int nterms);

struct F64Args {
    const double2 *pos;         // global positions
    const double4 *rec[2];      // per set, cell order: (x, y, a, a0)
    const double2 *set64[2];    // per set, local order: positions
    const int *cell_start[2];
    FieldGeom geom;
    double h_w, inv_hw, sig2, L, Rc, alpha, w1, w2, inv_tot;
    double2 *contrib;           // per sample
    double2 *y, *r;             // per sample: query point, owner - point
    Kper64 kc;
    // deformed step (all NULL / 0 otherwise): per set, cell order, the
    // records' quadratic forms (G11, 2 G12, G22, sigma_max^2) of F = jac64; per set,
    // local order, F; per sample (om, om0); the aux estimates
    // [sign][L, R, T, B][owner]
    const double4 *G64[2];
    const double4 *setjac64[2];
    double2 *om2;
    double2 *vel_aux;
    double aux_dx;
};

// G = F^-T F^-1 of F = (a b; c d) as (G11, 2 G12, G22, sigma_max(F)^2), FP64
// (sigma_max^2 the larger eigenvalue of F F^T, with a relative margin far
// above FP64 rounding: |F^-1 d| >= |d| / sigma_max)
__device__ __forceinline__ double4 quad_form_f64(double4 F)
{
    const double det = F.x * F.w - F.y * F.z;
    const double i2 = 1.0 / (det * det);
    const double p = F.x * F.x + F.y * F.y, q = F.z * F.z + F.w * F.w;
    const double r = F.x * F.z + F.y * F.w;
    const double smax2 = 0.5 * ((p + q) + sqrt((p - q) * (p - q) + 4.0 * r * r));
    return make_double4(q * i2, -2.0 * r * i2, p * i2, smax2 * (1.0 + 1e-12));
}

__global__ void k_f64_recs(int n_s, const int *__restrict__ count_dev,
                           const int *__restrict__ order, const int *__restrict__ sub,
                           const double2 *__restrict__ pos, const double *__restrict__ w,
                           const float *__restrict__ w0, double sg, double4 *__restrict__ rec,
                           double2 *__restrict__ set64, const double4 *__restrict__ jac64,
                           double4 *__restrict__ G64, double4 *__restrict__ setjac64)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int n = count_dev ? *count_dev : n_s;
    if (t >= n) return;
    const int gt = sub[order[t]];
    rec[t] = make_double4(pos[gt].x, pos[gt].y, sg * w[gt], fabs((double)w0[gt]));
    set64[t] = pos[sub[t]];
    if (jac64) {
        G64[t] = quad_form_f64(jac64[gt]);
        setjac64[t] = jac64[sub[t]];
    }
}

// Per sample: the draws, the point y (wrapped into the box) and the min-image
// displacement r = x_i - y, in FP64; key = the FP32 field cell of y in its
// sign's block of cells, or 2 nx ny for a sample without a query (empty set,
// branch-2 offset outside the box).
__global__ void __launch_bounds__(ADJ_THREADS) k_f64_gen(McD d, BinGeom g, F64Args a,
                                                         uint32_t *__restrict__ key)
{
    const size_t sid = (size_t)blockIdx.x * ADJ_THREADS + threadIdx.x;
    if (sid >= g.S) return;
    const uint32_t pair = div_ns(g, (uint32_t)sid);   // sid < 2^31 (bin_geom)
    const int k = (int)((uint32_t)sid - pair * g.NS);
    const int owner = (int)(pair >> 1), si = (int)(pair & 1);
    const SetD &S = d.set[si];
    const int n_s = S.count_dev ? *S.count_dev : S.n;
    const FieldGeom &fg = a.geom;
    uint32_t kk = 2u * (uint32_t)fg.nx * (uint32_t)fg.ny;
    double yx = 0.0, yy = 0.0, rx = 0.0, ry = 0.0;
    if (n_s > 0) {
        const bool b1 = k < d.N1;
        int t = 0, kb = b1 ? k : k - d.N1;   // t: target of a disk sample (0 = owner)
        if (!b1 && kb >= d.N2) {
            t = 1 + (kb - d.N2) / d.N2;
            kb = (kb - d.N2) - (t - 1) * d.N2;
        }
        float u[3];
        draws(d, owner, si, b1 ? 1 : 2 + t, kb, u);
        double2 xi = a.pos[owner];
        if (t) {   // the auxiliary target
            const float2 o = aux_offset(t, d.aux_dx);
            xi.x += (double)o.x;
            xi.y += (double)o.y;
        }
        const double L = a.L;
        bool live = true;
        if (b1) {
            const int j = guided_upper_bound(u[0], S.cdf, S.guide, S.glog2, n_s);
            const double2 xj = a.set64[si][j];
            const double rho = mass_inverse_f64((double)u[1]) * a.h_w;
            double sn, cs;
            sincos(2.0 * PI_D * (double)u[2], &sn, &cs);
            double jx = rho * cs, jy = rho * sn;
            if (a.setjac64[si]) {   // the deformed blob: F_j u
                const double4 F = a.setjac64[si][j];
                const double ex = F.x * jx + F.y * jy, ey = F.z * jx + F.w * jy;
                jx = ex;
                jy = ey;
            }
            yx = xj.x + jx;
            yy = xj.y + jy;
            rx = xi.x - yx;
            ry = xi.y - yy;
            if (rx > 0.5 * L) rx -= L; else if (rx < -0.5 * L) rx += L;
            if (ry > 0.5 * L) ry -= L; else if (ry < -0.5 * L) ry += L;
        } else {
            const double rho = a.Rc * (double)u[1];
            double sn, cs;
            sincos(2.0 * PI_D * (double)u[0], &sn, &cs);
            yx = xi.x + rho * cs;
            yy = xi.y + rho * sn;
            rx = -rho * cs;
            ry = -rho * sn;
            live = fabs(rx) <= 0.5 * L && fabs(ry) <= 0.5 * L;
        }
        if (live) {
            yx -= floor(yx / L) * L;
            yy -= floor(yy / L) * L;
            // the cell field_visit64 starts from
            const int cx = min((int)((float)yx * fg.inv_cx), fg.nx - 1);
            const int cy = min((int)((float)yy * fg.inv_cy), fg.ny - 1);
            kk = ((uint32_t)si * (uint32_t)fg.nx + (uint32_t)cx) * (uint32_t)fg.ny + (uint32_t)cy;
        }
    }
    key[sid] = kk;
    a.y[sid] = make_double2(yx, yy);
    a.r[sid] = make_double2(rx, ry);
}

// One thread per sample in cell order (order[t] = sample id, keys[t] its
// key): the lanes of a wave query points of one cell and walk the same
// records.  Each sample's sum runs in the field's cell order, so the
// contribution does not depend on which thread computes it.
__global__ void __launch_bounds__(ADJ_THREADS) k_f64_eval(McD d, BinGeom g, F64Args a,
                                                          const uint32_t *__restrict__ order,
                                                          const uint32_t *__restrict__ keys)
{
    const size_t t = (size_t)blockIdx.x * ADJ_THREADS + threadIdx.x;
    if (t >= g.S) return;
    const uint32_t sid = order[t];
    const uint32_t cells = (uint32_t)a.geom.nx * (uint32_t)a.geom.ny;
    const uint32_t kk = keys[t];
    double2 out = make_double2(0.0, 0.0);
    if (kk < 2u * cells) {
        const int si = kk >= cells ? 1 : 0;
        const double2 y = a.y[sid], r = a.r[sid];
        double om = 0.0, om0 = 0.0;
        const double4 *G64 = a.G64[si];
        auto fn = [&](const double4 &c, double dx, double dy, int k) {
            double w, gr;
            double r2 = dx * dx + dy * dy;
            if (G64) {   // deformed record: |F^-1 d|^2, zero beyond sigma_max h_w
                const double4 G = G64[k];
                if (r2 > G.w * a.h_w * a.h_w) return;
                r2 = fmax(G.x * dx * dx + G.y * dx * dy + G.z * dy * dy, 0.0);
            }
            spline_wg64(r2, a.inv_hw, w, gr);
            om += c.z * w;
            om0 += c.w * w;
        };
        field_visit64(a.geom, a.cell_start[si], a.rec[si], y.x, y.y, fn);
        om *= a.sig2;
        om0 *= a.sig2;
        if (a.om2) {   // the targets' contributions follow in k_f64_reduce5
            a.om2[sid] = make_double2(om, om0);
            return;
        }
        const double r2 = r.x * r.x + r.y * r.y, rn = sqrt(r2);
        if (om0 > 0.0 && rn > 0.0) {
            const double Om0 = *d.set[si].total_dev;
            const double D = a.w1 * om0 / Om0 + a.w2 * a.alpha / rn;
            double ksx, ksy;
            kper_smooth_f64(a.kc.c, KPER_TERMS_ADJ64, a.L, r.x, r.y, ksx, ksy);
            const double f = 1.0 / (2.0 * PI_D * r2);
            const double gk = om / D;
            out = make_double2((-r.y * f + ksx) * gk, (r.x * f + ksy) * gk);
        }
    }
    if (a.om2) a.om2[sid] = make_double2(0.0, 0.0);
    else a.contrib[sid] = out;
}

// one MIS contribution in FP64 (the terms of k_f64_eval)
__device__ __forceinline__ double2 f64_contrib(const F64Args &a, double om, double om0,
                                               double Om0, double rx, double ry)
{
    const double r2 = rx * rx + ry * ry, rn = sqrt(r2);
    if (!(om0 > 0.0 && rn > 0.0)) return make_double2(0.0, 0.0);
    const double D = a.w1 * om0 / Om0 + a.w2 * a.alpha / rn;
    double ksx, ksy;
    kper_smooth_f64(a.kc.c, KPER_TERMS_ADJ64, a.L, rx, ry, ksx, ksy);
    const double f = 1.0 / (2.0 * PI_D * r2);
    const double gk = om / D;
    return make_double2((-ry * f + ksx) * gk, (rx * f + ksy) * gk);
}

// per (owner, sign) with the auxiliary targets: sums in sample order at the
// five targets (a branch-1 sample at every target with r + offset, min
// image; a disk sample at its own); vel[si * n + owner] and
// vel_aux[(si * 4 + t - 1) * n + owner]
__global__ void k_f64_reduce5(McD d, BinGeom g, F64Args a, double2 *__restrict__ vel)
{
    const int pair = blockIdx.x * blockDim.x + threadIdx.x;
    const int n = d.n_owner;
    if (pair >= 2 * n) return;
    const int owner = pair >> 1, si = pair & 1;
    double sx[5] = {0.0, 0.0, 0.0, 0.0, 0.0}, sy[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    const SetD &S = d.set[si];
    const int n_s = S.count_dev ? *S.count_dev : S.n;
    const double L = a.L;
    if (n_s > 0) {
        const double Om0 = *S.total_dev;
        const size_t base = (size_t)pair * g.NS;
        for (uint32_t k = 0; k < g.NS; ++k) {
            const double2 o = a.om2[base + k], r = a.r[base + k];
            if ((int)k < d.N1) {
                for (int t = 0; t < 5; ++t) {
                    double rx = r.x, ry = r.y;
                    if (t) {
                        const float2 of = aux_offset(t, d.aux_dx);
                        rx += (double)of.x;
                        ry += (double)of.y;
                        if (rx > 0.5 * L) rx -= L; else if (rx < -0.5 * L) rx += L;
                        if (ry > 0.5 * L) ry -= L; else if (ry < -0.5 * L) ry += L;
                    }
                    const double2 c = f64_contrib(a, o.x, o.y, Om0, rx, ry);
                    sx[t] += c.x;
                    sy[t] += c.y;
                }
            } else {
                const int kb = (int)k - d.N1;
                const int t = kb < d.N2 ? 0 : 1 + (kb - d.N2) / d.N2;
                const double2 c = f64_contrib(a, o.x, o.y, Om0, r.x, r.y);
                sx[t] += c.x;
                sy[t] += c.y;
            }
        }
    }
    vel[(size_t)si * n + owner] = make_double2(sx[0] * a.inv_tot, sy[0] * a.inv_tot);
    for (int t = 1; t < 5; ++t)
        a.vel_aux[((size_t)si * 4 + (t - 1)) * n + owner] =
            make_double2(sx[t] * a.inv_tot, sy[t] * a.inv_tot);
}

// update_jacobian_from5 + renorm_det1 in FP64 (jacobian_from5's formula)
__global__ void k_f64_jac(int n, const double2 *__restrict__ va, double c,
                          double4 *__restrict__ jac64)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    double2 u[4];
    for (int t = 0; t < 4; ++t) {
        const double2 p = va[(size_t)t * n + j], q = va[(size_t)(4 + t) * n + j];
        u[t] = make_double2(p.x - q.x, p.y - q.y);
    }
    const double fx = 1.0 + (u[1].x - u[0].x) * c, fy = (u[2].x - u[3].x) * c;
    const double fz = (u[1].y - u[0].y) * c, fw = 1.0 + (u[2].y - u[3].y) * c;
    const double det = fx * fw - fy * fz;
    const double sc = 1.0 / sqrt(fmax(det, 1e-12));
    jac64[j] = make_double4(fx * sc, fy * sc, fz * sc, fw * sc);
}

// per (owner, sign): sum in sample order; vel[si * n + owner]
__global__ void k_f64_reduce(int n_owner, uint32_t NS, double inv_tot,
                             const double2 *__restrict__ contrib, double2 *__restrict__ vel)
{
    const int pair = blockIdx.x * blockDim.x + threadIdx.x;
    if (pair >= 2 * n_owner) return;
    double sx = 0.0, sy = 0.0;
    const size_t base = (size_t)pair * NS;
    for (uint32_t k = 0; k < NS; ++k) {
        sx += contrib[base + k].x;
        sy += contrib[base + k].y;
    }
    vel[(size_t)(pair & 1) * n_owner + (pair >> 1)] = make_double2(sx * inv_tot, sy * inv_tot);
}

__global__ void k_f64_advect(int n, double2 *__restrict__ pos, const double2 *__restrict__ vel,
                             double dt, double L, float *__restrict__ pos32)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double2 p = pos[i];
    p.x += dt * (vel[i].x - vel[(size_t)n + i].x);
    p.y += dt * (vel[i].y - vel[(size_t)n + i].y);
    p.x -= floor(p.x / L) * L;
    p.y -= floor(p.y / L) * L;
    pos[i] = p;
    if (pos32) {
        pos32[2 * (size_t)i] = (float)p.x;
        pos32[2 * (size_t)i + 1] = (float)p.y;
    }
}

__global__ void k_d2f_xy(int n, const double2 *__restrict__ p, float2 *__restrict__ q)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) q[i] = make_float2((float)p[i].x, (float)p[i].y);
}

// ---- grid vorticity (loss) ----------------------------------------------------------

__global__ void k_grid_query(FieldD f, float inv_hw, float sig2, int gnx, int gny,
                             float *__restrict__ out)
{
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= gnx * gny) return;
    const int a = idx / gny, b = idx - a * gny;
    const float zx = ((float)a + 0.5f) * (f.g.Lx / (float)gnx);
    const float zy = ((float)b + 0.5f) * (f.g.Ly / (float)gny);
    float acc = 0.f;
    auto fn = [&](const float4 &c, float dx, float dy, int) {
        float w, gr;
        spline_wg(fmaf(dx, dx, dy * dy), inv_hw, w, gr);
        acc = fmaf(c.z, w, acc);
    };
    field_visit(f, zx, zy, fn);
    out[idx] = sig2 * acc;
}

// grid points within h_w of each particle, by index range (periodic)
__global__ void k_grid_adjoint(int n, const float2 *__restrict__ pos, const float *__restrict__ w,
                               float h_w, float Lx, float Ly, int gnx, int gny,
                               const float *__restrict__ gout, float2 *__restrict__ grad_pos,
                               float *__restrict__ grad_w)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float2 p = pos[i];
    const float hx = Lx / (float)gnx, hy = Ly / (float)gny;
    const float inv_hw = 1.f / h_w, sig2 = 40.f / (7.f * PI_F * h_w * h_w);
    const int a0 = (int)ceilf((p.x - h_w) / hx - 0.5f), a1 = (int)floorf((p.x + h_w) / hx - 0.5f);
    const int b0 = (int)ceilf((p.y - h_w) / hy - 0.5f), b1 = (int)floorf((p.y + h_w) / hy - 0.5f);
    float A = 0.f, bx = 0.f, by = 0.f;
    for (int a = a0; a <= a1; ++a) {
        const int aw = ((a % gnx) + gnx) % gnx;
        const float dx = ((float)a + 0.5f) * hx - p.x;
        for (int b = b0; b <= b1; ++b) {
            const int bw = ((b % gny) + gny) % gny;
            const float dy = ((float)b + 0.5f) * hy - p.y;
            float wv, gr;
            spline_wg(fmaf(dx, dx, dy * dy), inv_hw, wv, gr);
            const float go = gout[aw * gny + bw];
            A = fmaf(go, wv, A);
            bx = fmaf(go * gr, dx, bx);
            by = fmaf(go * gr, dy, by);
        }
    }
    grad_w[i] += sig2 * A;
    float2 gp = grad_pos[i];
    const float wi = w[i];
    gp.x -= wi * sig2 * bx;   // d/dx_i W(|z - x_i|) = -W'(r)/r (z - x_i)
    gp.y -= wi * sig2 * by;
    grad_pos[i] = gp;
}

// FP64, brute force over all particles (checks)
__global__ void k_grid_f64(int n, const double2 *__restrict__ pos, const double *__restrict__ w,
                           double h_w, double Lx, double Ly, int gnx, int gny,
                           double *__restrict__ out)
{
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= gnx * gny) return;
    const int a = idx / gny, b = idx - a * gny;
    const double zx = ((double)a + 0.5) * (Lx / gnx), zy = ((double)b + 0.5) * (Ly / gny);
    const double inv_hw = 1.0 / h_w, hw2 = h_w * h_w;
    double acc = 0.0;
    for (int j = 0; j < n; ++j) {
        double dx = zx - pos[j].x, dy = zy - pos[j].y;
        dx -= rint(dx / Lx) * Lx;
        dy -= rint(dy / Ly) * Ly;
        const double r2 = dx * dx + dy * dy;
        if (r2 < hw2) {
            double wv, gr;
            spline_wg64(r2, inv_hw, wv, gr);
            acc += w[j] * wv;
        }
    }
    out[idx] = 40.0 / (7.0 * PI_D * h_w * h_w) * acc;
}

}  // namespace vpmc

using namespace vpmc;

extern "C" {

size_t vpmc_mc_adjoint_workspace_bytes(const vpmc_mc_desc *d)
{
    McD m;
    BinGeom g;
    if (adj_check(d, false, &m, &g) != VPMC_OK) return 0;
    return adj_ws_bytes(g, d->n_owner, false);
}

int vpmc_mc_adjoint(const vpmc_mc_desc *d, const int *sub_pos, const int *sub_neg,
                    const float *adj_vel, float *grad_pos, float *grad_w, void *workspace,
                    size_t workspace_bytes, void *stream)
{
    return mc_adjoint_impl(d, false, sub_pos, sub_neg, adj_vel, nullptr, grad_pos, grad_w,
                           nullptr, workspace, workspace_bytes, as_stream(stream));
}

size_t vpmc_step_adjoint_workspace_bytes(const vpmc_mc_desc *d)
{
    if (!d) return 0;
    const size_t a = vpmc_mc_adjoint_workspace_bytes(d);
    if (a == 0) return 0;
    const int m = d->set[0].n > d->set[1].n ? d->set[0].n : d->set[1].n;
    size_t f = vpmc_field_workspace_bytes(m, d->set[0].field.nx, d->set[0].field.ny);
    const size_t n = (size_t)(d->n_owner > 0 ? d->n_owner : 1);
    return align256(n * 8) + (a > f ? a : f);
}

int vpmc_step_adjoint(vpmc_mc_desc *d, int n_total, const float *pos_xy, const float *omega,
                      const int *sub_pos, const int *sub_neg, float dt, float *lambda,
                      float *grad_w, void *workspace, size_t workspace_bytes, void *stream)
{
    if (!d || n_total < 0 || d->n_owner != n_total || d->owner_xy != pos_xy) return VPMC_ERR_ARG;
    if (n_total == 0) return VPMC_OK;
    if (!lambda || !grad_w || !omega) return VPMC_ERR_ARG;
    hipStream_t s = as_stream(stream);
    Carve cv(workspace, workspace_bytes);
    float2 *av = cv.take<float2>((size_t)n_total);
    if (!av) return VPMC_ERR_WORKSPACE;
    char *rest = cv.base + cv.used;
    const size_t rest_bytes = cv.cap - cv.used;
    const int *subs[2] = {sub_pos, sub_neg};
    for (int k = 0; k < 2; ++k) {
        vpmc_sign_set &S = d->set[k];
        VPMC_TRY(field_build_impl(S.field.nx, S.field.ny, S.field.Lx, S.field.Ly, S.n, nullptr,
                                  subs[k], pos_xy, omega, 1, const_cast<float *>(S.field.pts),
                                  const_cast<int *>(S.field.cell_start),
                                  const_cast<int *>(S.field.order),
                                  const_cast<float *>(S.set_xy), rest, rest_bytes, s));
    }
    hipLaunchKernelGGL(k_scale_xy, dim3((unsigned)((n_total + 255) / 256)), dim3(256), 0, s,
                       n_total, (const float2 *)lambda, dt, av);
    VPMC_CHECK_LAUNCH();
    return mc_adjoint_impl(d, false, sub_pos, sub_neg, (const float *)av, nullptr, lambda,
                           grad_w, nullptr, rest, rest_bytes, s);
}

size_t vpmc_step_adjoint_deformed_workspace_bytes(const vpmc_mc_desc *d)
{
    McD m;
    BinGeom g;
    if (adj_check(d, true, &m, &g) != VPMC_OK) return 0;
    const size_t a = adj_ws_bytes(g, d->n_owner, true);
    const size_t e = vpmc_mc_binned_workspace_bytes(d);
    if (e == 0) return 0;
    const int mx = d->set[0].n > d->set[1].n ? d->set[0].n : d->set[1].n;
    const size_t f = vpmc_field_workspace_bytes(mx, d->set[0].field.nx, d->set[0].field.ny);
    const size_t n = (size_t)(d->n_owner > 0 ? d->n_owner : 1);
    size_t rest = a > e ? a : e;
    if (f > rest) rest = f;
    return align256(n * 8) + align256(4 * n * 8) + rest;
}

int vpmc_step_adjoint_deformed(vpmc_mc_desc *d, int n_total, const float *pos_xy,
                               const float *omega, const int *sub_pos, const int *sub_neg,
                               float dt, float *lambda, float *jac_bar, float *grad_w,
                               void *workspace, size_t workspace_bytes, void *stream)
{
    if (!d || n_total < 0 || d->n_owner != n_total || d->owner_xy != pos_xy) return VPMC_ERR_ARG;
    McD m;
    BinGeom g;
    VPMC_TRY(adj_check(d, true, &m, &g));
    if (n_total == 0) return VPMC_OK;
    if (!lambda || !jac_bar || !grad_w || !omega || !d->vel_sign) return VPMC_ERR_ARG;
    hipStream_t s = as_stream(stream);
    Carve cv(workspace, workspace_bytes);
    float2 *av = cv.take<float2>((size_t)n_total);
    float2 *aa = cv.take<float2>(4 * (size_t)n_total);
    if (!av || !aa) return VPMC_ERR_WORKSPACE;
    char *rest = cv.base + cv.used;
    const size_t rest_bytes = cv.cap - cv.used;
    const int *subs[2] = {sub_pos, sub_neg};
    for (int k = 0; k < 2; ++k) {   // the fields with this step's F, as vpmc_step
        vpmc_sign_set &S = d->set[k];
        VPMC_TRY(field_build_impl(S.field.nx, S.field.ny, S.field.Lx, S.field.Ly, S.n, nullptr,
                                  subs[k], pos_xy, omega, 1, const_cast<float *>(S.field.pts),
                                  const_cast<int *>(S.field.cell_start),
                                  const_cast<int *>(S.field.order),
                                  const_cast<float *>(S.set_xy), rest, rest_bytes, s, d->jac,
                                  const_cast<float *>(S.field.gq),
                                  const_cast<float *>(S.set_jac)));
    }
    // this step's five estimates (the aux ones set F'), then F''s adjoint
    VPMC_TRY(vpmc_mc_estimate_binned(d, rest, rest_bytes, stream));
    hipLaunchKernelGGL(k_jac_adj, dim3((unsigned)((n_total + 255) / 256)), dim3(256), 0, s,
                       n_total, (const float2 *)d->vel_aux, 0.5f * dt / d->aux_dx,
                       (float4 *)jac_bar, aa);
    hipLaunchKernelGGL(k_scale_xy, dim3((unsigned)((n_total + 255) / 256)), dim3(256), 0, s,
                       n_total, (const float2 *)lambda, dt, av);
    VPMC_CHECK_LAUNCH();
    return mc_adjoint_impl(d, true, sub_pos, sub_neg, (const float *)av, (const float *)aa,
                           lambda, grad_w, jac_bar, rest, rest_bytes, s);
}

size_t vpmc_step_f64_workspace_bytes(const vpmc_mc_desc *d)
{
    McD m;
    if (!d || make_mcd(d, &m) != VPMC_OK) return 0;
    BinGeom g;
    if (bin_geom(m, d->N1, d->N2, d->n_owner, &g) != VPMC_OK) return 0;
    const size_t n = (size_t)(d->n_owner > 0 ? d->n_owner : 1);
    const int mx = d->set[0].n > d->set[1].n ? d->set[0].n : d->set[1].n;
    const size_t nm = (size_t)(mx > 0 ? mx : 1);
    const size_t S = (size_t)(g.S > 0 ? g.S : 1);
    const size_t cells = (size_t)d->set[0].field.nx * (size_t)d->set[0].field.ny;
    const size_t fb = vpmc_field_workspace_bytes(mx, d->set[0].field.nx, d->set[0].field.ny);
    const size_t rb = radix_workspace_bytes((int)S, 4);
    // the deformed step's quadratic forms, F per set and aux estimates
    const size_t def = d->aux ? 4 * align256(nm * 32) + align256(8 * n * 16) : 0;
    return align256(n * 8) + 2 * align256(nm * 32) + 2 * align256(nm * 16) +
           3 * align256(S * 16) + 3 * align256(S * 4) + align256((2 * cells + 2) * 4) +
           align256(2 * n * 16) + (fb > rb ? fb : rb) + def + 4096;
}

// jac64 (deformed step, aux targets): F per particle in FP64, replaced by
// the F' of this step's five estimates before the move, as vpmc_step does
static int step_f64_impl(vpmc_mc_desc *d, int n_total, double *pos_xy, const double *w,
                         const float *omega0, const int *sub_pos, const int *sub_neg, double dt,
                         double *jac64, void *workspace, size_t workspace_bytes, hipStream_t s)
{
    if (!d || n_total < 0 || d->n_owner != n_total) return VPMC_ERR_ARG;
    if (d->mode != VPMC_MODE_CONSISTENT || !d->periodic || d->uniforms || d->Lx != d->Ly)
        return VPMC_ERR_ARG;
    const bool def = jac64 != nullptr;
    if (def != (d->aux != 0)) return VPMC_ERR_ARG;   // aux targets exactly with F
    if (def && (!(d->aux_dx > 0.f) || !d->set[0].field.gq || !d->set[1].field.gq))
        return VPMC_ERR_ARG;   // the fields' search radius covers the deformation
    if (n_total == 0) return VPMC_OK;
    if (!pos_xy || !w || !omega0 || !sub_pos || !sub_neg) return VPMC_ERR_ARG;
    McD m;
    VPMC_TRY(make_mcd(d, &m));
    BinGeom g;
    VPMC_TRY(bin_geom(m, d->N1, d->N2, d->n_owner, &g));
    Carve cv(workspace, workspace_bytes);
    const int mx = d->set[0].n > d->set[1].n ? d->set[0].n : d->set[1].n;
    const size_t nm = (size_t)(mx > 0 ? mx : 1);
    float2 *pos32 = cv.take<float2>((size_t)n_total);
    double4 *rec[2] = {cv.take<double4>(nm), cv.take<double4>(nm)};
    double2 *set64[2] = {cv.take<double2>(nm), cv.take<double2>(nm)};
    const size_t S = (size_t)(g.S > 0 ? g.S : 1);
    double2 *contrib = cv.take<double2>(S);
    double2 *ys = cv.take<double2>(S);
    double2 *rs = cv.take<double2>(S);
    uint32_t *key = cv.take<uint32_t>(S);
    uint32_t *order = cv.take<uint32_t>(S);
    uint32_t *key_sorted = cv.take<uint32_t>(S);
    const int fnx = d->set[0].field.nx, fny = d->set[0].field.ny;
    const size_t cells = (size_t)fnx * (size_t)fny;
    int *cstart = cv.take<int>(2 * cells + 2);
    double2 *vel = cv.take<double2>(2 * (size_t)n_total);
    double4 *G64[2] = {nullptr, nullptr}, *sj64[2] = {nullptr, nullptr};
    double2 *vaux = nullptr;
    if (def) {
        for (int k = 0; k < 2; ++k) {
            G64[k] = cv.take<double4>(nm);
            sj64[k] = cv.take<double4>(nm);
        }
        vaux = cv.take<double2>(8 * (size_t)n_total);
        if (!G64[0] || !G64[1] || !sj64[0] || !sj64[1] || !vaux) return VPMC_ERR_WORKSPACE;
    }
    const size_t fb = vpmc_field_workspace_bytes(mx, fnx, fny);
    const size_t rb = radix_workspace_bytes((int)S, 4);
    void *fws = cv.take<char>(fb > rb ? fb : rb);   // the field builds, then the sort
    if (!pos32 || !rec[0] || !rec[1] || !set64[0] || !set64[1] || !contrib || !ys || !rs ||
        !key || !order || !key_sorted || !cstart || !vel || !fws)
        return VPMC_ERR_WORKSPACE;
    if (2 * cells + 1 > 0x7fffffffu) return VPMC_ERR_ARG;
    const unsigned gn = (unsigned)((n_total + 255) / 256);
    hipLaunchKernelGGL(k_d2f_xy, dim3(gn), dim3(256), 0, s, n_total, (const double2 *)pos_xy,
                       pos32);
    VPMC_CHECK_LAUNCH();
    const int *subs[2] = {sub_pos, sub_neg};
    for (int k = 0; k < 2; ++k) {
        vpmc_sign_set &S = d->set[k];
        VPMC_TRY(field_build_impl(S.field.nx, S.field.ny, S.field.Lx, S.field.Ly, S.n, nullptr,
                                  subs[k], (const float *)pos32, omega0, 1,
                                  const_cast<float *>(S.field.pts),
                                  const_cast<int *>(S.field.cell_start),
                                  const_cast<int *>(S.field.order),
                                  const_cast<float *>(S.set_xy), fws, fb, s));
        if (S.n > 0) {
            hipLaunchKernelGGL(k_f64_recs, dim3((unsigned)((S.n + 255) / 256)), dim3(256), 0, s,
                               S.n, S.count_dev, S.field.order, subs[k], (const double2 *)pos_xy,
                               w, omega0, k ? -1.0 : 1.0, rec[k], set64[k],
                               (const double4 *)jac64, G64[k], sj64[k]);
            VPMC_CHECK_LAUNCH();
        }
    }
    F64Args a;
    a.pos = (const double2 *)pos_xy;
    for (int k = 0; k < 2; ++k) {
        a.rec[k] = rec[k];
        a.set64[k] = set64[k];
        a.cell_start[k] = d->set[k].field.cell_start;
        a.G64[k] = G64[k];
        a.setjac64[k] = sj64[k];
    }
    // the sample generation reads F through the FP64 copies only
    m.set[0].set_jac = m.set[1].set_jac = nullptr;
    a.om2 = def ? contrib : nullptr;   // (om, om0) in place of the contributions
    a.vel_aux = vaux;
    a.aux_dx = (double)d->aux_dx;
    a.geom = m.set[0].field.g;
    a.h_w = (double)d->h_w;
    a.inv_hw = 1.0 / a.h_w;
    a.sig2 = 40.0 / (7.0 * PI_D * a.h_w * a.h_w);
    a.L = (double)d->Lx;
    a.Rc = 0.5 * sqrt(2.0) * a.L;
    a.alpha = 1.0 / (2.0 * PI_D * a.Rc);
    a.w1 = (double)d->N1 / (double)(d->N1 + d->N2);
    a.w2 = (double)d->N2 / (double)(d->N1 + d->N2);
    a.inv_tot = 1.0 / (double)(d->N1 + d->N2);
    a.contrib = contrib;
    a.y = ys;
    a.r = rs;
    const double *kc = kper_coeffs_f64_host(KPER_TERMS_ADJ64);
    for (int i = 0; i < KPER_TERMS_ADJ64; ++i) a.kc.c[i] = kc[i];
    const unsigned gs = (unsigned)((g.S + ADJ_THREADS - 1) / ADJ_THREADS);
    hipLaunchKernelGGL(k_f64_gen, dim3(gs), dim3(ADJ_THREADS), 0, s, m, g, a, key);
    VPMC_CHECK_LAUNCH();
    const uint32_t dead = 2u * (uint32_t)cells;
    VPMC_TRY(radix_sort_indices(key, (int)g.S, key_bits_for(dead), order, key_sorted,
                                (int)dead, cstart, fws, fb > rb ? fb : rb, s));
    hipLaunchKernelGGL(k_f64_eval, dim3(gs), dim3(ADJ_THREADS), 0, s, m, g, a,
                       (const uint32_t *)order, (const uint32_t *)key_sorted);
    VPMC_CHECK_LAUNCH();
    const unsigned gp = (unsigned)((2 * (size_t)n_total + 255) / 256);
    if (def) {
        hipLaunchKernelGGL(k_f64_reduce5, dim3(gp), dim3(256), 0, s, m, g, a, vel);
        hipLaunchKernelGGL(k_f64_jac, dim3(gn), dim3(256), 0, s, n_total,
                           (const double2 *)vaux, 0.5 * dt / (double)d->aux_dx,
                           (double4 *)jac64);
    } else {
        hipLaunchKernelGGL(k_f64_reduce, dim3(gp), dim3(256), 0, s, n_total, g.NS, a.inv_tot,
                           (const double2 *)contrib, vel);
    }
    hipLaunchKernelGGL(k_f64_advect, dim3(gn), dim3(256), 0, s, n_total, (double2 *)pos_xy,
                       (const double2 *)vel, dt, a.L, (float *)nullptr);
    VPMC_CHECK_LAUNCH();
    return VPMC_OK;
}

int vpmc_step_f64(vpmc_mc_desc *d, int n_total, double *pos_xy, const double *w,
                  const float *omega0, const int *sub_pos, const int *sub_neg, double dt,
                  void *workspace, size_t workspace_bytes, void *stream)
{
    if (d && d->aux) return VPMC_ERR_ARG;   // the deformed step: vpmc_step_f64_deformed
    return step_f64_impl(d, n_total, pos_xy, w, omega0, sub_pos, sub_neg, dt, nullptr,
                         workspace, workspace_bytes, as_stream(stream));
}

int vpmc_step_f64_deformed(vpmc_mc_desc *d, int n_total, double *pos_xy, const double *w,
                           const float *omega0, const int *sub_pos, const int *sub_neg,
                           double dt, double *jac64, void *workspace, size_t workspace_bytes,
                           void *stream)
{
    if (!jac64) return VPMC_ERR_ARG;
    return step_f64_impl(d, n_total, pos_xy, w, omega0, sub_pos, sub_neg, dt, jac64, workspace,
                         workspace_bytes, as_stream(stream));
}

size_t vpmc_grid_vorticity_workspace_bytes(int n, int nx, int ny)
{
    const size_t m = (size_t)(n > 0 ? n : 1);
    return align256(m * 16) + align256(((size_t)nx * ny + 1) * 4) +
           vpmc_field_workspace_bytes(n, nx, ny) + 1024;
}

int vpmc_grid_vorticity(int n, const float *pos_xy, const float *w, float h_w, float Lx,
                        float Ly, int cells_per_hw, int gnx, int gny, float *out, void *workspace,
                        size_t workspace_bytes, void *stream)
{
    if (n < 0 || gnx <= 0 || gny <= 0 || !out || (n > 0 && (!pos_xy || !w))) return VPMC_ERR_ARG;
    hipStream_t s = as_stream(stream);
    int nx, ny;
    VPMC_TRY(vpmc_field_grid(h_w, Lx, Ly, cells_per_hw, &nx, &ny));
    Carve cv(workspace, workspace_bytes);
    float *pts = cv.take<float>(4 * (size_t)(n > 0 ? n : 1));
    int *cs = cv.take<int>((size_t)nx * ny + 1);
    const size_t fb = vpmc_field_workspace_bytes(n, nx, ny);
    void *fws = cv.take<char>(fb);
    if (!pts || !cs || !fws) return VPMC_ERR_WORKSPACE;
    VPMC_TRY(field_build_impl(nx, ny, Lx, Ly, n, nullptr, nullptr, pos_xy, w, 0, pts, cs, nullptr,
                              nullptr, fws, fb, s));
    vpmc_field fh{nx, ny, Lx, Ly, n, pts, cs, nullptr};
    FieldGeom g;
    VPMC_TRY(field_geom(&fh, h_w, &g));
    FieldD f{(const float4 *)pts, cs, g};
    hipLaunchKernelGGL(k_grid_query, dim3((unsigned)((gnx * gny + 255) / 256)), dim3(256), 0, s, f,
                       1.f / h_w, 40.f / (7.f * PI_F * h_w * h_w), gnx, gny, out);
    VPMC_CHECK_LAUNCH();
    return VPMC_OK;
}

int vpmc_grid_vorticity_adjoint(int n, const float *pos_xy, const float *w, float h_w, float Lx,
                                float Ly, int gnx, int gny, const float *grad_out,
                                float *grad_pos, float *grad_w, void *stream)
{
    if (n < 0 || gnx <= 0 || gny <= 0 || !(h_w > 0.f) || 2.f * h_w >= Lx || 2.f * h_w >= Ly)
        return VPMC_ERR_ARG;
    if (n == 0) return VPMC_OK;
    if (!pos_xy || !w || !grad_out || !grad_pos || !grad_w) return VPMC_ERR_ARG;
    hipLaunchKernelGGL(k_grid_adjoint, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                       as_stream(stream), n, (const float2 *)pos_xy, w, h_w, Lx, Ly, gnx, gny,
                       grad_out, (float2 *)grad_pos, grad_w);
    VPMC_CHECK_LAUNCH();
    return VPMC_OK;
}

int vpmc_grid_vorticity_f64(int n, const double *pos_xy, const double *w, double h_w, double Lx,
                            double Ly, int gnx, int gny, double *out, void *stream)
{
    if (n < 0 || gnx <= 0 || gny <= 0 || !out || (n > 0 && (!pos_xy || !w))) return VPMC_ERR_ARG;
    hipLaunchKernelGGL(k_grid_f64, dim3((unsigned)((gnx * gny + 255) / 256)), dim3(256), 0,
                       as_stream(stream), n, (const double2 *)pos_xy, w, h_w, Lx, Ly, gnx, gny,
                       out);
    VPMC_CHECK_LAUNCH();
    return VPMC_OK;
}

}  // extern "C"
