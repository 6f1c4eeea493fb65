/*
 * vpmc_oracle.c -- CPU restatement of the reference's vortex-particle hot path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker).  Only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg may load this library; the product library
 * libvpmc.so never links or calls it.
 *
 * Pinning (each function is checked by something other than a retyped
 * formula; the checks live in tests/test_oracle.py):
 *   orc_spline_*        closed forms: int W dA = 1, continuity at q = 1/2,
 *                       W(0) = 40/(7 pi h^2); and oracle/_ref (the reference's
 *                       own cubicSplinePDF compiled from its sources).
 *   orc_mass_*          numerical quadrature of 2 pi r W(r); inverse round trip.
 *   orc_fill_cells /    oracle/_ref (FillCells_cpu / CountingSort_cpu) bit-exact,
 *   orc_counting_sort   std::stable_sort by key, the reference's invariants.
 *   orc_eval_vorticity  oracle/_ref evalVorticityAbs_cpu, brute force min-image.
 *   orc_build_subsets   numpy cumsum in FP64, invariants cdf[last] = 1.
 *   orc_sample_idx      numpy.searchsorted(side='right').
 *   orc_philox4x32_10   Random123 known-answer vectors and rocRAND's engine
 *                       (tests/golden/philox_kat.json, written by
 *                       tests/golden/make_philox_vectors.cpp).
 *   orc_kper_* /        periodicity of K_per across the cell edges, the
 *   orc_kernel_smooth_d Eisenstein closed form G4 = Gamma(1/4)^8/(960 pi^2),
 *                       brute-force lattice sums, and convergence of
 *                       orc_direct_blob to the analytic Taylor-Green velocity.
 *   orc_eval_vorticity  with F (the CUDA query's deformed kernel): F = I and
 *   (deformed)          rotations equal the undeformed FP64 brute force; one
 *                       blob at x_j + F u equals |w| W(|u|) (closed form).
 *   orc_mc_velocity     (reference mode) oracle/_ref monte_carlo_bs_cpu replayed
 *                       on the same std::mt19937 draws; (consistent mode) its
 *                       Monte Carlo mean converges to orc_direct_blob.
 *   orc_direct_eps      brute force in numpy; odd symmetry.
 *   orc_cg_rbf          dense numpy solve and the Taylor-Green lattice closed
 *                       form w = b / lambda_11.
 */
#include "vpmc_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* include/utilities.hpp:5 redefines M_PI as a float literal; the reference's
   src/cpu files see that float value. */
#define PI_F 3.14159265358979323846f
#define PI_D 3.14159265358979323846

/* ------------------------------------------------------------------------- */
/* include/utilities.hpp:7-17                                                 */
float orc_spline_f(float r, float h_w)
{
    const float pi = 3.141592653589793f;
    const float sig2 = 40.f / (7.f * pi * h_w * h_w);
    float q = r / h_w;
    if (q >= 0.f && q <= 0.5f) return sig2 * (6.f * (q * q * q - q * q) + 1.f);
    if (q > 0.5f && q < 1.f) return sig2 * 2.f * (1.f - q) * (1.f - q) * (1.f - q);
    return 0.f;
}

double orc_spline_d(double r, double h_w)
{
    const double sig2 = 40.0 / (7.0 * PI_D * h_w * h_w);
    double q = r / h_w;
    if (q >= 0.0 && q <= 0.5) return sig2 * (6.0 * (q * q * q - q * q) + 1.0);
    if (q > 0.5 && q < 1.0) return sig2 * 2.0 * (1.0 - q) * (1.0 - q) * (1.0 - q);
    return 0.0;
}

/* Fraction of W's mass within radius q*h_w: (80/7) int_0^q s f(s) ds with
   f the normalised spline; 7/80 is the full radial integral. */
double orc_mass_fraction(double q)
{
    if (q <= 0.0) return 0.0;
    if (q >= 1.0) return 1.0;
    double q2 = q * q;
    if (q <= 0.5) return (80.0 / 7.0) * (0.5 * q2 - 1.5 * q2 * q2 + 1.2 * q2 * q2 * q);
    return (80.0 / 7.0) * (q2 - 2.0 * q2 * q + 1.5 * q2 * q2 - 0.4 * q2 * q2 * q - 0.0125);
}

/* Inverse of orc_mass_fraction by bracketed Newton (FP64).  Inner piece
   q <= 1/2 solves q^2/2 - 1.5 q^4 + 1.2 q^5 = 7u/80; the outer piece is
   written in t = 1 - q: t^4/2 - 0.4 t^5 = 7(1-u)/80, well conditioned as
   u -> 1.  The HIP library runs the same iteration. */
double orc_mass_inverse(double u)
{
    if (u <= 0.0) return 0.0;
    if (u >= 1.0) return 1.0;
    const double u_half = 11.0 / 14.0;   /* orc_mass_fraction(0.5) */
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
    } else {
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
}

/* ------------------------------------------------------------------------- */
/* src/cpu/hash.cpp:6-34                                                      */
void orc_fill_cells(int rshx, int rshy, float Lx, float Ly, int n, const int *sub,
                    const float *px, const float *py, int *cell_num, int *cell_acc,
                    int *vp_cell)
{
    int cell_cnt = rshx * rshy;
    memset(cell_num, 0, sizeof(int) * (size_t)cell_cnt);
    const float inv_hx = rshx / Lx;
    const float inv_hy = rshy / Ly;
    for (int i = 0; i < n; ++i) {
        int j = sub[i];
        int ix = (int)floorf(px[j] * inv_hx);
        int iy = (int)floorf(py[j] * inv_hy);
        ix = ix < 0 ? 0 : (ix > rshx - 1 ? rshx - 1 : ix);
        iy = iy < 0 ? 0 : (iy > rshy - 1 ? rshy - 1 : iy);
        int cell = ix * rshy + iy;
        vp_cell[i] = cell;
        ++cell_num[cell];
    }
    cell_acc[0] = 0;
    for (int c = 1; c < cell_cnt; ++c) cell_acc[c] = cell_acc[c - 1] + cell_num[c - 1];
}

/* src/cpu/hash.cpp:36-47 -- reverse walk gives ascending index within a cell */
void orc_counting_sort(int n, int *cell_num, const int *cell_acc, const int *vp_cell,
                       int *vp_map)
{
    for (int i = n - 1; i >= 0; --i) {
        int c = vp_cell[i];
        int dst = --cell_num[c] + cell_acc[c];
        vp_map[dst] = i;
    }
}

/* ------------------------------------------------------------------------- */
/* src/cpu/evaluate_vorticity.cpp:8-75                                        */
float orc_eval_vorticity(const orc_view *v, float qx, float qy, int sign, int use_signed,
                         int exclude_center, float h_w, int periodic)
{
    float px = 0.f, py = 0.f;
    if (periodic) {
        px = qx - floorf(qx / v->Lx) * v->Lx;
        py = qy - floorf(qy / v->Ly) * v->Ly;
    } else {
        if (qx < 0.f || qx > v->Lx || qy < 0.f || qy > v->Ly) return 0.f;
        /* src/cpu/evaluate_vorticity.cpp:16 leaves the point at {0,0} on this
           branch (a defect); the CUDA query include/vorticity_query.cuh:26 uses
           p, which is restated here. */
        px = qx; py = qy;
    }
    float sum = 0.f;
#define ORC_VISIT(I)                                                                \
    do {                                                                            \
        float dx = v->px[I] - px, dy = v->py[I] - py;                               \
        if (periodic && !v->vp_sort) {                                              \
            /* naive view: rshx = 0, the fold below would divide by zero    \
               (src/cpu/evaluate_vorticity.cpp:30-37); plain minimum image */ \
            if (dx > 0.5f * v->Lx) dx -= v->Lx;                                     \
            if (dx < -0.5f * v->Lx) dx += v->Lx;                                    \
            if (dy > 0.5f * v->Ly) dy -= v->Ly;                                     \
            if (dy < -0.5f * v->Ly) dy += v->Ly;                                    \
        } else if (periodic) {                                                      \
            const float cellx = v->Lx / v->rshx, celly = v->Ly / v->rshy;           \
            if (dx > v->Lx - 2 * cellx) dx -= v->Lx;                                \
            if (dx < -v->Lx + 2 * cellx) dx += v->Lx;                               \
            if (dy > v->Ly - 2 * celly) dy -= v->Ly;                                \
            if (dy < -v->Ly + 2 * celly) dy += v->Ly;                               \
        }                                                                           \
        float r2 = dx * dx + dy * dy;                                               \
        if (!(r2 > h_w * h_w || (exclude_center && r2 < 1e-12f))) {                 \
            float wv = v->w[I];                                                     \
            float wval = use_signed ? wv : fabsf(wv);                               \
            if (!((sign > 0 && wv <= 0) || (sign < 0 && wv >= 0)))                  \
                sum += wval * orc_spline_f(sqrtf(r2), h_w);                         \
        }                                                                           \
    } while (0)

    /* the CUDA step's deformed kernel: d = x_I - p by minimum image,
       dloc = inv(F_I) d, rad = |dloc|, skipped beyond 2 h_w
       (include/vorticity_query.cuh:52-76) */
#define ORC_VISIT_F(I)                                                              \
    do {                                                                            \
        const float *Fi = v->F + 4 * (size_t)(I);                                   \
        float det = Fi[0] * Fi[3] - Fi[1] * Fi[2];                                  \
        float dx = v->px[I] - px, dy = v->py[I] - py;                               \
        if (periodic) {                                                             \
            if (dx > 0.5f * v->Lx) dx -= v->Lx;                                     \
            if (dx < -0.5f * v->Lx) dx += v->Lx;                                    \
            if (dy > 0.5f * v->Ly) dy -= v->Ly;                                     \
            if (dy < -0.5f * v->Ly) dy += v->Ly;                                    \
        }                                                                           \
        float lx = (Fi[3] * dx - Fi[1] * dy) / det;                                 \
        float ly = (-Fi[2] * dx + Fi[0] * dy) / det;                                \
        float rad = hypotf(lx, ly);                                                 \
        if (!(rad > 2.f * h_w) && !(exclude_center && rad < 1e-6f)) {               \
            float wv = v->w[I];                                                     \
            float wval = use_signed ? wv : fabsf(wv);                               \
            if (!((sign > 0 && wv <= 0) || (sign < 0 && wv >= 0)))                  \
                sum += wval * orc_spline_f(rad, h_w);                               \
        }                                                                           \
    } while (0)

    if (v->F && v->vp_sort && v->cell_acc) {   /* include/vorticity_query.cuh:31-49 */
        int cx = (int)(px / v->Lx * v->rshx);
        int cy = (int)(py / v->Ly * v->rshy);
        if (cx > v->rshx - 1) cx = v->rshx - 1;
        if (cy > v->rshy - 1) cy = v->rshy - 1;
        for (int ox = -1; ox <= 1; ++ox)
            for (int oy = -1; oy <= 1; ++oy) {
                int gx = cx + ox, gy = cy + oy;
                if (periodic) {
                    gx = (gx % v->rshx + v->rshx) % v->rshx;
                    gy = (gy % v->rshy + v->rshy) % v->rshy;
                } else if (gx < 0 || gy < 0 || gx >= v->rshx || gy >= v->rshy) {
                    continue;
                }
                int c = gx * v->rshy + gy;
                int beg = v->cell_acc[c];
                int end = (c + 1 < v->rshx * v->rshy) ? v->cell_acc[c + 1] : v->cnt_view;
                for (int k = beg; k < end; ++k) {
                    int I = v->sub[v->vp_sort[k]];
                    ORC_VISIT_F(I);
                }
            }
    } else if (v->F) {
        for (int I = 0; I < v->N; ++I) ORC_VISIT_F(I);
    } else if (v->vp_sort && v->cell_acc) {
        int cx = (int)(px * v->rshx / v->Lx);
        int cy = (int)(py * v->rshy / v->Ly);
        if (cx > v->rshx - 1) cx = v->rshx - 1;
        if (cy > v->rshy - 1) cy = v->rshy - 1;
        for (int ox = -1; ox <= 1; ++ox)
            for (int oy = -1; oy <= 1; ++oy) {
                int gx = cx + ox, gy = cy + oy;
                if (periodic) {
                    gx = (gx % v->rshx + v->rshx) % v->rshx;
                    gy = (gy % v->rshy + v->rshy) % v->rshy;
                } else if (gx < 0 || gy < 0 || gx >= v->rshx || gy >= v->rshy) {
                    continue;
                }
                int c = gx * v->rshy + gy;
                int beg = v->cell_acc[c];
                int end = (c + 1 < v->rshx * v->rshy) ? v->cell_acc[c + 1] : v->cnt_view;
                for (int k = beg; k < end; ++k) {
                    int I = v->sub[v->vp_sort[k]];
                    ORC_VISIT(I);
                }
            }
    } else {
        for (int I = 0; I < v->N; ++I) ORC_VISIT(I);
    }
#undef ORC_VISIT
#undef ORC_VISIT_F
    return sum;
}

/* ------------------------------------------------------------------------- */
/* src/cpu/build_particle_sets.cpp:6-42, prefix sums and totals in FP64       */
void orc_build_subsets(int n, const float *omega, int *sub_pos, int *sub_neg,
                       double *cdf_pos, double *cdf_neg, int *cnt, double *cum)
{
    int np = 0, nn = 0;
    double cp = 0.0, cn = 0.0;
    for (int i = 0; i < n; ++i) {
        float w = omega[i];
        if (w > 0) { sub_pos[np] = i; cp += (double)w; cdf_pos[np++] = cp; }
        else if (w < 0) { sub_neg[nn] = i; cn -= (double)w; cdf_neg[nn++] = cn; }
    }
    for (int j = 0; j < np; ++j) cdf_pos[j] /= cp;
    for (int j = 0; j < nn; ++j) cdf_neg[j] /= cn;
    cnt[0] = np; cnt[1] = nn; cum[0] = cp; cum[1] = cn;
}

/* src/cpu/monte_carlo_sampling.cpp:23-33: std::upper_bound, j == N -> N-1 */
int orc_sample_idx(float u, const double *cdf, int n)
{
    if (n <= 0) return -1;
    int lo = 0, hi = n;
    double ud = (double)u;
    while (lo < hi) {
        int mid = lo + ((hi - lo) >> 1);
        if (ud < cdf[mid]) hi = mid; else lo = mid + 1;
    }
    return lo == n ? n - 1 : lo;
}

/* ------------------------------------------------------------------------- */
/* Philox4x32-10, Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as
   easy as 1, 2, 3", SC'11: multipliers 0xD2511F53 / 0xCD9E8D57, Weyl key
   increments 0x9E3779B9 / 0xBB67AE85, ten rounds.                           */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        uint32_t n1 = (uint32_t)p1;
        uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        uint32_t n3 = (uint32_t)p0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
        k0 += 0x9E3779B9u; k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

float orc_u01(uint32_t x) { return (float)(x >> 8) * 0x1p-24f; }

/* ------------------------------------------------------------------------- */
/* Doubly periodic Biot-Savart kernel on the square lattice L(Z + iZ).
   Per unit circulation, u - i v = F(z) / (2 pi i) with
   F(z) = zeta(z) - (pi / L^2) conj(z)  (Weierstrass zeta, quasi-periods
   eta1 = pi/L, eta2 = -i pi/L by the Legendre relation), doubly periodic and
   equal to 1/z + (smooth, harmonic part) near 0.  The smooth part uses the
   Laurent series zeta(z) - 1/z = -sum_m G_{4m} z^{4m-1}; on the square lattice
   only G_{4m} survive.  G_4 from the Eisenstein q-series
   G_4 = (pi^4/45) E_4(i), E_4 = 1 + 240 sum sigma_3(n) q^n, q = exp(-2 pi);
   G_{4m}, m >= 2, from the absolutely convergent lattice sum. */
void orc_kper_coeffs(int nterms, double *c)
{
    if (nterms <= 0) return;
    {
        double q = exp(-2.0 * PI_D), qn = 1.0, s = 0.0;
        for (int n = 1; n <= 12; ++n) {
            qn *= q;
            double sig = 0.0;
            for (int d = 1; d <= n; ++d)
                if (n % d == 0) sig += (double)d * d * d;
            s += sig * qn;
        }
        double pi4 = PI_D * PI_D * PI_D * PI_D;
        c[0] = (pi4 / 45.0) * (1.0 + 240.0 * s);
    }
    for (int m = 2; m <= nterms; ++m) {
        int R = (m == 2) ? 600 : 160;
        double acc = 0.0;
        /* one representative of each rotation orbit: a >= 1, b >= 0 */
        for (int a = 1; a <= R; ++a)
            for (int b = 0; b <= R; ++b) {
                /* (a + i b)^(-4m) = conj(a+ib)^(4m) / (a^2+b^2)^(4m) */
                double re = a, im = -(double)b;
                double pr = 1.0, pi_ = 0.0;
                /* z^4 then power m */
                double z2r = re * re - im * im, z2i = 2 * re * im;
                double z4r = z2r * z2r - z2i * z2i, z4i = 2 * z2r * z2i;
                for (int k = 0; k < m; ++k) {
                    double t = pr * z4r - pi_ * z4i;
                    pi_ = pr * z4i + pi_ * z4r;
                    pr = t;
                }
                double mod2 = (double)a * a + (double)b * b;
                acc += pr / pow(mod2, 4.0 * m);
            }
        c[m - 1] = 4.0 * acc;
    }
}

/* smooth (harmonic) part of the periodic kernel at displacement r = x - y,
   r in the min-image cell: K_per(r) = K(r) + K_s(r),
   K(r) = (-r_y, r_x) / (2 pi |r|^2). */
static void kernel_smooth_series(double rx, double ry, double L, int nterms,
                                 const double *c, double *kx, double *ky)
{
    double sr = rx / L, si = ry / L;
    double s2r = sr * sr - si * si, s2i = 2 * sr * si;
    double wr = s2r * s2r - s2i * s2i, wi = 2 * s2r * s2i;       /* s^4 */
    double hr = c[nterms - 1], hi = 0.0;
    for (int m = nterms - 2; m >= 0; --m) {                     /* Horner in w */
        double t = hr * wr - hi * wi + c[m];
        hi = hr * wi + hi * wr;
        hr = t;
    }
    double s3r = s2r * sr - s2i * si, s3i = s2r * si + s2i * sr;
    double Ar = hr * s3r - hi * s3i + PI_D * sr;                 /* + pi conj(s) */
    double Ai = hr * s3i + hi * s3r - PI_D * si;
    double inv = 1.0 / (2.0 * PI_D * L);
    *kx = -Ai * inv;
    *ky = -Ar * inv;
}

void orc_kernel_smooth_d(double rx, double ry, double L, int nterms, const double *c,
                         double *kx, double *ky)
{
    kernel_smooth_series(rx, ry, L, nterms, c, kx, ky);
}

/* ------------------------------------------------------------------------- */
/* Estimator.  Reference mode restates src/cpu/monte_carlo_bs.cpp:11-128 with
   aux = 0 and src/cpu/monte_carlo_sampling.cpp:23-45,
   include/monte_carlo_bs.hpp:60-102.  Consistent mode is the build's
   unbiased estimator of int_Q K_per(x - y) omega_s(y) dy (no reference
   counterpart; the same samplers, the pdfs matched to them). */

typedef struct { const orc_mc_params *P; int owner, sign_idx; const float *U;
                 uint32_t blk[4]; int blk_branch, blk_k; } ustream;

static void philox_draw(const orc_mc_params *P, int owner, int sign_idx, int branch, int k,
                        int block, float out[4])
{
    uint32_t ctr[4] = { (uint32_t)k, (uint32_t)owner, P->step,
                        (P->sim << 8) | ((uint32_t)sign_idx << 7) | ((uint32_t)branch << 4) |
                            (uint32_t)block };
    uint32_t key[2] = { (uint32_t)(P->seed & 0xffffffffu), (uint32_t)(P->seed >> 32) };
    uint32_t x[4];
    orc_philox4x32_10(ctr, key, x);
    for (int i = 0; i < 4; ++i) out[i] = orc_u01(x[i]);
}

static int branch1_draws(const orc_mc_params *P)
{
    return (P->mode == ORC_MODE_REFERENCE && !P->periodic) ? 1 : 3;
}

static void get_uniforms(const orc_mc_params *P, int owner, int sign_idx, int branch, int k,
                         float u[4])
{
    if (P->U) {
        int b1 = branch1_draws(P), b2 = P->aux_draws ? 10 : 2;
        size_t per = (size_t)P->N1 * b1 + (size_t)P->N2 * b2;
        const float *base = P->U + ((size_t)owner * 2 + sign_idx) * per;
        if (branch == 1) {
            for (int i = 0; i < b1; ++i) u[i] = base[(size_t)k * b1 + i];
        } else {
            const float *q = base + (size_t)P->N1 * b1 + (size_t)k * b2;
            u[0] = q[0]; u[1] = q[1];
        }
    } else {
        philox_draw(P, owner, sign_idx, branch, k, 0, u);
    }
}

static float disk_radius_ref(float Lx, float Ly)
{
    /* include/monte_carlo_bs.hpp:73-89: the largest image offset, sqrt(Lx^2+Ly^2)
       for either branch of `periodic` */
    float maxR2 = 0.f;
    for (int dx = -1; dx <= 1; ++dx)
        for (int dy = -1; dy <= 1; ++dy) {
            float ix = dx * Lx, iy = dy * Ly;
            float r2 = ix * ix + iy * iy;
            if (r2 > maxR2) maxR2 = r2;
        }
    return sqrtf(maxR2);
}

void orc_mc_velocity(const orc_mc_params *P, const orc_sign_set *S, int sign_idx, int owner,
                     float dstx, float dsty, const float *px, const float *py, float *out)
{
    const int N1 = P->N1, N2 = P->N2;
    const int sign = sign_idx == 0 ? 1 : -1;
    float w1 = (float)N1 / (float)(N1 + N2);
    float w2 = (float)N2 / (float)(N1 + N2);
    float sx = 0.f, sy = 0.f;
    out[0] = out[1] = 0.f;
    if (S->n <= 0) return;   /* empty set: the reference indexes sub[-1]; we return 0 */

    if (P->mode == ORC_MODE_REFERENCE) {
        float cum = (float)S->omega;
        if (P->periodic) cum *= 9.f;
        float R = disk_radius_ref(P->Lx, P->Ly);
        float alpha = 1.f / (2.f * PI_F * R);
        for (int k = 0; k < N1 + N2; ++k) {
            float u[4];
            float spx, spy;
            if (k < N1) {
                get_uniforms(P, owner, sign_idx, 1, k, u);
                int idx = orc_sample_idx(u[0], S->cdf, S->n);
                int pI = S->sub[idx];
                spx = px[pI]; spy = py[pI];
                if (P->periodic) {
                    float dxs = 0.f, dys = 0.f;
                    if (u[1] < 1.f / 3.f) dxs = -1.f; else if (u[1] > 2.f / 3.f) dxs = 1.f;
                    if (u[2] < 1.f / 3.f) dys = -1.f; else if (u[2] > 2.f / 3.f) dys = 1.f;
                    spx += dxs * P->Lx; spy += dys * P->Ly;
                }
            } else {
                get_uniforms(P, owner, sign_idx, 2, k - N1, u);
                float phi = 2.f * PI_F * u[0];
                float r = R * sqrtf(u[1]);
                spx = dstx + r * cosf(phi);
                spy = dsty + r * sinf(phi);
            }
            float om = orc_eval_vorticity(&S->view, spx, spy, sign, 0, 0, P->h_w, P->periodic);
            if (om <= 0.f) continue;
            /* mis_contrib, src/cpu/monte_carlo_bs.cpp:11-26 */
            float rx = dstx - spx, ry = dsty - spy;
            float rn = sqrtf(rx * rx + ry * ry);
            if (rn < 1e-6f) rn = 1e-6f;
            float inv_r2 = 1.f / (rx * rx + ry * ry + 1e-8f);
            float kx = -ry * inv_r2 / (2 * PI_F);
            float ky = rx * inv_r2 / (2 * PI_F);
            float denom = w1 * om / cum + w2 * alpha / rn;
            float g = om / denom;
            sx += kx * g; sy += ky * g;
        }
    } else {
        const int per = P->periodic;
        float Om = (float)S->omega;
        float Rc = per ? 0.5f * sqrtf(P->Lx * P->Lx + P->Ly * P->Ly)
                       : sqrtf(P->Lx * P->Lx + P->Ly * P->Ly);
        float alpha = 1.f / (2.f * PI_F * Rc);
        for (int k = 0; k < N1 + N2; ++k) {
            float u[4];
            float spx, spy, rx, ry;
            if (k < N1) {
                get_uniforms(P, owner, sign_idx, 1, k, u);
                int idx = orc_sample_idx(u[0], S->cdf, S->n);
                int pI = S->sub[idx];
                float rho = (float)orc_mass_inverse((double)u[1]) * P->h_w;
                float th = 2.f * PI_F * u[2];
                float jx = rho * cosf(th), jy = rho * sinf(th);
                if (S->view.F) {   /* F_j u, src/cuda/monte_carlo_bs.cu:210-212 */
                    const float *F = S->view.F + 4 * (size_t)pI;
                    float ex = F[0] * jx + F[1] * jy, ey = F[2] * jx + F[3] * jy;
                    jx = ex;
                    jy = ey;
                }
                spx = px[pI] + jx;
                spy = py[pI] + jy;
                rx = dstx - spx; ry = dsty - spy;
                if (per) {
                    if (rx > 0.5f * P->Lx) rx -= P->Lx; else if (rx < -0.5f * P->Lx) rx += P->Lx;
                    if (ry > 0.5f * P->Ly) ry -= P->Ly; else if (ry < -0.5f * P->Ly) ry += P->Ly;
                }
            } else {
                get_uniforms(P, owner, sign_idx, 2, k - N1, u);
                float phi = 2.f * PI_F * u[0];
                float rho = Rc * u[1];
                float cx = rho * cosf(phi), cy = rho * sinf(phi);
                spx = dstx + cx; spy = dsty + cy;
                rx = -cx; ry = -cy;
                if (per && (fabsf(rx) > 0.5f * P->Lx || fabsf(ry) > 0.5f * P->Ly)) continue;
            }
            float om = orc_eval_vorticity(&S->view, spx, spy, sign, 0, 0, P->h_w, per);
            if (om <= 0.f) continue;
            float rn = sqrtf(rx * rx + ry * ry);
            if (rn == 0.f) continue;
            float g = om * rn / (w1 * om * rn / Om + w2 * alpha);   /* omega / (w1 p1 + w2 p2) */
            float f = g / (2.f * PI_F * rn * rn);
            float kx = -ry * f, ky = rx * f;
            if (per && P->kper_c && P->nterms > 0) {
                double ksx, ksy;
                kernel_smooth_series(rx, ry, P->Lx, P->nterms, P->kper_c, &ksx, &ksy);
                kx += (float)ksx * g; ky += (float)ksy * g;
            }
            sx += kx; sy += ky;
        }
    }
    float invTot = 1.f / (float)(N1 + N2);
    out[0] = sx * invTot;
    out[1] = sy * invTot;
}

void orc_step(const orc_mc_params *P, const orc_sign_set *pos_set, const orc_sign_set *neg_set,
              int n_owner, float *px, float *py, float dt, float *vel, int nthreads)
{
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 4)
#endif
    for (int i = 0; i < n_owner; ++i) {
        float vp[2], vn[2];
        orc_mc_velocity(P, pos_set, 0, i, px[i], py[i], px, py, vp);
        orc_mc_velocity(P, neg_set, 1, i, px[i], py[i], px, py, vn);
        vel[2 * i] = vp[0] - vn[0];
        vel[2 * i + 1] = vp[1] - vn[1];
    }
    /* src/cpu/simulation_step.cpp:52-71 */
    for (int i = 0; i < n_owner; ++i) {
        px[i] += vel[2 * i] * dt;
        py[i] += vel[2 * i + 1] * dt;
        if (P->periodic) {
            px[i] = px[i] - floorf(px[i] / P->Lx) * P->Lx;
            py[i] = py[i] - floorf(py[i] / P->Ly) * P->Ly;
        }
    }
}

/* ------------------------------------------------------------------------- */
/* src/cuda/naive_biot_savart.cu:15-40 in FP64; wrap_min_image via rint      */
void orc_direct_eps(int n_probe, const double *qx, const double *qy, int skip_self, int n,
                    const double *px, const double *py, const double *g, double L,
                    int periodic, double eps, double *u, int nthreads)
{
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(static)
#endif
    for (int i = 0; i < n_probe; ++i) {
        double ux = 0.0, uy = 0.0, e2 = eps * eps;
        for (int j = 0; j < n; ++j) {
            if (skip_self && j == i) continue;
            double dx = qx[i] - px[j], dy = qy[i] - py[j];
            if (periodic) { dx -= rint(dx / L) * L; dy -= rint(dy / L) * L; }
            double coef = 0.5 * (1.0 / PI_D) / (dx * dx + dy * dy + e2);
            ux += g[j] * (-dy * coef);
            uy += g[j] * (dx * coef);
        }
        u[2 * i] = ux; u[2 * i + 1] = uy;
    }
}

/* exact velocity of sum_j g_j W(|y - x_j|) under K_per: by the 2-D shell
   theorem the free part of a radially symmetric blob is K(r) M(|r|/h); the
   smooth part is harmonic on the blob, so its blob average is its value at
   the centre (mean-value property). */
void orc_direct_blob(int n_probe, const double *qx, const double *qy, int n, const double *px,
                     const double *py, const double *g, double h_w, double L, int nterms,
                     const double *kper_c, double *u, int nthreads)
{
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(static)
#endif
    for (int i = 0; i < n_probe; ++i) {
        double ux = 0.0, uy = 0.0;
        for (int j = 0; j < n; ++j) {
            double rx = qx[i] - px[j], ry = qy[i] - py[j];
            rx -= rint(rx / L) * L; ry -= rint(ry / L) * L;
            double r2 = rx * rx + ry * ry;
            if (r2 > 0.0) {
                double m = orc_mass_fraction(sqrt(r2) / h_w);
                double f = m / (2.0 * PI_D * r2);
                ux += g[j] * (-ry * f);
                uy += g[j] * (rx * f);
            }
            if (nterms > 0) {
                double kx, ky;
                kernel_smooth_series(rx, ry, L, nterms, kper_c, &kx, &ky);
                ux += g[j] * kx; uy += g[j] * ky;
            }
        }
        u[2 * i] = ux; u[2 * i + 1] = uy;
    }
}

/* ------------------------------------------------------------------------- */
/* src/cuda/cg_rbf_init.cu:11-31 (matvec, sign = 0, use_signed, diagonal kept) */
void orc_rbf_matvec(int n, const double *px, const double *py, double h_w, double L,
                    int periodic, const double *x, double *y, int nthreads)
{
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(static)
#endif
    for (int i = 0; i < n; ++i) {
        double s = 0.0;
        for (int j = 0; j < n; ++j) {
            double dx = px[j] - px[i], dy = py[j] - py[i];
            if (periodic) { dx -= rint(dx / L) * L; dy -= rint(dy / L) * L; }
            double r2 = dx * dx + dy * dy;
            if (r2 < h_w * h_w) s += x[j] * orc_spline_d(sqrt(r2), h_w);
        }
        y[i] = s;
    }
}

static double dotd(int n, const double *a, const double *b)
{
    double s = 0.0;
    for (int i = 0; i < n; ++i) s += a[i] * b[i];
    return s;
}

/* src/cuda/cg_rbf_init.cu:53-116: w0 = b / W(0), plain CG, stop at
   ||r|| / ||b|| < tol.  Returns iterations; -(k+1) on p.Ap <= 0 breakdown. */
int orc_cg_rbf(int n, const double *px, const double *py, double h_w, double L, int periodic,
               const double *b, double *w, double tol, int max_iter, double *rel_res,
               int nthreads)
{
    double *r = (double *)malloc(sizeof(double) * n), *p = (double *)malloc(sizeof(double) * n);
    double *Ap = (double *)malloc(sizeof(double) * n);
    double diag = orc_spline_d(0.0, h_w);
    for (int i = 0; i < n; ++i) w[i] = b[i] / diag;
    orc_rbf_matvec(n, px, py, h_w, L, periodic, w, Ap, nthreads);
    for (int i = 0; i < n; ++i) { r[i] = b[i] - Ap[i]; p[i] = r[i]; }
    double bn = sqrt(dotd(n, b, b));
    double r2_old = dotd(n, r, r);
    int ret = max_iter;
    *rel_res = bn > 0 ? sqrt(r2_old) / bn : 0.0;
    if (bn == 0.0 || *rel_res < tol) { ret = 0; goto done; }
    for (int k = 0; k < max_iter; ++k) {
        orc_rbf_matvec(n, px, py, h_w, L, periodic, p, Ap, nthreads);
        double pAp = dotd(n, p, Ap);
        if (!(pAp > 0.0)) { ret = -(k + 1); break; }
        double alpha = r2_old / pAp;
        for (int i = 0; i < n; ++i) { w[i] += alpha * p[i]; r[i] -= alpha * Ap[i]; }
        double r2_new = dotd(n, r, r);
        *rel_res = sqrt(r2_new) / bn;
        if (*rel_res < tol) { ret = k + 1; break; }
        double beta = r2_new / r2_old;
        for (int i = 0; i < n; ++i) p[i] = r[i] + beta * p[i];
        r2_old = r2_new;
    }
done:
    free(r); free(p); free(Ap);
    return ret;
}

/* ------------------------------------------------------------------------- */
/* include/analytical_flow.hpp:12-31                                           */
double orc_tg_vorticity(double x, double y)
{
    return 4.0 * PI_D * sin(2.0 * PI_D * x) * sin(2.0 * PI_D * y);
}

void orc_tg_velocity(double x, double y, double *u, double *v)
{
    *u = sin(2.0 * PI_D * x) * cos(2.0 * PI_D * y);
    *v = -cos(2.0 * PI_D * x) * sin(2.0 * PI_D * y);
}

/* double shear layer: u = tanh((y-1/4)/rho) (y <= 1/2), tanh((3/4-y)/rho)
   otherwise; v = delta sin(2 pi x); omega = dv/dx - du/dy */
double orc_shear_vorticity(double x, double y, double rho, double delta)
{
    double base = 2.0 * PI_D * delta * cos(2.0 * PI_D * x);
    if (y <= 0.5) {
        double c = cosh((y - 0.25) / rho);
        return base - 1.0 / (rho * c * c);
    } else {
        double c = cosh((0.75 - y) / rho);
        return base + 1.0 / (rho * c * c);
    }
}
