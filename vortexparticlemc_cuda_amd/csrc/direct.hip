// direct.hip -- FP64 direct Biot-Savart sums at probe points (error oracles).
//
// One thread per probe; source particles are staged through LDS in tiles of
// 256 so every wave reads them with uniform addresses.
#include <math.h>

#include <mutex>

#include "vpmc_internal.hpp"

namespace vpmc {

constexpr int DS_THREADS = 256;
constexpr int KPER_TERMS_F64 = 24;

struct KperC {
    double c[KPER_TERMS_F64];
};

// src/cuda/naive_biot_savart.cu:15-40 in FP64
__global__ void __launch_bounds__(DS_THREADS)
    k_direct_eps(int nprobe, const double2 *__restrict__ q, int skip_self, int n,
                 const double2 *__restrict__ pos, const double *__restrict__ g, double L,
                 int periodic, double eps2, double2 *__restrict__ u)
{
    __shared__ double3 tile[DS_THREADS];
    int i = blockIdx.x * DS_THREADS + threadIdx.x;
    double2 p = i < nprobe ? q[i] : make_double2(0.0, 0.0);
    double ux = 0.0, uy = 0.0;
    for (int t0 = 0; t0 < n; t0 += DS_THREADS) {
        int j = t0 + threadIdx.x;
        if (j < n) tile[threadIdx.x] = make_double3(pos[j].x, pos[j].y, g[j]);
        __syncthreads();
        int m = min(DS_THREADS, n - t0);
        for (int k = 0; k < m; ++k) {
            if (skip_self && t0 + k == i) continue;
            double3 s = tile[k];
            double dx = p.x - s.x, dy = p.y - s.y;
            if (periodic) {
                dx -= rint(dx / L) * L;
                dy -= rint(dy / L) * L;
            }
            double coef = 0.5 * (1.0 / PI_D) / (dx * dx + dy * dy + eps2);
            ux += s.z * (-dy * coef);
            uy += s.z * (dx * coef);
        }
        __syncthreads();
    }
    if (i < nprobe) u[i] = make_double2(ux, uy);
}

__device__ __forceinline__ double mass_fraction(double q)
{
    if (q >= 1.0) return 1.0;
    double q2 = q * q;
    if (q <= 0.5) return (80.0 / 7.0) * (0.5 * q2 - 1.5 * q2 * q2 + 1.2 * q2 * q2 * q);
    return (80.0 / 7.0) * (q2 - 2.0 * q2 * q + 1.5 * q2 * q2 - 0.4 * q2 * q2 * q - 0.0125);
}

// exact periodic velocity of a cubic-spline blob: free part K(r) M(|r|/h)
// (2-D shell theorem) plus the harmonic lattice correction at the centre
// (mean-value property).
__global__ void __launch_bounds__(DS_THREADS)
    k_direct_blob(int nprobe, const double2 *__restrict__ q, int n,
                  const double2 *__restrict__ pos, const double *__restrict__ g, double h_w,
                  double L, KperC kc, double2 *__restrict__ u)
{
    __shared__ double3 tile[DS_THREADS];
    int i = blockIdx.x * DS_THREADS + threadIdx.x;
    double2 p = i < nprobe ? q[i] : make_double2(0.0, 0.0);
    double ux = 0.0, uy = 0.0;
    const double inv_h = 1.0 / h_w, inv_2piL = 1.0 / (2.0 * PI_D * L);
    for (int t0 = 0; t0 < n; t0 += DS_THREADS) {
        int j = t0 + threadIdx.x;
        if (j < n) tile[threadIdx.x] = make_double3(pos[j].x, pos[j].y, g[j]);
        __syncthreads();
        int m = min(DS_THREADS, n - t0);
        for (int k = 0; k < m; ++k) {
            double3 s = tile[k];
            double rx = p.x - s.x, ry = p.y - s.y;
            rx -= rint(rx / L) * L;
            ry -= rint(ry / L) * L;
            double r2 = rx * rx + ry * ry;
            if (r2 > 0.0) {
                double f = mass_fraction(sqrt(r2) * inv_h) / (2.0 * PI_D * r2);
                ux += s.z * (-ry * f);
                uy += s.z * (rx * f);
            }
            double sr = rx / L, si = ry / L;
            double s2r = sr * sr - si * si, s2i = 2.0 * sr * si;
            double wr = s2r * s2r - s2i * s2i, wi = 2.0 * s2r * s2i;
            double hr = kc.c[KPER_TERMS_F64 - 1], hi = 0.0;
            for (int mm = KPER_TERMS_F64 - 2; mm >= 0; --mm) {
                double t = hr * wr - hi * wi + kc.c[mm];
                hi = hr * wi + hi * wr;
                hr = t;
            }
            double s3r = s2r * sr - s2i * si, s3i = s2r * si + s2i * sr;
            double Ar = hr * s3r - hi * s3i + PI_D * sr;
            double Ai = hr * s3i + hi * s3r - PI_D * si;
            ux += s.z * (-Ai * inv_2piL);
            uy += s.z * (-Ar * inv_2piL);
        }
        __syncthreads();
    }
    if (i < nprobe) u[i] = make_double2(ux, uy);
}

static const KperC &kper_f64()
{
    static KperC k;
    static std::once_flag once;
    std::call_once(once, [] {
        double q = exp(-2.0 * PI_D), qn = 1.0, s = 0.0;
        for (int nn = 1; nn <= 12; ++nn) {
            qn *= q;
            double sig = 0.0;
            for (int dd = 1; dd <= nn; ++dd)
                if (nn % dd == 0) sig += (double)dd * dd * dd;
            s += sig * qn;
        }
        double pi4 = PI_D * PI_D * PI_D * PI_D;
        k.c[0] = (pi4 / 45.0) * (1.0 + 240.0 * s);
        for (int m = 2; m <= KPER_TERMS_F64; ++m) {
            int R = (m == 2) ? 600 : 160;
            double acc = 0.0;
            for (int a = 1; a <= R; ++a)
                for (int b = 0; b <= R; ++b) {
                    double re = a, im = -(double)b;
                    double z2r = re * re - im * im, z2i = 2 * re * im;
                    double z4r = z2r * z2r - z2i * z2i, z4i = 2 * z2r * z2i;
                    double pr = 1.0, pim = 0.0;
                    for (int kk = 0; kk < m; ++kk) {
                        double t = pr * z4r - pim * z4i;
                        pim = pr * z4i + pim * z4r;
                        pr = t;
                    }
                    double mod2 = (double)a * a + (double)b * b;
                    acc += pr / pow(mod2, 4.0 * m);
                }
            k.c[m - 1] = 4.0 * acc;
        }
    });
    return k;
}

}  // namespace vpmc

using namespace vpmc;

extern "C" {

int vpmc_direct_sum_eps(int nprobe, const double *probe_xy, int skip_self, int n,
                        const double *pos_xy, const double *g, double L, int periodic,
                        double eps, double *u_out, void *stream)
{
    if (nprobe < 0 || n < 0 || !(L > 0.0)) return VPMC_ERR_ARG;
    if (nprobe == 0) return VPMC_OK;
    if (!probe_xy || !u_out || (n > 0 && (!pos_xy || !g))) return VPMC_ERR_ARG;
    hipLaunchKernelGGL(k_direct_eps, dim3((nprobe + DS_THREADS - 1) / DS_THREADS),
                       dim3(DS_THREADS), 0, as_stream(stream), nprobe, (const double2 *)probe_xy,
                       skip_self, n, (const double2 *)pos_xy, g, L, periodic, eps * eps,
                       (double2 *)u_out);
    VPMC_CHECK_LAUNCH();
    return VPMC_OK;
}

int vpmc_direct_sum_blob(int nprobe, const double *probe_xy, int n, const double *pos_xy,
                         const double *g, double h_w, double L, double *u_out, void *stream)
{
    if (nprobe < 0 || n < 0 || !(L > 0.0) || !(h_w > 0.0)) return VPMC_ERR_ARG;
    if (nprobe == 0) return VPMC_OK;
    if (!probe_xy || !u_out || (n > 0 && (!pos_xy || !g))) return VPMC_ERR_ARG;
    hipLaunchKernelGGL(k_direct_blob, dim3((nprobe + DS_THREADS - 1) / DS_THREADS),
                       dim3(DS_THREADS), 0, as_stream(stream), nprobe, (const double2 *)probe_xy,
                       n, (const double2 *)pos_xy, g, h_w, L, kper_f64(), (double2 *)u_out);
    VPMC_CHECK_LAUNCH();
    return VPMC_OK;
}

}  // extern "C"
