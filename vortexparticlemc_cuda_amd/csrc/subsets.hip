// subsets.hip -- sign partition, FP64 |w| CDF and its guide table.
//
// The reference builds these with Thrust count_if/copy_if/inclusive_scan in
// FP32 (src/cuda/build_particle_sets.cu:18-86).  Here: flag scans give the
// stable compaction, the |w| prefix is an FP64 scan with a fixed association
// order, and the normalisation reads the subset size from device memory, so
// nothing synchronises with the host.
#include "vpmc_internal.hpp"

namespace vpmc {

__global__ void k_sign_flags(int n, const float *__restrict__ w, int *__restrict__ fp,
                             int *__restrict__ fn)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float v = w[i];
    fp[i] = v > 0.f ? 1 : 0;
    fn[i] = v < 0.f ? 1 : 0;
}

__global__ void k_sign_compact(int n, const float *__restrict__ w, const int *__restrict__ op,
                               const int *__restrict__ on, int *__restrict__ sub_pos,
                               int *__restrict__ sub_neg, double *__restrict__ ap,
                               double *__restrict__ an)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float v = w[i];
    if (v > 0.f) { sub_pos[op[i]] = i; ap[op[i]] = (double)v; }
    else if (v < 0.f) { sub_neg[on[i]] = i; an[on[i]] = -(double)v; }
}

// zero the tails so a full-length scan equals the subset prefix
__global__ void k_zero_tail(int n, const int *__restrict__ counts, double *__restrict__ ap,
                            double *__restrict__ an)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (i >= counts[0]) ap[i] = 0.0;
    if (i >= counts[1]) an[i] = 0.0;
}

__global__ void k_normalise(int n, const int *__restrict__ counts, const double *__restrict__ cp,
                            const double *__restrict__ cn, double *__restrict__ cdf_pos,
                            double *__restrict__ cdf_neg, double *__restrict__ totals)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    int np = counts[0], nn = counts[1];
    double tp = np > 0 ? cp[np - 1] : 0.0, tn = nn > 0 ? cn[nn - 1] : 0.0;
    if (i == 0) { totals[0] = tp; totals[1] = tn; }
    if (i < np) cdf_pos[i] = cp[i] / tp;
    if (i < nn) cdf_neg[i] = cn[i] / tn;
}

__global__ void k_guide(const double *__restrict__ cdf, const int *__restrict__ count_dev,
                        int m_log2, int *__restrict__ guide)
{
    long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    long long M = 1ll << m_log2;
    if (k > M) return;
    int n = *count_dev;
    double target = (double)k / (double)M;   // exact: M is a power of two
    int lo = 0, hi = n;
    while (lo < hi) {
        int mid = lo + ((hi - lo) >> 1);
        if (target < cdf[mid]) hi = mid; else lo = mid + 1;
    }
    guide[k] = lo;
}

__global__ void k_sample_idx(int nu, const float *__restrict__ u, const double *__restrict__ cdf,
                             const int *__restrict__ count_dev, const int *__restrict__ guide,
                             int glog2, int *__restrict__ out)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nu) return;
    int n = *count_dev;
    out[i] = n > 0 ? guided_upper_bound(u[i], cdf, guide, glog2, n) : -1;
}

}  // namespace vpmc

using namespace vpmc;

extern "C" {

size_t vpmc_subsets_workspace_bytes(int n)
{
    size_t m = (size_t)(n > 0 ? n : 1);
    return 2 * align256(m * 4) + 2 * align256(m * 8) + scan_workspace_bytes(n, sizeof(double));
}

int vpmc_build_subsets(int n, const float *omega, int *sub_pos, int *sub_neg, double *cdf_pos,
                       double *cdf_neg, int *counts, double *totals, void *workspace,
                       size_t workspace_bytes, void *stream)
{
    if (n < 0 || !counts || !totals) return VPMC_ERR_ARG;
    hipStream_t s = as_stream(stream);
    if (n == 0) {
        VPMC_TRY(fill_i32(counts, 0, 2, s));
        return scan_f64(nullptr, nullptr, 0, true, totals, nullptr, 0, s) == VPMC_OK
                   ? scan_f64(nullptr, nullptr, 0, true, totals + 1, nullptr, 0, s)
                   : VPMC_ERR_HIP;
    }
    if (!omega || !sub_pos || !sub_neg || !cdf_pos || !cdf_neg) return VPMC_ERR_ARG;
    Carve cv(workspace, workspace_bytes);
    int *fp = cv.take<int>(n), *fn = cv.take<int>(n);
    double *ap = cv.take<double>(n), *an = cv.take<double>(n);
    void *sws = cv.take<char>(scan_workspace_bytes(n, sizeof(double)));
    if (!fp || !fn || !ap || !an || !sws) return VPMC_ERR_WORKSPACE;
    size_t swb = scan_workspace_bytes(n, sizeof(double));
    unsigned gb = (unsigned)((n + 255) / 256);
    hipLaunchKernelGGL(k_sign_flags, dim3(gb), dim3(256), 0, s, n, omega, fp, fn);
    VPMC_CHECK_LAUNCH();
    VPMC_TRY(scan_i32(fp, fp, n, false, counts, sws, swb, s));
    VPMC_TRY(scan_i32(fn, fn, n, false, counts + 1, sws, swb, s));
    hipLaunchKernelGGL(k_sign_compact, dim3(gb), dim3(256), 0, s, n, omega, (const int *)fp,
                       (const int *)fn, sub_pos, sub_neg, ap, an);
    hipLaunchKernelGGL(k_zero_tail, dim3(gb), dim3(256), 0, s, n, (const int *)counts, ap, an);
    VPMC_CHECK_LAUNCH();
    VPMC_TRY(scan_f64(ap, ap, n, true, nullptr, sws, swb, s));
    VPMC_TRY(scan_f64(an, an, n, true, nullptr, sws, swb, s));
    hipLaunchKernelGGL(k_normalise, dim3(gb), dim3(256), 0, s, n, (const int *)counts,
                       (const double *)ap, (const double *)an, cdf_pos, cdf_neg, totals);
    VPMC_CHECK_LAUNCH();
    return VPMC_OK;
}

int vpmc_build_guide(const double *cdf, const int *count_dev, int m_log2, int *guide,
                     void *stream)
{
    if (!cdf || !count_dev || !guide || m_log2 < 0 || m_log2 > 30) return VPMC_ERR_ARG;
    long long M = 1ll << m_log2;
    unsigned gb = (unsigned)((M + 1 + 255) / 256);
    hipLaunchKernelGGL(k_guide, dim3(gb), dim3(256), 0, as_stream(stream), cdf, count_dev,
                       m_log2, guide);
    VPMC_CHECK_LAUNCH();
    return VPMC_OK;
}

int vpmc_sample_idx(int nu, const float *u, const double *cdf, const int *count_dev,
                    const int *guide, int m_log2, int *idx_out, void *stream)
{
    if (nu < 0 || m_log2 < 0 || m_log2 > 30) return VPMC_ERR_ARG;
    if (nu == 0) return VPMC_OK;
    if (!u || !cdf || !count_dev || !guide || !idx_out) return VPMC_ERR_ARG;
    hipLaunchKernelGGL(k_sample_idx, dim3((nu + 255) / 256), dim3(256), 0, as_stream(stream), nu,
                       u, cdf, count_dev, guide, m_log2, idx_out);
    VPMC_CHECK_LAUNCH();
    return VPMC_OK;
}

}  // extern "C"
