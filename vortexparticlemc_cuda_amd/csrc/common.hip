// common.hip -- status reporting and version of libvpmc.
#include "vpmc_internal.hpp"

namespace vpmc {
static thread_local int g_last_hip_error = 0;
void set_hip_error(hipError_t e) { g_last_hip_error = (int)e; }

// FP32 FMA issue-rate probe: 16 independent dependency chains per lane (more
// than the VALU latency), iters rounds of 16 FMAs; the result is stored only
// when it equals an impossible value, which keeps every chain live.  Both
// multiplicands live in VGPRs (a lane-dependent term the compiler cannot
// fold): with a scalar operand the FMA issues at half the rate on gfx950,
// while the pair loops of the evaluation kernels take VGPR operands.
constexpr int PROBE_CHAINS = 16;
__global__ void __launch_bounds__(256) k_fma_probe(int iters, float a, float b,
                                                   float *__restrict__ out)
{
    const float tiny = (float)threadIdx.x * 1e-30f;
    const float av = a + tiny * tiny, bv = b + tiny;   // per-lane registers
    a = av;
    b = bv;
    float x[PROBE_CHAINS];
#pragma unroll
    for (int k = 0; k < PROBE_CHAINS; ++k) x[k] = (float)threadIdx.x * 1e-3f + (float)k;
#pragma unroll 4
    for (int i = 0; i < iters; ++i) {
#pragma unroll
        for (int k = 0; k < PROBE_CHAINS; ++k) x[k] = fmaf(x[k], a, b);
    }
    float s = 0.f;
#pragma unroll
    for (int k = 0; k < PROBE_CHAINS; ++k) s += x[k];
    if (s == -1.25f) out[blockIdx.x] = s;
}
}  // namespace vpmc

extern "C" {

const char *vpmc_status_string(int status)
{
    switch (status) {
    case VPMC_OK: return "ok";
    case VPMC_ERR_ARG: return "invalid argument";
    case VPMC_ERR_HIP: return "HIP runtime error";
    case VPMC_ERR_WORKSPACE: return "workspace too small";
    case VPMC_ERR_GRID: return "cell grid cannot cover the kernel support without overlap";
    case VPMC_ERR_CG_BREAKDOWN: return "CG breakdown (p.Ap <= 0)";
    case VPMC_ERR_NOT_CONVERGED: return "CG did not converge";
    default: return "unknown status";
    }
}

int vpmc_last_hip_error(void) { return vpmc::g_last_hip_error; }

int vpmc_version(void) { return 100; }

int vpmc_probe_fma(int blocks, int iters, float *out, void *stream)
{
    if (blocks <= 0 || iters <= 0 || !out) return VPMC_ERR_ARG;
    hipLaunchKernelGGL(vpmc::k_fma_probe, dim3((unsigned)blocks), dim3(256), 0,
                       vpmc::as_stream(stream), iters, 0.999f, 1e-3f, out);
    VPMC_CHECK_LAUNCH();
    return VPMC_OK;
}

}  // extern "C"
