// common.hip -- status reporting and version of libvpmc.
#include "vpmc_internal.hpp"

namespace vpmc {
static thread_local int g_last_hip_error = 0;
void set_hip_error(hipError_t e) { g_last_hip_error = (int)e; }
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

}  // extern "C"
