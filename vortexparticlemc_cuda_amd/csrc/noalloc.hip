// noalloc.hip -- vpmc_step_noalloc: vpmc_step without a caller workspace and
// without any allocation, for callers that cannot allocate (code under HIP
// graph capture, serving loops that forbid hipMalloc).
//
// The step's scratch (the field builds' keys and counts, then the binned
// estimator's sample buffers) lives in one file-scope device array of
// NOALLOC_BYTES.  Host code reaches it through the address that
// hipGetSymbolAddress returns for the current device on the first call there
// (one address per device, kept for the life of the process), never through
// the array's name.  The estimator's owner chunks are sized so that one chunk
// fits the array; the chunk size does not change any result, so the step is
// bit-identical to vpmc_step with a workspace of its own.  The step clears
// what must start at zero (the field histograms) itself; every other scratch
// byte is written before it is read.  Calls that overlap on one device race
// on the array.
#include <mutex>

#include "binned_core.hpp"

namespace vpmc {

constexpr size_t NOALLOC_BYTES = (size_t)VPMC_NOALLOC_SCRATCH_BYTES;
constexpr int NOALLOC_MAX_DEVICES = 64;

static __device__ __align__(256) unsigned char scratch[NOALLOC_BYTES];

static void *g_scratch[NOALLOC_MAX_DEVICES];
static std::mutex g_scratch_mu;

// the current device's copy of the array (looked up once per device)
static int scratch_for_current_device(void **out)
{
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) {
        set_hip_error(e);
        return VPMC_ERR_HIP;
    }
    if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) return VPMC_ERR_ARG;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch[dev]) {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) {
            set_hip_error(e);
            return VPMC_ERR_HIP;
        }
        if (!p) return VPMC_ERR_HIP;
        g_scratch[dev] = p;
    }
    *out = g_scratch[dev];
    return VPMC_OK;
}

// the largest owners-per-chunk whose binned workspace fits `cap` bytes (at
// most the descriptor's own chunk_samples), as chunk_samples; 0 if not even
// one owner's samples fit
static uint64_t fitted_chunk_samples(vpmc_mc_desc d, size_t cap)
{
    const uint64_t per_owner = 2 * (uint64_t)samples_per_pair(d.N1, d.N2, d.aux);
    uint64_t limit = d.chunk_samples ? d.chunk_samples : ((uint64_t)1 << 30);
    if (limit > ((uint64_t)1 << 31) - 1) limit = ((uint64_t)1 << 31) - 1;
    uint64_t hi = limit / per_owner;
    if (hi > (uint64_t)d.n_owner) hi = (uint64_t)d.n_owner;
    if (hi == 0) return 0;
    auto fits = [&](uint64_t owners) {
        d.chunk_samples = owners * per_owner;
        const size_t b = vpmc_mc_binned_workspace_bytes(&d);
        return b > 0 && b <= cap;
    };
    if (fits(hi)) return hi * per_owner;
    if (!fits(1)) return 0;
    uint64_t lo = 1;   // fits(lo), !fits(hi)
    while (hi - lo > 1) {
        const uint64_t mid = lo + (hi - lo) / 2;
        if (fits(mid)) lo = mid; else hi = mid;
    }
    return lo * per_owner;
}

}  // namespace vpmc

using namespace vpmc;

extern "C" int vpmc_step_noalloc(vpmc_mc_desc *d, int n_total, float *pos_xy, const float *omega,
                                 const int *sub_pos, const int *sub_neg, float *vel_xy, float dt,
                                 void *stream)
{
    if (!d || n_total < 0) return VPMC_ERR_ARG;
    // the fields' sort scratch (both sets use it in turn)
    const int n_set = d->set[0].n > d->set[1].n ? d->set[0].n : d->set[1].n;
    if (vpmc_field_workspace_bytes(n_set, d->set[0].field.nx, d->set[0].field.ny) >
        NOALLOC_BYTES)
        return VPMC_ERR_WORKSPACE;
    vpmc_mc_desc dd = *d;
    if (dd.algo == VPMC_ALGO_BINNED && dd.n_owner > 0) {
        const uint64_t cs = fitted_chunk_samples(dd, NOALLOC_BYTES);
        if (cs == 0) return VPMC_ERR_WORKSPACE;   // one owner's samples exceed the array
        dd.chunk_samples = cs;
    }
    void *ws = nullptr;
    VPMC_TRY(scratch_for_current_device(&ws));
    return vpmc_step(&dd, n_total, pos_xy, omega, sub_pos, sub_neg, vel_xy, dt, ws, NOALLOC_BYTES,
                     stream);
}
