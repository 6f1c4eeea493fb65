// scan.hip -- device-wide prefix sums (int32, FP64) and fill kernels.
//
// Three passes with a fixed partition (SCAN_CHUNK elements per block,
// SCAN_ITEMS contiguous per thread): block reduce -> scan of block totals in
// one workgroup -> block scan with carry-in.  The association order depends
// only on n, so FP64 results are reproducible run to run.  Wave-level scans
// are written directly on wave64 shuffles.
#include "vpmc_internal.hpp"

namespace vpmc {

template <class T> __device__ __forceinline__ T wave_incl_scan(T v, int lane)
{
#pragma unroll
    for (int o = 1; o < WAVE; o <<= 1) {
        T u = __shfl_up(v, o, WAVE);
        if (lane >= o) v += u;
    }
    return v;
}

// exclusive scan of one value per thread across the block; returns the
// exclusive prefix, *block_total receives the sum
template <class T> __device__ __forceinline__ T block_excl_scan(T v, T *block_total)
{
    __shared__ T wtot[SCAN_THREADS / WAVE];
    const int lane = threadIdx.x & (WAVE - 1), wid = threadIdx.x / WAVE;
    T incl = wave_incl_scan(v, lane);
    if (lane == WAVE - 1) wtot[wid] = incl;
    __syncthreads();
    T off = T(0), tot = T(0);
#pragma unroll
    for (int w = 0; w < SCAN_THREADS / WAVE; ++w) {
        if (w < wid) off += wtot[w];
        tot += wtot[w];
    }
    __syncthreads();
    *block_total = tot;
    return off + incl - v;
}

template <class T>
__global__ void __launch_bounds__(SCAN_THREADS) k_scan_reduce(const T *__restrict__ in, int n,
                                                             T *__restrict__ part)
{
    const size_t base = (size_t)blockIdx.x * SCAN_CHUNK + (size_t)threadIdx.x * SCAN_ITEMS;
    T s = T(0);
#pragma unroll
    for (int i = 0; i < SCAN_ITEMS; ++i)
        if (base + i < (size_t)n) s += in[base + i];
    T tot;
    block_excl_scan(s, &tot);
    if (threadIdx.x == 0) part[blockIdx.x] = tot;
}

template <class T>
__global__ void __launch_bounds__(SCAN_THREADS) k_scan_parts(T *__restrict__ part, int nb)
{
    T carry = T(0);
    for (int b0 = 0; b0 < nb; b0 += SCAN_THREADS) {
        int i = b0 + threadIdx.x;
        T v = i < nb ? part[i] : T(0);
        T tot;
        T ex = block_excl_scan(v, &tot);
        if (i < nb) part[i] = carry + ex;
        carry += tot;
    }
}

template <class T>
__global__ void __launch_bounds__(SCAN_THREADS) k_scan_down(const T *__restrict__ in,
                                                           T *__restrict__ out, int n,
                                                           const T *__restrict__ part,
                                                           int inclusive, T *__restrict__ total)
{
    const size_t base = (size_t)blockIdx.x * SCAN_CHUNK + (size_t)threadIdx.x * SCAN_ITEMS;
    T v[SCAN_ITEMS];
    T s = T(0);
#pragma unroll
    for (int i = 0; i < SCAN_ITEMS; ++i) {
        v[i] = (base + i < (size_t)n) ? in[base + i] : T(0);
        s += v[i];
    }
    T tot;
    T ex = block_excl_scan(s, &tot);
    T run = (part ? part[blockIdx.x] : T(0)) + ex;
#pragma unroll
    for (int i = 0; i < SCAN_ITEMS; ++i) {
        T nv = run + v[i];
        if (base + i < (size_t)n) out[base + i] = inclusive ? nv : run;
        run = nv;
    }
    if (total && blockIdx.x == gridDim.x - 1 && threadIdx.x == SCAN_THREADS - 1) *total = run;
}

template <class T> __global__ void k_fill(T *__restrict__ p, T v, size_t n)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        p[i] = v;
}

size_t scan_workspace_bytes(int n, size_t elem)
{
    int nb = scan_blocks(n > 0 ? n : 1);
    return align256((size_t)nb * elem);
}

template <class T>
static int scan_impl(const T *in, T *out, int n, bool inclusive, T *total, void *ws,
                     size_t ws_bytes, hipStream_t s)
{
    if (n <= 0) {
        if (total) {   // a fill kernel, not a memset node, so graphs replay it
            hipLaunchKernelGGL(k_fill<T>, dim3(1), dim3(64), 0, s, total, T(0), (size_t)1);
            VPMC_CHECK_LAUNCH();
        }
        return VPMC_OK;
    }
    int nb = scan_blocks(n);
    T *part = nullptr;
    if (nb > 1) {
        if (!ws || ws_bytes < (size_t)nb * sizeof(T)) return VPMC_ERR_WORKSPACE;
        part = static_cast<T *>(ws);
        hipLaunchKernelGGL(k_scan_reduce<T>, dim3(nb), dim3(SCAN_THREADS), 0, s, in, n, part);
        hipLaunchKernelGGL(k_scan_parts<T>, dim3(1), dim3(SCAN_THREADS), 0, s, part, nb);
    }
    hipLaunchKernelGGL(k_scan_down<T>, dim3(nb), dim3(SCAN_THREADS), 0, s, in, out, n,
                       (const T *)part, inclusive ? 1 : 0, total);
    VPMC_CHECK_LAUNCH();
    return VPMC_OK;
}

int scan_i32(const int *in, int *out, int n, bool inclusive, int *total, void *ws,
             size_t ws_bytes, hipStream_t s)
{
    return scan_impl<int>(in, out, n, inclusive, total, ws, ws_bytes, s);
}

int scan_f64(const double *in, double *out, int n, bool inclusive, double *total, void *ws,
             size_t ws_bytes, hipStream_t s)
{
    return scan_impl<double>(in, out, n, inclusive, total, ws, ws_bytes, s);
}

int fill_i32(int *p, int v, size_t n, hipStream_t s)
{
    if (n == 0) return VPMC_OK;
    unsigned g = (unsigned)((n + 255) / 256);
    if (g > 4096) g = 4096;
    hipLaunchKernelGGL(k_fill<int>, dim3(g), dim3(256), 0, s, p, v, n);
    VPMC_CHECK_LAUNCH();
    return VPMC_OK;
}

int fill_f32(float *p, float v, size_t n, hipStream_t s)
{
    if (n == 0) return VPMC_OK;
    unsigned g = (unsigned)((n + 255) / 256);
    if (g > 4096) g = 4096;
    hipLaunchKernelGGL(k_fill<float>, dim3(g), dim3(256), 0, s, p, v, n);
    VPMC_CHECK_LAUNCH();
    return VPMC_OK;
}

}  // namespace vpmc
