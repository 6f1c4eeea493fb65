// hash.hip -- counting-sort spatial hash on the reference's cell grid.
//
// FillCells: FP32 cell keys (src/cuda/hash.cu:17-22, src/cpu/hash.cpp:20-25)
// and the cell histogram, then its exclusive scan (scan.hip).  The reference
// grid holds few cells (36 at 200 x 200, hwr 16; 4096 at 1024 x 1024, hwr 8)
// with hundreds to thousands of particles each, so one global atomic per
// particle would serialise on a few counters: each block counts HASH_ITEMS x
// 256 particles into an LDS histogram and adds its non-zero bins to the
// global one.  Grids beyond HASH_LDS_CELLS cells (a few particles per cell)
// count into the global histogram directly.
//
// CountingSort: the CPU path's stable order (ascending local index within a
// cell, src/cpu/hash.cpp:36-47) is the stable sort of the cell keys carrying
// their local indices: the LSD radix sort of sort.hip (per-block digit
// histograms, one scan, ballot-ranked scatter), with no atomics and no
// per-cell ranking loop.  cell_num is cleared afterwards, the state the
// reference's atomic countdown leaves it in (src/cuda/hash.cu:27-39).
#include <algorithm>

#include "vpmc_internal.hpp"

namespace vpmc {

constexpr int HASH_THREADS = 256;
constexpr int HASH_ITEMS = 8;
constexpr int HASH_LDS_CELLS = 4096;

template <bool LDS>
__global__ void __launch_bounds__(HASH_THREADS)
    k_hash_keys(int rshx, int rshy, float inv_hx, float inv_hy, int n,
                const int *__restrict__ sub, const float2 *__restrict__ pos,
                int *__restrict__ cell_num, int *__restrict__ vp_cell)
{
    __shared__ int h[LDS ? HASH_LDS_CELLS : 1];
    const int cells = rshx * rshy;
    if (LDS) {
        for (int c = threadIdx.x; c < cells; c += HASH_THREADS) h[c] = 0;
        __syncthreads();
    }
    const int base = (int)blockIdx.x * (HASH_THREADS * HASH_ITEMS) + (int)threadIdx.x;
#pragma unroll
    for (int r = 0; r < HASH_ITEMS; ++r) {
        const int i = base + r * HASH_THREADS;
        if (i < n) {
            const int I = sub ? sub[i] : i;
            const float2 p = pos[I];
            // src/cuda/hash.cu:17-22 / src/cpu/hash.cpp:20-25: floor(x * (rsh/L)), clamp
            int ix = (int)floorf(p.x * inv_hx);
            int iy = (int)floorf(p.y * inv_hy);
            ix = min(max(ix, 0), rshx - 1);
            iy = min(max(iy, 0), rshy - 1);
            const int cell = ix * rshy + iy;
            vp_cell[i] = cell;
            if (LDS) atomicAdd(&h[cell], 1);
            else atomicAdd(cell_num + cell, 1);
        }
    }
    if (LDS) {
        __syncthreads();
        for (int c = threadIdx.x; c < cells; c += HASH_THREADS)
            if (h[c]) atomicAdd(cell_num + c, h[c]);
    }
}

}  // namespace vpmc

using namespace vpmc;

extern "C" {

size_t vpmc_hash_workspace_bytes(int n, int cells)
{
    const size_t m = (size_t)(n > 0 ? n : 1);
    const size_t c = (size_t)(cells > 0 ? cells : 1);
    const size_t fill = scan_workspace_bytes((int)c, sizeof(int));
    const size_t sort = align256(m * 4) + align256((c + 1) * 4) + radix_workspace_bytes((int)m, 4);
    return std::max(fill, sort);
}

int vpmc_fill_cells(int rshx, int rshy, float Lx, float Ly, int n, const int *sub,
                    const float *pos_xy, int *cell_num, int *cell_acc, int *vp_cell,
                    void *workspace, size_t workspace_bytes, void *stream)
{
    if (rshx <= 0 || rshy <= 0 || n < 0 || !cell_num || !cell_acc) return VPMC_ERR_ARG;
    if (n > 0 && (!pos_xy || !vp_cell)) return VPMC_ERR_ARG;
    hipStream_t s = as_stream(stream);
    const int cells = rshx * rshy;
    VPMC_TRY(fill_i32(cell_num, 0, (size_t)cells, s));
    // include/spatial_hash.hpp / src/cuda/hash.cu:49-50: inv_h = rsh / L in FP32
    const float inv_hx = (float)rshx / Lx, inv_hy = (float)rshy / Ly;
    if (n > 0) {
        const dim3 grid((unsigned)((n + HASH_THREADS * HASH_ITEMS - 1) / (HASH_THREADS * HASH_ITEMS)));
        if (cells <= HASH_LDS_CELLS)
            hipLaunchKernelGGL(k_hash_keys<true>, grid, dim3(HASH_THREADS), 0, s, rshx, rshy,
                               inv_hx, inv_hy, n, sub, (const float2 *)pos_xy, cell_num, vp_cell);
        else
            hipLaunchKernelGGL(k_hash_keys<false>, grid, dim3(HASH_THREADS), 0, s, rshx, rshy,
                               inv_hx, inv_hy, n, sub, (const float2 *)pos_xy, cell_num, vp_cell);
        VPMC_CHECK_LAUNCH();
    }
    return scan_i32(cell_num, cell_acc, cells, false, nullptr, workspace, workspace_bytes, s);
}

int vpmc_counting_sort(int n, int cells, int *cell_num, const int *cell_acc,
                       const int *vp_cell, int *vp_map, void *workspace,
                       size_t workspace_bytes, void *stream)
{
    if (n < 0 || cells <= 0) return VPMC_ERR_ARG;
    if (n == 0) return VPMC_OK;
    if (!cell_num || !cell_acc || !vp_cell || !vp_map) return VPMC_ERR_ARG;
    hipStream_t s = as_stream(stream);
    Carve cv(workspace, workspace_bytes);
    uint32_t *keys_sorted = cv.take<uint32_t>((size_t)n);
    int *start = cv.take<int>((size_t)cells + 1);
    const size_t rb = radix_workspace_bytes(n, 4);
    void *rws = cv.take<char>(rb);
    if (!keys_sorted || !start || !rws) return VPMC_ERR_WORKSPACE;
    // keys lie in [0, cells): start[] receives the same offsets as cell_acc
    VPMC_TRY(radix_sort_indices((const uint32_t *)vp_cell, n, key_bits_for((uint32_t)(cells - 1)),
                                (uint32_t *)vp_map, keys_sorted, cells, start, rws, rb, s));
    return fill_i32(cell_num, 0, (size_t)cells, s);
}

}  // extern "C"
