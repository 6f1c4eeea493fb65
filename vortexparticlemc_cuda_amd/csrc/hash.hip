// hash.hip -- counting-sort spatial hash on the reference's cell grid.
//
// Pass 1 computes FP32 cell keys and an atomic histogram, pass 2 scans the
// histogram (scan.hip), pass 3 scatters with a per-cell atomic countdown
// (the reference's counting_sort_kernel, src/cuda/hash.cu:27-39, which leaves
// cell_num at 0), and pass 4 restores the CPU path's stable order
// (src/cpu/hash.cpp:42-46) by ranking each element among its cell's members.
// Pass 4 costs the cell occupancy per element; cells hold O(10-1000)
// particles here.
#include "vpmc_internal.hpp"

namespace vpmc {

__global__ void k_hash_keys(int rshx, int rshy, float inv_hx, float inv_hy, int n,
                            const int *__restrict__ sub, const float2 *__restrict__ pos,
                            int *__restrict__ cell_num, int *__restrict__ vp_cell)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int I = sub ? sub[i] : i;
    float2 p = pos[I];
    // src/cuda/hash.cu:17-22 / src/cpu/hash.cpp:20-25: floor(x * (rsh/L)), clamp
    int ix = (int)floorf(p.x * inv_hx);
    int iy = (int)floorf(p.y * inv_hy);
    ix = min(max(ix, 0), rshx - 1);
    iy = min(max(iy, 0), rshy - 1);
    int cell = ix * rshy + iy;
    vp_cell[i] = cell;
    atomicAdd(cell_num + cell, 1);
}

__global__ void k_hash_scatter(int n, const int *__restrict__ vp_cell, int *__restrict__ cell_num,
                               const int *__restrict__ cell_acc, int *__restrict__ tmp)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int c = vp_cell[i];
    int slot = atomicSub(cell_num + c, 1) - 1;
    tmp[cell_acc[c] + slot] = i;
}

__global__ void k_hash_fixup(int n, int cells, const int *__restrict__ vp_cell,
                             const int *__restrict__ cell_acc, const int *__restrict__ tmp,
                             int *__restrict__ vp_map)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    int i = tmp[t];
    int c = vp_cell[i];
    int beg = cell_acc[c];
    int end = (c + 1 < cells) ? cell_acc[c + 1] : n;
    int rank = 0;
    for (int k = beg; k < end; ++k) rank += (tmp[k] < i) ? 1 : 0;
    vp_map[beg + rank] = i;
}

}  // namespace vpmc

using namespace vpmc;

extern "C" {

size_t vpmc_hash_workspace_bytes(int n, int cells)
{
    return scan_workspace_bytes(cells, sizeof(int)) + align256((size_t)(n > 0 ? n : 1) * 4);
}

int vpmc_fill_cells(int rshx, int rshy, float Lx, float Ly, int n, const int *sub,
                    const float *pos_xy, int *cell_num, int *cell_acc, int *vp_cell,
                    void *workspace, size_t workspace_bytes, void *stream)
{
    if (rshx <= 0 || rshy <= 0 || n < 0 || !cell_num || !cell_acc) return VPMC_ERR_ARG;
    if (n > 0 && (!pos_xy || !vp_cell)) return VPMC_ERR_ARG;
    hipStream_t s = as_stream(stream);
    int cells = rshx * rshy;
    VPMC_TRY(fill_i32(cell_num, 0, (size_t)cells, s));
    // include/spatial_hash.hpp / src/cuda/hash.cu:49-50: inv_h = rsh / L in FP32
    float inv_hx = (float)rshx / Lx, inv_hy = (float)rshy / Ly;
    if (n > 0) {
        hipLaunchKernelGGL(k_hash_keys, dim3((n + 255) / 256), dim3(256), 0, s, rshx, rshy,
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
    cv.take<char>(scan_workspace_bytes(cells, sizeof(int)));
    int *tmp = cv.take<int>((size_t)n);
    if (!tmp) return VPMC_ERR_WORKSPACE;
    hipLaunchKernelGGL(k_hash_scatter, dim3((n + 255) / 256), dim3(256), 0, s, n, vp_cell,
                       cell_num, cell_acc, tmp);
    hipLaunchKernelGGL(k_hash_fixup, dim3((n + 255) / 256), dim3(256), 0, s, n, cells, vp_cell,
                       cell_acc, (const int *)tmp, vp_map);
    VPMC_CHECK_LAUNCH();
    return VPMC_OK;
}

}  // extern "C"
