// spatial_hash.hpp -- cell-list hash of one particle set.
// Replaces include/spatial_hash.hpp:7-73 for the GPU path.  The *_cuda free
// functions keep the reference's signatures (device pointers) and run the
// HIP kernels of libvpmc; CountingSort_cuda is stable (ascending local index
// within a cell, the order of CountingSort_cpu).  The CPU/OpenMP variants of
// the reference are not part of this library.
#pragma once

#include "device_vector_alias.hpp"

struct SpatialHash {
    int rshx{}, rshy{};
    dvec<int> cell_num, cell_acc;   // rshx*rshy
    dvec<int> vp_cell, vp_sort;     // N_max

    void resize_cells(int rx, int ry)
    {
        rshx = rx;
        rshy = ry;
        size_t cells = size_t(rx) * ry;
        cell_num.resize(cells);
        cell_acc.resize(cells);
    }
    void resize_particles(int N_max)
    {
        vp_cell.resize(N_max);
        vp_sort.resize(N_max);
    }
};

// cell keys ix*rshy+iy of pos[sub[i]], histogram cell_num, exclusive scan
// cell_acc (src/cuda/hash.cu:42-69).  Synchronises the device, as the
// reference does.
void FillCells_cuda(int rshx, int rshy, float Lx, float Ly, int N, const int *sub,
                    const float2 *pos, int *cell_num, int *cell_acc, int *vp_cell);

// vp_map = local indices grouped by cell (src/cuda/hash.cu:71-91, stable);
// cell_num is consumed (ends at zero).
void CountingSort_cuda(int N, int *cell_num, const int *cell_acc, const int *vp_cell,
                       int *vp_map);
