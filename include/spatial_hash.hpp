// spatial_hash.hpp -- cell-list hash of one particle set.
// Replaces include/spatial_hash.hpp:7-73.  Every variant keeps the
// reference's signature and runs the HIP kernels of libvpmc: the *_cuda
// functions take device pointers, the *_cpu / *_omp functions host spans
// (copied to the device and back).  CountingSort is stable in all of them
// (ascending local index within a cell, the order of CountingSort_cpu; the
// reference's CountingSort_omp writes from a wrong base, hash_omp.cpp:65).
#pragma once

#include <span>

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

// src/cpu/hash.cpp:6-34 / src/omp/hash_omp.cpp:12-54: keys of (px, py)[sub[i]]
// for i < N_current, histogram, exclusive scan; host spans
void FillCells_cpu(int rshx, int rshy, float Lx, float Ly, int N_current,
                   std::span<const int> sub, std::span<const float> pos_x,
                   std::span<const float> pos_y, std::span<int> cell_num, std::span<int> cell_acc,
                   std::span<int> vp_cell);
void FillCells_omp(int rshx, int rshy, float Lx, float Ly, int N_current,
                   std::span<const int> sub, std::span<const float> pos_x,
                   std::span<const float> pos_y, std::span<int> cell_num, std::span<int> cell_acc,
                   std::span<int> vp_cell);

// src/cpu/hash.cpp:36-47 / src/omp/hash_omp.cpp:56-69: stable counting sort of
// the local indices by cell; cell_num is consumed (ends at zero); host spans
void CountingSort_cpu(int N_current, std::span<int> cell_num, std::span<const int> cell_acc,
                      std::span<const int> vp_cell, std::span<int> vp_map);
void CountingSort_omp(int N_current, std::span<int> cell_num, std::span<const int> cell_acc,
                      std::span<const int> vp_cell, std::span<int> vp_map);
