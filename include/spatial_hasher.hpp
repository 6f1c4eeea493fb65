// spatial_hasher.hpp -- backend-selecting hash builder.
// Replaces include/spatial_hasher.hpp:6-21 (whose implementation,
// src/spatial_hasher.cu, does not build).  Backend::CUDA names the GPU
// backend and runs on HIP; Backend::HIP is an alias.  CPU and OMP are not part
// of this library: selecting them throws std::invalid_argument.
#pragma once

#include <stdexcept>

#include "spatial_hash.hpp"

enum class Backend { CPU, OMP, CUDA, HIP = CUDA };

class SpatialHasher {
public:
    SpatialHasher(int rshx, int rshy, Backend b) : backend_(b)
    {
        if (b != Backend::CUDA)
            throw std::invalid_argument("SpatialHasher: only the GPU backend (Backend::CUDA / "
                                        "Backend::HIP) is provided by libvpmc");
        hash_.resize_cells(rshx, rshy);
    }

    void resize(int N_max) { hash_.resize_particles(N_max); }

    // pos: device float2 positions; sub: device local -> global indices
    template <typename VecPos, typename VecSub>
    void build(const VecPos &pos, const VecSub &sub, float Lx, float Ly)
    {
        int N = static_cast<int>(sub.size());
        if ((size_t)N > hash_.vp_cell.size()) hash_.resize_particles(N);
        FillCells_cuda(hash_.rshx, hash_.rshy, Lx, Ly, N, raw_ptr(sub), raw_ptr(pos),
                       raw_ptr(hash_.cell_num), raw_ptr(hash_.cell_acc), raw_ptr(hash_.vp_cell));
        CountingSort_cuda(N, raw_ptr(hash_.cell_num), raw_ptr(hash_.cell_acc),
                          raw_ptr(hash_.vp_cell), raw_ptr(hash_.vp_sort));
    }

    const SpatialHash &data() const { return hash_; }

private:
    SpatialHash hash_;
    Backend backend_;
};
