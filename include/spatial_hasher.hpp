// spatial_hasher.hpp -- backend-selecting hash builder.
// Replaces include/spatial_hasher.hpp:6-21 (whose implementation,
// src/spatial_hasher.cu, does not build).  Every backend builds on the GPU:
// Backend::CUDA names the GPU backend (Backend::HIP is an alias), and CPU /
// OMP are kept so that the reference's callers compile and get the same
// hash.  build() takes device containers (dvec) or host ones (std::vector or
// anything with data() and size(); copied to the device first).
#pragma once

#include <type_traits>

#include "spatial_hash.hpp"

enum class Backend { CPU, OMP, CUDA, HIP = CUDA };

namespace vpmc_detail {
template <class T> struct is_dvec : std::false_type {};
template <class T> struct is_dvec<dvec<T>> : std::true_type {};

// device pointer to the container's elements (host containers are copied to tmp)
template <class T, class Vec> const T *device_view(const Vec &v, dvec<T> &tmp)
{
    using E = std::remove_cv_t<std::remove_reference_t<decltype(*v.data())>>;
    static_assert(sizeof(E) == sizeof(T), "element layout must match");
    if constexpr (is_dvec<Vec>::value) {
        return reinterpret_cast<const T *>(raw_ptr(v));
    } else {
        tmp.assign_host(reinterpret_cast<const T *>(v.data()), v.size());
        return raw_ptr(tmp);
    }
}
}  // namespace vpmc_detail

class SpatialHasher {
public:
    SpatialHasher(int rshx, int rshy, Backend b) : backend_(b) { hash_.resize_cells(rshx, rshy); }

    void resize(int N_max) { hash_.resize_particles(N_max); }

    // pos: float2-layout positions (all particles); sub: local -> global indices
    template <typename VecPos, typename VecSub>
    void build(const VecPos &pos, const VecSub &sub, float Lx, float Ly)
    {
        int N = static_cast<int>(sub.size());
        if ((size_t)N > hash_.vp_cell.size()) hash_.resize_particles(N);
        dvec<float2> tp;
        dvec<int> ts;
        const float2 *p = vpmc_detail::device_view<float2>(pos, tp);
        const int *s = vpmc_detail::device_view<int>(sub, ts);
        FillCells_cuda(hash_.rshx, hash_.rshy, Lx, Ly, N, s, p, raw_ptr(hash_.cell_num),
                       raw_ptr(hash_.cell_acc), raw_ptr(hash_.vp_cell));
        CountingSort_cuda(N, raw_ptr(hash_.cell_num), raw_ptr(hash_.cell_acc),
                          raw_ptr(hash_.vp_cell), raw_ptr(hash_.vp_sort));
    }

    const SpatialHash &data() const { return hash_; }
    Backend backend() const { return backend_; }

private:
    SpatialHash hash_;
    Backend backend_;
};
