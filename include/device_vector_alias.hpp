// device_vector_alias.hpp -- device memory container of the C++ API.
//
// Replaces include/device_vector_alias.hpp of the reference, where dvec<T> is
// thrust::device_vector<T> (:26).  Here dvec<T> is a small RAII owner of a
// hipMalloc allocation with the subset of the interface the path uses:
// resize (new elements zeroed), size, data, begin/end, copy, host transfer.
// Vector types are HIP's float2 / float4 (member w, as in CUDA).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#ifndef HD
#define HD __host__ __device__
#endif

namespace vpmc_detail {
inline void hip_check(hipError_t e, const char *what)
{
    if (e != hipSuccess)
        throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}
}  // namespace vpmc_detail

template <typename T> class dvec {
public:
    dvec() = default;
    explicit dvec(size_t n) { resize(n); }
    explicit dvec(const std::vector<T> &h) { assign_host(h.data(), h.size()); }
    dvec(const dvec &o) { *this = o; }
    dvec(dvec &&o) noexcept { swap(o); }
    ~dvec() { release(); }

    dvec &operator=(const dvec &o)
    {
        if (this == &o) return *this;
        resize(o.n_);
        if (o.n_)
            vpmc_detail::hip_check(hipMemcpy(p_, o.p_, o.n_ * sizeof(T), hipMemcpyDeviceToDevice),
                                   "dvec copy");
        return *this;
    }
    dvec &operator=(dvec &&o) noexcept
    {
        if (this != &o) {
            release();
            swap(o);
        }
        return *this;
    }

    void resize(size_t n)
    {
        if (n > cap_) {
            T *q = nullptr;
            vpmc_detail::hip_check(hipMalloc(&q, n * sizeof(T)), "dvec hipMalloc");
            if (n_)
                vpmc_detail::hip_check(hipMemcpy(q, p_, n_ * sizeof(T), hipMemcpyDeviceToDevice),
                                       "dvec grow");
            if (p_) (void)hipFree(p_);
            p_ = q;
            cap_ = n;
        }
        if (n > n_)
            vpmc_detail::hip_check(hipMemset(p_ + n_, 0, (n - n_) * sizeof(T)), "dvec zero");
        n_ = n;
    }
    void clear() { n_ = 0; }
    HD size_t size() const { return n_; }
    bool empty() const { return n_ == 0; }
    HD T *data() { return p_; }
    HD const T *data() const { return p_; }
    T *begin() { return p_; }
    T *end() { return p_ + n_; }
    const T *begin() const { return p_; }
    const T *end() const { return p_ + n_; }

    void assign_host(const T *h, size_t n)
    {
        resize(n);
        if (n)
            vpmc_detail::hip_check(hipMemcpy(p_, h, n * sizeof(T), hipMemcpyHostToDevice),
                                   "dvec upload");
    }
    std::vector<T> to_host() const
    {
        std::vector<T> h(n_);
        if (n_)
            vpmc_detail::hip_check(hipMemcpy(h.data(), p_, n_ * sizeof(T), hipMemcpyDeviceToHost),
                                   "dvec download");
        return h;
    }

private:
    void release()
    {
        if (p_) (void)hipFree(p_);
        p_ = nullptr;
        n_ = cap_ = 0;
    }
    void swap(dvec &o) noexcept
    {
        std::swap(p_, o.p_);
        std::swap(n_, o.n_);
        std::swap(cap_, o.cap_);
    }
    T *p_ = nullptr;
    size_t n_ = 0, cap_ = 0;
};

template <typename Vec> HD auto raw_ptr(Vec &v) { return v.data(); }
template <typename Vec> HD auto raw_ptr(const Vec &v) { return v.data(); }
