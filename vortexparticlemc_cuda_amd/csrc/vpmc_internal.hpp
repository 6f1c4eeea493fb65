// vpmc_internal.hpp -- shared device helpers and host plumbing for libvpmc.
#pragma once
#include <vector>

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "vpmc.h"

namespace vpmc {

constexpr float PI_F = 3.14159265358979323846f;
constexpr double PI_D = 3.14159265358979323846;
constexpr int WAVE = 64;

// ---- status ---------------------------------------------------------------
void set_hip_error(hipError_t e);
#define VPMC_CHECK_LAUNCH()                                                    \
    do {                                                                       \
        hipError_t e_ = hipGetLastError();                                     \
        if (e_ != hipSuccess) { ::vpmc::set_hip_error(e_); return VPMC_ERR_HIP; } \
    } while (0)
#define VPMC_TRY(x)                                                            \
    do {                                                                       \
        int s_ = (x);                                                          \
        if (s_ != VPMC_OK) return s_;                                          \
    } while (0)

inline hipStream_t as_stream(void *s) { return reinterpret_cast<hipStream_t>(s); }
inline size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

// records a caller's timing event (hipEvent_t) on s; no-op for NULL
inline int record_event(void *ev, hipStream_t s)
{
    if (!ev) return VPMC_OK;
    hipError_t e = hipEventRecord(reinterpret_cast<hipEvent_t>(ev), s);
    if (e != hipSuccess) { set_hip_error(e); return VPMC_ERR_HIP; }
    return VPMC_OK;
}

// bump allocator over a caller workspace
struct Carve {
    char *base;
    size_t cap, used = 0;
    Carve(void *p, size_t c) : base(static_cast<char *>(p)), cap(c) {}
    template <class T> T *take(size_t count) {
        size_t b = align256(count * sizeof(T));
        if (used + b > cap) return nullptr;
        T *p = reinterpret_cast<T *>(base + used);
        used += b;
        return p;
    }
};

// A binned sample as the sort moves it (8 bytes): the query point as 16-bit
// offsets inside its fine cell (x in the high half) and the sample id.  The
// cell is the sample's tile, known from its key, so the offsets carry the
// point to c / 2^17 (c the cell side): below the FP32 spacing of coordinates
// near 1 for cells up to 2^-7, a few ulp on the coarsest grids used.
struct SRec {
    uint32_t xy;
    uint32_t id;
};

constexpr float SREC_Q = 65536.f;

// offsets of point (qx, qy) in cell (cx, cy) of side (cw, ch), clamped into
// the cell (the key's cell comes from qx * inv_cx, a rounding apart)
__host__ __device__ inline uint32_t srec_pack(float qx, float qy, int cx, int cy, float cw,
                                              float ch)
{
    const float ux = (qx - (float)cx * cw) * (SREC_Q / cw);
    const float uy = (qy - (float)cy * ch) * (SREC_Q / ch);
    const int ix = ux < 0.f ? 0 : (ux > 65535.f ? 65535 : (int)ux);
    const int iy = uy < 0.f ? 0 : (uy > 65535.f ? 65535 : (int)uy);
    return ((uint32_t)ix << 16) | (uint32_t)iy;
}

// the point relative to its cell's centre: the midpoint of the offset's bin
__host__ __device__ inline void srec_offset(uint32_t xy, float cw, float ch, float &dx, float &dy)
{
    dx = ((float)(xy >> 16) + (0.5f - 0.5f * SREC_Q)) * (cw * (1.f / SREC_Q));
    dy = ((float)(xy & 0xffffu) + (0.5f - 0.5f * SREC_Q)) * (ch * (1.f / SREC_Q));
}

// ---- scans (scan.hip) -------------------------------------------------------
constexpr int SCAN_THREADS = 256;
constexpr int SCAN_ITEMS = 8;
constexpr int SCAN_CHUNK = SCAN_THREADS * SCAN_ITEMS;
inline int scan_blocks(int n) { return (n + SCAN_CHUNK - 1) / SCAN_CHUNK; }
size_t scan_workspace_bytes(int n, size_t elem);
// out[i] = sum_{j<i} in[j] (exclusive) or sum_{j<=i} (inclusive);
// total (optional, device) receives the grand total.  In-place allowed.
int scan_i32(const int *in, int *out, int n, bool inclusive, int *total, void *ws,
             size_t ws_bytes, hipStream_t s);
int scan_f64(const double *in, double *out, int n, bool inclusive, double *total, void *ws,
             size_t ws_bytes, hipStream_t s);
int fill_i32(int *p, int v, size_t n, hipStream_t s);
// stable LSD radix sort of tile keys (sort.hip)
size_t radix_workspace_bytes(int n, size_t rec_bytes = 16);   // rec_bytes: 16 (float4), 8 (SRec), 4 (indices)
int radix_sort_to_records(const uint32_t *keys, int n, int key_bits, const float2 *q,
                          const float4 *recs0, float4 *rec, uint32_t *keys_sorted, int nkeys,
                          int *start, void *ws, size_t ws_bytes, hipStream_t s);
// the same with 8-byte sample records (point offsets, sample id)
int radix_sort_samples(const uint32_t *keys, int n, int key_bits, const SRec *recs0, SRec *rec,
                       uint32_t *keys_sorted, int nkeys, int *start, void *ws, size_t ws_bytes,
                       hipStream_t s);
// the same carrying only the input positions: idx = the stable argsort of keys
// (workspace radix_workspace_bytes(n, 4))
int radix_sort_indices(const uint32_t *keys, int n, int key_bits, uint32_t *idx,
                       uint32_t *keys_sorted, int nkeys, int *start, void *ws, size_t ws_bytes,
                       hipStream_t s);
// segmented variant: blocks from radix_seg_blocks (segments = contiguous
// ranges seg_off[g] .. seg_off[g+1] whose keys lie in [seg_base[g],
// seg_base[g] + 2^local_bits)); keys globally sorted on return
int radix_seg_blocks(const std::vector<int> &seg_off, const std::vector<uint32_t> &seg_base,
                     std::vector<int4> *blk, std::vector<uint32_t> *blk_base);
size_t radix_seg_workspace_bytes(int n, int nblk);
int radix_sort_segmented(const uint32_t *keys, int n, int local_bits, const SRec *recs0,
                         SRec *rec, uint32_t *keys_sorted, int nkeys, int *start,
                         const int4 *blk, const uint32_t *blk_base, int nblk, void *ws,
                         size_t ws_bytes, hipStream_t s);
inline int key_bits_for(uint32_t max_key)
{
    int b = 1;
    while (b < 32 && (max_key >> b) != 0u) ++b;
    return b;
}
int fill_f32(float *p, float v, size_t n, hipStream_t s);

// ---- field (field.hip) -------------------------------------------------------
struct FieldGeom {           // device-side copy of vpmc_field's geometry
    int nx, ny;
    float Lx, Ly;
    float inv_cx, inv_cy;    // nx / Lx, ny / Ly
    float cw, ch;            // Lx / nx, Ly / ny
    int reach;               // columns each side
    float rs, rs2;           // search radius (h_w, times the deformation bound
                             // with deformed records, plus rounding margin)
    float rs1;               // h_w plus the same margin (one record's reach
                             // is sigma_max(F) rs1)
};

// G = F^-T F^-1 of F = (a b; c d), stored (G11, 2 G12, G22, sigma_max(F)^2):
// |F^-1 d|^2 = G11 dx^2 + 2 G12 dx dy + G22 dy^2, and W(|F^-1 d|) > 0 only
// for |d| < sigma_max(F) h_w.  sigma_max^2 is the larger eigenvalue of
// F F^T = (p r; r q): (p + q + sqrt((p - q)^2 + 4 r^2)) / 2, a sum of
// squares under the root (no cancellation near rotations).  Every multiply-add
// is an explicit fmaf, so that each kernel inlining this rounds alike (with
// implicit contraction the compiler may fuse a different product per kernel).
__device__ __forceinline__ float4 quad_form(float4 F)
{
    const float det = fmaf(F.x, F.w, -(F.y * F.z));
    const float i2 = 1.f / (det * det);
    const float p = fmaf(F.x, F.x, F.y * F.y), q = fmaf(F.z, F.z, F.w * F.w);
    const float r = fmaf(F.x, F.z, F.y * F.w);
    const float pq = p - q;
    const float smax2 = 0.5f * ((p + q) + sqrtf(fmaf(pq, pq, 4.f * (r * r))));
    return make_float4(q * i2, -2.f * (r * i2), p * i2, smax2 * (1.f + 1e-5f));
}
int field_geom(const vpmc_field *f, float h_w, FieldGeom *g);
int field_build_impl(int nx, int ny, float Lx, float Ly, int n, const int *n_dev,
                     const int *sub, const float *pos_xy, const float *w, int abs_w,
                     float *pts, int *cell_start, int *order, float *set_xy, void *ws,
                     size_t ws_bytes, hipStream_t s, const float *jac = nullptr,
                     float *gq = nullptr, float *set_jac = nullptr);

// ---- device math --------------------------------------------------------------
__device__ __forceinline__ float u01(uint32_t x) { return (float)(x >> 8) * 0x1p-24f; }

// Philox4x32-10 (Salmon et al., SC'11)
__device__ __forceinline__ uint4 philox10(uint4 c, uint2 k)
{
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        // one 32 x 32 -> 64-bit product per multiplier gives both halves (a
        // single v_mad_u64_u32 instead of a mulhi and a mullo)
        const uint64_t p0 = (uint64_t)0xD2511F53u * c.x, p1 = (uint64_t)0xCD9E8D57u * c.z;
        const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        c = make_uint4(hi1 ^ c.y ^ k.x, lo1, hi0 ^ c.w ^ k.y, lo0);
        k.x += 0x9E3779B9u;
        k.y += 0xBB67AE85u;
    }
    return c;
}

// include/utilities.hpp:7-17, branch-free: W = sig2 (2 t^3 - 8 u^3) with
// t = max(1-q, 0), u = max(1/2-q, 0) (equal to 1 - 6q^2 + 6q^3 for q <= 1/2).
__device__ __forceinline__ float spline_sig2(float h_w)
{
    return 40.f / (7.f * PI_F * h_w * h_w);
}

__device__ __forceinline__ double spline_d(double r, double h_w)
{
    const double sig2 = 40.0 / (7.0 * PI_D * h_w * h_w);
    double q = r / h_w;
    if (q <= 0.5) return sig2 * (6.0 * (q * q * q - q * q) + 1.0);
    if (q < 1.0) return sig2 * 2.0 * (1.0 - q) * (1.0 - q) * (1.0 - q);
    return 0.0;
}

// std::upper_bound(cdf, cdf+n, u) through a guide table (guide[k] =
// upper_bound(cdf, k / 2^glog2)); j == n -> n-1 as sample_vorticity_idx_cdf
// does (src/cpu/monte_carlo_sampling.cpp:23-33).  Any float is accepted:
// injected draws may equal 1 (std::uniform_real_distribution<float> can return
// 1.0f) and the bucket index is clamped to the table, so u >= 1 gives n-1 and
// u < 0 gives 0 exactly as upper_bound does.
// the guide bucket of draw u (0 or M - 1 outside [0, 1), NaN included: no
// cdf value exceeds such a draw)
__device__ __forceinline__ int guide_bucket(float u, int glog2)
{
    const int M = 1 << glog2;
    if (u >= 0.f && u < 1.f) return min((int)(u * (float)M), M - 1);
    return (u < 0.f) ? 0 : M - 1;
}

__device__ __forceinline__ int guided_upper_bound(float u, const double *__restrict__ cdf,
                                                  const int *__restrict__ guide, int glog2, int n)
{
    const int kb = guide_bucket(u, glog2);
    int lo = guide[kb], hi = guide[kb + 1];
    double ud = (double)u;
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (ud < cdf[mid]) hi = mid; else lo = mid + 1;
    }
    return lo >= n ? n - 1 : lo;
}

// wave-level sums (deterministic butterfly)
template <class T> __device__ __forceinline__ T wave_sum(T v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, WAVE);
    return v;
}

}  // namespace vpmc
