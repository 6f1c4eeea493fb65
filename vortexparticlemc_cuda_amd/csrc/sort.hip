// sort.hip -- stable LSD radix sort of (key, value) pairs without global atomics.
//
// Random-key binning with one global atomic per element is bound by the
// device's atomic throughput (about 2e10/s on MI355X, 1 ms for 2e7 keys).
// Here every pass over an 8-bit digit is:
//   hist     one block per RS_SUPER contiguous keys: digit histogram in LDS,
//            written digit-major (hist[d * nblk + b]);
//   scan     exclusive prefix of hist: the output offset of (digit, block);
//   scatter  the block re-reads its keys in order, RS_TILE keys per round;
//            the rank of a key among equal digits of its round comes from
//            wave ballots (eight ballots give the lanes sharing a digit) and
//            per-wave counts in LDS, and a running offset per digit carries
//            the block's position from round to round, so the pass is stable
//            and the result is the unique stable order.
// A block covers RS_ROUNDS rounds so that the digit-major histogram, whose
// 256 entries per block are written and read at a stride of nblk (one cache
// line each), holds a third of the entries that one block per round would
// need; the LDS of the scatter stays that of one round.  A round of 6144
// keys (24 a thread) writes digit runs of 24 keys on average: longer runs
// fill more of each cache line they touch (with 8-byte records, 4096-key
// rounds took 7% longer a pass; 78 KB of LDS keeps two blocks a CU).
// The sample records (8 bytes: point offsets and id) move with the keys in
// every pass, staged through LDS; the fields' 16-byte records use the same
// passes with a gather.
//
// Segmented form: when the input is already grouped into segments (an
// ensemble's simulations, each owning the key range [base_s, base_s + 2^b)),
// blocks are cut per segment and the histogram is laid out [segment][digit]
// [block of the segment]; one global exclusive scan of it then yields
// offsets that keep every segment in place, and only the b local bits need
// passes (two 8-bit passes instead of three for 2^20 global keys).
#include <algorithm>
#include <vector>

#include "vpmc_internal.hpp"

namespace vpmc {

constexpr int RS_THREADS = 256;
constexpr int RS_ITEMS = 24;
constexpr int RS_TILE = RS_THREADS * RS_ITEMS;   // keys per round
constexpr int RS_ROUNDS = 3;                     // rounds per block
constexpr int RS_SUPER = RS_TILE * RS_ROUNDS;    // keys per block
constexpr int RS_BINS = 256;                     // 8-bit digits
constexpr int RS_WAVES = RS_THREADS / WAVE;
static_assert(RS_THREADS == RS_BINS, "one thread per digit");

// a block's input range, histogram column (digit d at h0 + d * hstride) and
// key base; blk == NULL: block b of the plain sort
struct RsBlk {
    int base, cnt, h0, hstride;
    uint32_t kbase;
};
// Workgroups are dealt round-robin to the 8 XCDs.  Block b of the sort is
// workgroup (b % per) * 8 + b / per, so that each XCD runs a contiguous range
// of blocks: block b's digit runs end in the cache lines where block b+1's
// runs of the same digits begin, and both then fill those lines in one L2.
__device__ __forceinline__ int rs_logical_block(int nblk)
{
    const int per = (nblk + 7) >> 3;
    return (int)(blockIdx.x & 7) * per + (int)(blockIdx.x >> 3);
}
inline unsigned rs_grid(int nblk) { return 8u * (unsigned)((nblk + 7) >> 3); }

__device__ __forceinline__ RsBlk rs_block(int b, const int4 *__restrict__ blk,
                                          const uint32_t *__restrict__ blk_base, int n, int nblk)
{
    if (blk) {
        const int4 v = blk[b];
        return RsBlk{v.x, v.y, v.z, v.w, blk_base[b]};
    }
    const int base = b * RS_SUPER;
    return RsBlk{base, min(RS_SUPER, n - base), b, nblk, 0u};
}

// LDS copies of the digit histogram: lane l counts into copy l % RS_HCOPY,
// so that lanes of one wave whose keys share a digit (frequent: samples
// cluster where the vorticity is) rarely add to the same LDS word
constexpr int RS_HCOPY = 16;

__device__ __forceinline__ void rs_hist_add(int *h, uint32_t key, uint32_t kbase, int shift, int c)
{
    atomicAdd(&h[(((key - kbase) >> shift) & (RS_BINS - 1)) * RS_HCOPY + c], 1);
}

__global__ void __launch_bounds__(RS_THREADS)
    k_rs_hist(const uint32_t *__restrict__ keys, int n, int shift, int nblk,
              const int4 *__restrict__ blk, const uint32_t *__restrict__ blk_base,
              int *__restrict__ hist)
{
    const int lb = rs_logical_block(nblk);
    if (lb >= nblk) return;   // uniform over the workgroup
    __shared__ __align__(16) int h[RS_BINS * RS_HCOPY];
    for (int i = threadIdx.x; i < RS_BINS * RS_HCOPY; i += RS_THREADS) h[i] = 0;
    __syncthreads();
    const RsBlk B = rs_block(lb, blk, blk_base, n, nblk);
    const uint32_t *kp = keys + B.base;
    const int c = (int)threadIdx.x & (RS_HCOPY - 1);
    if (((reinterpret_cast<uintptr_t>(kp)) & 15) == 0) {
        const uint4 *k4 = reinterpret_cast<const uint4 *>(kp);
        const int n4 = B.cnt >> 2;
        for (int r = 0; r < RS_SUPER / 4 / RS_THREADS; ++r) {
            const int e = r * RS_THREADS + threadIdx.x;
            if (e < n4) {
                const uint4 v = k4[e];
                rs_hist_add(h, v.x, B.kbase, shift, c);
                rs_hist_add(h, v.y, B.kbase, shift, c);
                rs_hist_add(h, v.z, B.kbase, shift, c);
                rs_hist_add(h, v.w, B.kbase, shift, c);
            }
        }
        const int e = 4 * n4 + (int)threadIdx.x;
        if (threadIdx.x < 4 && e < B.cnt) rs_hist_add(h, kp[e], B.kbase, shift, c);
    } else {
        for (int r = 0; r < RS_SUPER / RS_THREADS; ++r) {
            const int e = r * RS_THREADS + threadIdx.x;
            if (e < B.cnt) rs_hist_add(h, kp[e], B.kbase, shift, c);
        }
    }
    __syncthreads();
    const int4 *row = reinterpret_cast<const int4 *>(h + threadIdx.x * RS_HCOPY);
    int sum = 0;
#pragma unroll
    for (int k = 0; k < RS_HCOPY / 4; ++k) {
        const int4 v = row[k];
        sum += v.x + v.y + v.z + v.w;
    }
    hist[(size_t)B.h0 + (size_t)threadIdx.x * B.hstride] = sum;
}

// exclusive prefix over the block of one value per thread; *total = sum
__device__ __forceinline__ int rs_block_excl_scan(int v, int *wsum, int *total)
{
    const int lane = threadIdx.x & (WAVE - 1), wid = threadIdx.x / WAVE;
    int x = v;
#pragma unroll
    for (int o = 1; o < WAVE; o <<= 1) {
        int y = __shfl_up(x, o, WAVE);
        if (lane >= o) x += y;
    }
    if (lane == WAVE - 1) wsum[wid] = x;
    __syncthreads();
    int base = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < RS_WAVES; ++w) {
        base += (w < wid) ? wsum[w] : 0;
        tot += wsum[w];
    }
    *total = tot;
    return base + x - v;
}

// One round = RS_TILE consecutive input positions of the block's range; wave
// w owns the contiguous quarter [w * RS_TILE/4, (w+1) * RS_TILE/4) and ranks
// its keys in element order against a wave-private digit histogram in LDS
// (no barrier inside the wave's loop: one wave's LDS operations complete in
// order).  One barrier later, per-digit offsets across waves give every key
// its stable position inside the round; keys and local input indices are
// placed in LDS in digit order and written out in that order, so that a
// digit's elements of this round go to one contiguous range, which starts at
// offs[d * nblk + b] in the first round and right after the previous round's
// range of digit d in the next.
// The records travel with the keys.  STG = false: after the ranking, each
// thread gathers the records of its output positions from the round's input
// slice.  STG = true (sample records and indices after the first pass): each
// thread loads the records of its own input positions with its keys, in
// coalesced loads, and places them in LDS at their ranked position, from
// where they are stored in order; this replaces the gather (one cache line
// per record) with LDS traffic at twice the LDS footprint.  Without input
// records (recs_in == NULL: the first pass of the fields' sort and of the
// hash) the record of input position i is built from it, (q[i].x, q[i].y,
// i, 0) or i, and gathered.
constexpr int RS_WITEMS = RS_TILE / RS_WAVES / WAVE;   // keys per lane

// the first pass's record of input position i: from its query point q[i]
// (float4, SRec), or the position itself (uint32_t: the hash's local index)
template <class R>
__device__ __forceinline__ R rs_rec_from(const float2 *__restrict__ q, uint32_t i);
template <>
__device__ __forceinline__ float4 rs_rec_from<float4>(const float2 *__restrict__ q, uint32_t i)
{
    const float2 p = q[i];
    return make_float4(p.x, p.y, __uint_as_float(i), 0.f);
}
template <>
__device__ __forceinline__ SRec rs_rec_from<SRec>(const float2 *__restrict__, uint32_t i)
{
    return SRec{0u, i};   // sample sorts always pass their records
}
template <>
__device__ __forceinline__ uint32_t rs_rec_from<uint32_t>(const float2 *__restrict__, uint32_t i)
{
    return i;
}

template <class R, bool STG>
__global__ void __launch_bounds__(RS_THREADS)
    k_rs_scatter(const uint32_t *__restrict__ keys_in, const R *__restrict__ recs_in,
                 const float2 *__restrict__ q, int n, int shift, int nblk,
                 const int4 *__restrict__ blk, const uint32_t *__restrict__ blk_base,
                 const int *__restrict__ offs, uint32_t *__restrict__ keys_out,
                 R *__restrict__ recs_out)
{
    const int lb = rs_logical_block(nblk);
    if (lb >= nblk) return;   // uniform over the workgroup
    __shared__ uint32_t lk[RS_TILE];
    __shared__ uint16_t li[STG ? 1 : RS_TILE];   // local input index (< RS_TILE)
    // STG: the records in the round's digit order, as 32-bit words
    constexpr int RW = (int)(sizeof(R) / 4);
    __shared__ uint32_t lr[STG ? RS_TILE * RW : 1];
    __shared__ int wh[RS_WAVES][RS_BINS];   // per-wave digit counts, then offsets
    __shared__ int gofs[RS_BINS];           // global position - local position (this round)
    __shared__ int wsum[RS_WAVES];
    const int t = threadIdx.x, lane = t & (WAVE - 1), wid = t / WAVE;
    const RsBlk B = rs_block(lb, blk, blk_base, n, nblk);
    const uint32_t kb = B.kbase;
    // running output position of digit t for this block's next round
    int run = offs[(size_t)B.h0 + (size_t)t * B.hstride];
    const unsigned long long lt = (1ull << lane) - 1ull;
    for (int round = 0; round < RS_ROUNDS; ++round) {
        const int base = B.base + round * RS_TILE;
        const int cnt = min(RS_TILE, B.cnt - round * RS_TILE);
        if (cnt <= 0) break;   // uniform over the workgroup
#pragma unroll
        for (int w = 0; w < RS_WAVES; ++w) wh[w][t] = 0;
        __syncthreads();
        uint32_t kr[RS_WITEMS];
        int rk[RS_WITEMS];
        uint32_t rw[STG ? RS_WITEMS : 1][RW];
#pragma unroll
        for (int r = 0; r < RS_WITEMS; ++r) {
            const int e = (wid * RS_WITEMS + r) * WAVE + lane;
            const bool valid = e < cnt;
            kr[r] = valid ? keys_in[base + e] : kb;
            if constexpr (STG) {
                // one load of the whole record (dwordx2 for 8 bytes); a lane
                // past the round's end reads position base (cnt >= 1) and its
                // words are never placed
                const R v = recs_in[base + (valid ? e : 0)];
                __builtin_memcpy(rw[r], &v, sizeof(R));
            }
        }
#pragma unroll
        for (int r = 0; r < RS_WITEMS; ++r) {
            const int e = (wid * RS_WITEMS + r) * WAVE + lane;
            const bool valid = e < cnt;
            const int d = (int)(((kr[r] - kb) >> shift) & (RS_BINS - 1));
            unsigned long long peers = __ballot(valid);
#pragma unroll
            for (int b = 0; b < 8; ++b) {
                const bool bit = (d >> b) & 1;
                const unsigned long long m = __ballot(bit);
                peers &= bit ? m : ~m;
            }
            const int before = __popcll(peers & lt);
            const int prior = valid ? wh[wid][d] : 0;
            __builtin_amdgcn_wave_barrier();   // every lane's read precedes the leader's write
            rk[r] = prior + before;
            if (valid && before == 0) wh[wid][d] = prior + __popcll(peers);
            __builtin_amdgcn_wave_barrier();
        }
        __syncthreads();
        // digit t: exclusive offsets of the waves inside the round
        int c[RS_WAVES], sum = 0;
#pragma unroll
        for (int w = 0; w < RS_WAVES; ++w) {
            c[w] = wh[w][t];
            sum += c[w];
        }
        int tot;
        const int ls = rs_block_excl_scan(sum, wsum, &tot);
        gofs[t] = run - ls;
        run += sum;
        int acc = ls;
#pragma unroll
        for (int w = 0; w < RS_WAVES; ++w) {
            wh[w][t] = acc;   // column t: read above by this thread only
            acc += c[w];
        }
        __syncthreads();
#pragma unroll
        for (int r = 0; r < RS_WITEMS; ++r) {
            const int e = (wid * RS_WITEMS + r) * WAVE + lane;
            if (e < cnt) {
                const int pos = wh[wid][((kr[r] - kb) >> shift) & (RS_BINS - 1)] + rk[r];
                lk[pos] = kr[r];
                if constexpr (STG) {
#pragma unroll
                    for (int w = 0; w < RW; ++w) lr[pos * RW + w] = rw[r][w];
                } else {
                    li[pos] = (uint16_t)e;
                }
            }
        }
        __syncthreads();
        if constexpr (STG) {
#pragma unroll
            for (int j = 0; j < RS_ITEMS; ++j) {
                const int i = j * RS_THREADS + t;
                if (i < cnt) {
                    const uint32_t k = lk[i];
                    const int p = gofs[((k - kb) >> shift) & (RS_BINS - 1)] + i;
                    keys_out[p] = k;
                    uint32_t *dst = reinterpret_cast<uint32_t *>(recs_out + p);
#pragma unroll
                    for (int w = 0; w < RW; ++w) dst[w] = lr[i * RW + w];
                }
            }
        } else {
            // all RS_ITEMS record loads in flight before the stores
            uint32_t ko[RS_ITEMS];
            int po[RS_ITEMS];
            R rv[RS_ITEMS];
#pragma unroll
            for (int j = 0; j < RS_ITEMS; ++j) {
                const int i = j * RS_THREADS + t;
                po[j] = -1;
                if (i < cnt) {
                    ko[j] = lk[i];
                    const int src = base + li[i];
                    po[j] = gofs[((ko[j] - kb) >> shift) & (RS_BINS - 1)] + i;
                    rv[j] = recs_in ? recs_in[src] : rs_rec_from<R>(q, (uint32_t)src);
                }
            }
#pragma unroll
            for (int j = 0; j < RS_ITEMS; ++j) {
                if (po[j] >= 0) {
                    keys_out[po[j]] = ko[j];
                    recs_out[po[j]] = rv[j];
                }
            }
        }
        __syncthreads();   // lk, li/lr, gofs and wh are rewritten by the next round
    }
}

// start[k] = lower_bound(keys, keys + n, k) for k in [0, nkeys], one thread
// per key: runs of empty keys cost nothing extra (a position-parallel form,
// where the thread at a run's start fills the gap before it, writes a sign
// set's empty region of a field, thousands of cells, from one thread)
__global__ void k_rs_bounds_search(const uint32_t *__restrict__ keys, int n, int nkeys,
                                   int *__restrict__ start)
{
    const int k = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (k > nkeys) return;
    uint32_t lo = 0, hi = (uint32_t)n;
    while (lo < hi) {
        const uint32_t mid = lo + ((hi - lo) >> 1);
        if (keys[mid] < (uint32_t)k) lo = mid + 1;
        else hi = mid;
    }
    start[k] = (int)lo;
}

static int rs_nblk(int n) { return (n + RS_SUPER - 1) / RS_SUPER; }

size_t radix_workspace_bytes(int n, size_t rec_bytes)
{
    const size_t m = (size_t)(n > 0 ? n : 1);
    const size_t h = (size_t)RS_BINS * rs_nblk(n > 0 ? n : 1);
    return align256(m * 4) + align256(m * rec_bytes) + 2 * align256(h * 4) +
           scan_workspace_bytes((int)h, sizeof(int)) + 1024;
}

// Sorts keys[0..n) stably (key_bits significant bits) and moves with them
// the 16-byte records of the input positions i: recs0[i] if recs0 is given,
// else (q[i].x, q[i].y, i bits, 0).  rec and keys_sorted receive the sorted
// records and keys, start[0..nkeys] the first position with key >= k (every
// key must lie in [0, nkeys]).  keys and recs0 are left unchanged.  Pass p
// writes the final buffers when P-1-p is even, the workspace copies
// otherwise, so the last pass always ends in the outputs.
template <class R>
static int radix_sort_impl(const uint32_t *keys, int n, int key_bits, const float2 *q,
                           const R *recs0, R *rec, uint32_t *keys_sorted, int nkeys,
                           int *start, const int4 *blk, const uint32_t *blk_base, int nblk,
                           void *ws, size_t ws_bytes, hipStream_t s)
{
    if (n < 0 || key_bits < 1 || key_bits > 32) return VPMC_ERR_ARG;
    if (n == 0) return fill_i32(start, 0, (size_t)nkeys + 1, s);
    const size_t h = (size_t)RS_BINS * nblk;
    if (h >= ((size_t)1 << 31)) return VPMC_ERR_ARG;
    Carve cv(ws, ws_bytes);
    uint32_t *kb = cv.take<uint32_t>((size_t)n);
    R *rb = cv.take<R>((size_t)n);
    int *hist = cv.take<int>(h);
    int *offs = cv.take<int>(h);
    const size_t sb = scan_workspace_bytes((int)h, sizeof(int));
    void *sws = cv.take<char>(sb);
    if (!kb || !rb || !hist || !offs || !sws) return VPMC_ERR_WORKSPACE;
    const int passes = (key_bits + 7) / 8;
    const uint32_t *kin = keys;
    const R *rin = recs0;
    for (int p = 0; p < passes; ++p) {
        const bool to_out = ((passes - 1 - p) & 1) == 0;
        uint32_t *kout = to_out ? keys_sorted : kb;
        R *rout = to_out ? rec : rb;
        hipLaunchKernelGGL(k_rs_hist, dim3(rs_grid(nblk)), dim3(RS_THREADS), 0, s, kin, n, 8 * p, nblk,
                           blk, blk_base, hist);
        VPMC_CHECK_LAUNCH();
        VPMC_TRY(scan_i32(hist, offs, (int)h, false, nullptr, sws, sb, s));
        // records staged through LDS after the first pass (16-byte field
        // records keep the gather: their staging would leave one block per CU)
        if constexpr (sizeof(R) <= sizeof(SRec)) {
            if (rin) {
                hipLaunchKernelGGL((k_rs_scatter<R, true>), dim3(rs_grid(nblk)), dim3(RS_THREADS), 0,
                                   s, kin, rin, q, n, 8 * p, nblk, blk, blk_base,
                                   (const int *)offs, kout, rout);
                VPMC_CHECK_LAUNCH();
                kin = kout;
                rin = rout;
                continue;
            }
        }
            hipLaunchKernelGGL((k_rs_scatter<R, false>), dim3(rs_grid(nblk)), dim3(RS_THREADS), 0, s,
                               kin, rin, q, n, 8 * p, nblk, blk, blk_base, (const int *)offs, kout,
                               rout);
        VPMC_CHECK_LAUNCH();
        kin = kout;
        rin = rout;
    }
    hipLaunchKernelGGL(k_rs_bounds_search, dim3((unsigned)(((size_t)nkeys + 1 + 255) / 256)),
                       dim3(256), 0, s, (const uint32_t *)keys_sorted, n, nkeys, start);
    VPMC_CHECK_LAUNCH();
    return VPMC_OK;
}

int radix_sort_to_records(const uint32_t *keys, int n, int key_bits, const float2 *q,
                          const float4 *recs0, float4 *rec, uint32_t *keys_sorted, int nkeys,
                          int *start, void *ws, size_t ws_bytes, hipStream_t s)
{
    return radix_sort_impl<float4>(keys, n, key_bits, q, recs0, rec, keys_sorted, nkeys, start,
                                   nullptr, nullptr, rs_nblk(n), ws, ws_bytes, s);
}

int radix_sort_samples(const uint32_t *keys, int n, int key_bits, const SRec *recs0, SRec *rec,
                       uint32_t *keys_sorted, int nkeys, int *start, void *ws, size_t ws_bytes,
                       hipStream_t s)
{
    return radix_sort_impl<SRec>(keys, n, key_bits, nullptr, recs0, rec, keys_sorted, nkeys, start,
                                 nullptr, nullptr, rs_nblk(n), ws, ws_bytes, s);
}

int radix_sort_indices(const uint32_t *keys, int n, int key_bits, uint32_t *idx,
                       uint32_t *keys_sorted, int nkeys, int *start, void *ws, size_t ws_bytes,
                       hipStream_t s)
{
    return radix_sort_impl<uint32_t>(keys, n, key_bits, nullptr, nullptr, idx, keys_sorted, nkeys,
                                     start, nullptr, nullptr, rs_nblk(n), ws, ws_bytes, s);
}

int radix_seg_blocks(const std::vector<int> &seg_off, const std::vector<uint32_t> &seg_base,
                     std::vector<int4> *blk, std::vector<uint32_t> *blk_base)
{
    blk->clear();
    blk_base->clear();
    const int ns = (int)seg_base.size();
    if ((int)seg_off.size() != ns + 1) return VPMC_ERR_ARG;
    long long hbase = 0;
    for (int g = 0; g < ns; ++g) {
        const int b0 = seg_off[g], b1 = seg_off[g + 1];
        if (b1 < b0) return VPMC_ERR_ARG;
        const int nb = (b1 - b0 + RS_SUPER - 1) / RS_SUPER;
        for (int k = 0; k < nb; ++k) {
            const int beg = b0 + k * RS_SUPER;
            blk->push_back(make_int4(beg, std::min(RS_SUPER, b1 - beg), (int)(hbase + k), nb));
            blk_base->push_back(seg_base[g]);
        }
        hbase += (long long)RS_BINS * nb;
        if (hbase >= ((long long)1 << 31)) return VPMC_ERR_ARG;
    }
    return VPMC_OK;
}

size_t radix_seg_workspace_bytes(int n, int nblk)
{
    const size_t m = (size_t)(n > 0 ? n : 1);
    const size_t h = (size_t)RS_BINS * (nblk > 0 ? nblk : 1);
    return align256(m * 4) + align256(m * sizeof(SRec)) + 2 * align256(h * 4) +
           scan_workspace_bytes((int)h, sizeof(int)) + 1024;
}

int radix_sort_segmented(const uint32_t *keys, int n, int local_bits, const SRec *recs0,
                         SRec *rec, uint32_t *keys_sorted, int nkeys, int *start,
                         const int4 *blk, const uint32_t *blk_base, int nblk, void *ws,
                         size_t ws_bytes, hipStream_t s)
{
    if (!blk || !blk_base || nblk <= 0) return VPMC_ERR_ARG;
    return radix_sort_impl<SRec>(keys, n, local_bits, nullptr, recs0, rec, keys_sorted, nkeys,
                                 start, blk, blk_base, nblk, ws, ws_bytes, s);
}

}  // namespace vpmc
