// binned_core.hpp -- device building blocks of the binned estimator, shared
// by the single-simulation entry points (binned.hip) and the ensemble
// (ensemble.hip) so that both compute bit-identical results per simulation.
// The algorithm is described in binned.hip.
#pragma once
#include "mc_common.hpp"

namespace vpmc {

constexpr int BIN_THREADS = 256;

struct BinGeom {
    int ncell, nx, ny;
    float inv_cx, inv_cy;
    float cw, ch, rs2;        // cell size, squared search radius (field geometry)
    float rs2_1;              // squared h_w reach (deformed records scale it by sigma_max^2)
    int ntile, per_xcd;       // 2 * ncell tiles; tiles per XCD in the launch order
    int cols;                 // stencil half-width (columns)
    int ext[MAX_COL_EXT];     // rows each side for column offset ox (index ox+cols), -1: none
    uint32_t NS;              // N1 + N2
    size_t S;                 // samples
    // sample id -> (owner, sign) pair without a 64-bit division: for ids
    // below 2^31, id / NS = umulhi(id, ns_mul) >> (ns_l - 1), ns_mul =
    // ceil(2^(31 + ns_l) / NS), ns_l = ceil(log2 NS) (ns_mul = 0: NS = 1)
    uint32_t ns_mul;
    int ns_l;
};

// id / NS for 0 <= id < 2^31 (exact: the rounding error of ns_mul times id
// stays below 2^(31 + ns_l) / NS)
__device__ __forceinline__ uint32_t div_ns(const BinGeom &g, uint32_t id)
{
    return g.ns_mul == 0u ? id : (__umulhi(id, g.ns_mul) >> (g.ns_l - 1));
}

inline void set_ns_division(BinGeom *g)
{
    const uint32_t d = g->NS;
    if (d <= 1u) {
        g->ns_mul = 0u;
        g->ns_l = 0;
        return;
    }
    int l = 0;
    while ((1ull << l) < d) ++l;   // ceil(log2 d)
    g->ns_l = l;
    g->ns_mul = (uint32_t)((((unsigned __int128)1 << (31 + l)) + d - 1) / d);
}

// grid of the tiles and stencil for descriptor geometry (host)
int bin_geom(const McD &m, int N1, int N2, int n_owner, BinGeom *g);

// Tile chunks.  Branch-1 samples follow |w|, so the tiles at a vortex core
// hold tens of times the average sample count; one workgroup per tile would
// make such a tile the critical path of the launch.  A tile's samples are
// evaluated in chunks of TILE_CHUNK_BATCHES batches of THREADS * SPL samples
// (the last chunk up to twice that), one workgroup per chunk (each stages
// the tile's candidates): a sample's
// terms and their order do not depend on which workgroup evaluates it, and a
// chunk boundary falls on a batch boundary, so the tile's last (partly
// filled) wave is the same with and without chunks.
constexpr int TILE_CHUNK_BATCHES = 8;
// deformed tiles: three times the samples (the aux targets' disk samples)
// and a costlier staging of the 2 h_w stencil, so longer chunks
constexpr int TILE_CHUNK_DEF = 3;

struct TileChunks {
    const int *coff;    // ntile + 1 chunk offsets; coff[ntile] = chunks
    const int4 *ctab;   // chunk -> (tile, first sample, end, 0)
    int chunk;          // samples per chunk
};

// chunks of a tile of cnt samples: cnt / chunk (at least one for a tile with
// samples), the last taking the remainder, so that a chunk holds `chunk` to
// 2 chunk - 1 samples and a tile just past `chunk` is not split off into a
// nearly empty chunk that stages the whole stencil again
__device__ __forceinline__ int tile_nch(int cnt, int chunk)
{
    return cnt > 0 ? max(1, cnt / chunk) : 0;
}

// chunks of a launch over ntile tiles and S samples: an upper bound the host
// sizes the grid and the buffers by (each chunk past a tile's first holds at
// least `chunk` samples)
inline size_t tile_chunks_max(int ntile, size_t S, int chunk)
{
    return (size_t)ntile + S / (size_t)chunk;
}

// workgroup blockIdx.x of a grid of 8 * ceil(max / 8) -> the chunk's tile and
// sample range (the chunk count is a scalar load every workgroup of a CU
// shares, the chunk's entry one load); chunk c is taken by XCD
// c / ceil(C / 8), so each XCD runs a contiguous range of tiles (their
// stencils share its L2).  false (uniform over the workgroup): no chunk.
__device__ __forceinline__ bool chunk_range(const TileChunks &tc, int ntile, int &tile, int &beg,
                                            int &end)
{
    const int C = tc.coff[ntile];
    const int per = (C + 7) >> 3;
    const int i = (int)(blockIdx.x >> 3);
    if (i >= per) return false;
    const int c = (int)(blockIdx.x & 7) * per + i;
    if (c >= C) return false;
    const int4 e = tc.ctab[c];
    tile = e.x;
    beg = e.y;
    end = e.z;
    return true;
}

// host: chunk offsets and the chunk table from the per-tile chunk counts in
// coff[0 .. ntile] (coff[ntile] = 0 on entry; scanned in place) and the tile
// offsets; scan workspace: scan_workspace_bytes(ntile + 1, 4)
int tile_chunks_build(int *coff, int ntile, const int *start, int chunk, int4 *ctab,
                      void *scan_ws, size_t scan_bytes, hipStream_t s);

// Samples sid[0], sid[1] of one thread (count 2), or sid[0] alone (count 1):
// draws, query point, MIS displacement and tile key of each.  Writes
// key[out] (tile_base + local tile, or dead_key without a query),
// rec0[out] = (the point's offsets in its cell, out) (the record the sort
// moves) and r[out], NaN for a sample without a query: the final pass reads
// no vorticity for it, so om needs no clearing (om[out] is written by the
// evaluation for every live sample).  The two samples' loads overlap
// (gen_two); the query point, displacement and record of every sample come
// from one copy of the code (a loop that is not unrolled), so a sample's
// bits do not depend on its place in the pair.
template <int MODE, bool PER>
__device__ __forceinline__ void bin_gen_pair(const McD &d, const BinGeom &g, const size_t *sid,
                                             const size_t *out, int count, uint32_t tile_base,
                                             uint32_t dead_key, uint32_t *__restrict__ key,
                                             SRec *__restrict__ rec0, float2 *__restrict__ r)
{
    int owner[2], si[2], k[2], n[2];
#pragma unroll
    for (int q = 0; q < 2; ++q) {
        const size_t sq = sid[count > 1 ? q : 0];
        const uint32_t pair = div_ns(g, (uint32_t)sq);   // sid < 2^31 (bin_geom)
        k[q] = (int)((uint32_t)sq - pair * g.NS);
        owner[q] = (int)(pair >> 1);
        si[q] = (int)(pair & 1);
        const SetD &S = d.set[si[q]];
        n[q] = S.count_dev ? *S.count_dev : S.n;
    }
    GenTwo g2;
    gen_two<MODE, PER>(d, d.set[si[0]], d.set[si[1]], n[0], n[1], owner[0], owner[1], si[0],
                       si[1], k[0], k[1], d.owner[owner[0]], d.owner[owner[1]], g2);
#pragma unroll 1
    for (int q = 0; q < count; ++q) {
        const bool o = q != 0;
        GenState st = g2.st[0];
        if (o) st = g2.st[1];
        const int jq = o ? g2.j[1] : g2.j[0], sq = o ? si[1] : si[0], nq = o ? n[1] : n[0];
        const float2 xq = o ? g2.xj[1] : g2.xj[0];
        const float4 Fq = o ? g2.F[1] : g2.F[0];
        const size_t oq = o ? out[1] : out[0];
        float qx = 0.f, qy = 0.f, rx = 0.f, ry = 0.f;
        const bool live = nq > 0 && gen_finish<MODE, PER>(d, st, jq, xq, Fq,
                                                          d.set[sq].set_jac != nullptr, qx, qy,
                                                          rx, ry, nullptr, nullptr);
        r[oq] = live ? make_float2(rx, ry) : make_float2(__builtin_nanf(""), __builtin_nanf(""));
        if (!live) {
            rec0[oq] = SRec{0u, (uint32_t)oq};
            key[oq] = dead_key;
            continue;
        }
        const int cx = min(max((int)(qx * g.inv_cx), 0), g.nx - 1);
        const int cy = min(max((int)(qy * g.inv_cy), 0), g.ny - 1);
        rec0[oq] = SRec{srec_pack(qx, qy, cx, cy, g.cw, g.ch), (uint32_t)oq};
        key[oq] = tile_base + (uint32_t)(sq * g.ncell + cx * g.ny + cy);
    }
}

// The pair loop evaluates the spline in units where r'^2 = 3 q^2 (q = r / h_w,
// coordinates scaled by sqrt(3) / h_w): with t = max(1 - q, 0),
//   W(q) / (2 sig2) = t^3 - 4 max(t - 1/2, 0)^3 = min(t^3, 1/2 - 3 q^2 t),
// since the difference t^3 - (1/2 - 3 q^2 t) is 4 (t - 1/2)^3, negative for
// q > 1/2 and positive below, and both pieces vanish at t = 0 except the
// second, which the min drops.  Per pair: r'^2 (3), sqrt, t (clamped FMA),
// 1/2 - r'^2 t, t^3 (2), min, accumulate: 10 VALU instructions.
constexpr float SQRT3 = 1.7320508075688772f;
constexpr float INV_SQRT3 = 0.5773502691896258f;

// Tile evaluation comes in workgroup sizes THREADS = 64, 128, 256 (chosen by
// the host for the samples per tile); a pass stages up to CAP = 4 THREADS
// candidates in LDS (more candidates take several passes).
template <int THREADS> struct TileCfg {
    static constexpr int WAVES = THREADS / WAVE;
    static constexpr int CAP = 4 * THREADS;
    static constexpr int MAXCOL = THREADS < MAX_COL_EXT ? THREADS : MAX_COL_EXT;
    static constexpr int MAXP = 3 * MAXCOL;   // <= 3 runs per stencil column
};

// exclusive block-wide prefix of v (one value per thread); *total gets the sum
template <int THREADS>
__device__ __forceinline__ int block_excl_scan(int v, int *wsum, int *total)
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
    for (int w = 0; w < TileCfg<THREADS>::WAVES; ++w) {
        base += (w < wid) ? wsum[w] : 0;
        tot += wsum[w];
    }
    __syncthreads();   // wsum may be reused by the caller
    *total = tot;
    return base + x - v;
}

// The tile's stencil as runs ("pieces") of the sorted particle array, in the
// order column -cols..cols and, per column, [rows wrapped below] [rows
// wrapped above] [rows in range]; p_off is the running candidate count.
template <int THREADS, bool DEF = false> struct TileLds {
    using Cfg = TileCfg<THREADS>;
    float4 cand[Cfg::CAP];
    float4 cand_g[DEF ? Cfg::CAP : 1];   // deformed records: quadratic forms
    int p_beg[Cfg::MAXP];
    int p_off[Cfg::MAXP + 1];
    int wsum[Cfg::WAVES];
};

// Periodic image shift of stencil piece `lo` (column lo / 3 of the tile at
// column cx; kind lo % 3: rows wrapped below, rows wrapped above, rows in
// range), recomputed rather than kept in LDS so that a 256-thread tile's LDS
// stays under 20 KB (8 workgroups per CU)
__device__ __forceinline__ float2 piece_shift(const BinGeom &g, int cx, int lo, float Lx, float Ly)
{
    const int j = lo / 3, k = lo - 3 * j;
    const int colw = cx + j - g.cols;
    const float sx = colw < 0 ? -Lx : (colw >= g.nx ? Lx : 0.f);
    const float sy = k == 0 ? -Ly : (k == 1 ? Ly : 0.f);
    return make_float2(sx, sy);
}

// 2 cols + 1 <= TileCfg<THREADS>::MAXCOL (checked on the host)
template <int THREADS, bool DEF = false>
__device__ __forceinline__ int tile_pieces(const BinGeom &g, const int *__restrict__ cell_start,
                                           int cx, int cy, float Lx, float Ly,
                                           TileLds<THREADS, DEF> &L)
{
    const int ncol = 2 * g.cols + 1;
    const int j = threadIdx.x;
    int len[3] = {0, 0, 0}, pb[3] = {0, 0, 0};
    if (j < ncol) {
        const int e = g.ext[j];
        if (e >= 0) {
            int colw = cx + j - g.cols;   // image shifts: piece_shift
            if (colw < 0) colw += g.nx;
            else if (colw >= g.nx) colw -= g.nx;
            const int *csc = cell_start + (size_t)colw * g.ny;
            int r0 = cy - e, r1 = cy + e;
            if (r0 < 0) {   // rows ny+r0 .. ny-1: particle image y - Ly
                pb[0] = csc[g.ny + r0];
                len[0] = csc[g.ny] - pb[0];
                r0 = 0;
            }
            if (r1 >= g.ny) {   // rows 0 .. r1-ny: particle image y + Ly
                pb[1] = csc[0];
                len[1] = csc[r1 - g.ny + 1] - pb[1];
                r1 = g.ny - 1;
            }
            pb[2] = csc[r0];
            len[2] = csc[r1 + 1] - pb[2];
        }
    }
    int total;
    int off = block_excl_scan<THREADS>(len[0] + len[1] + len[2], L.wsum, &total);
    if (j < ncol) {
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            L.p_beg[3 * j + k] = pb[k];
            L.p_off[3 * j + k] = off;
            off += len[k];
        }
    }
    if (j == 0) L.p_off[3 * ncol] = total;
    __syncthreads();
    return total;
}

// stages the candidates with raw index in [r0, r1) (r1 - r0 <= CAP) that
// can reach the cell: cand[0 .. n) in raw order, padded with zero weights to
// a multiple of 4; returns n (padded).  Starts and ends with a barrier.
// Records hold c' = ksc (c - o) (o the cell centre, ksc = sqrt(3) / h_w):
// (-2 c'x, -2 c'y, |c'|^2, w).  DEF: the quadratic form of the deformed
// blob in the same units, split over cand and cand_g (tile_eval_def).
template <int THREADS, bool DEF = false>
__device__ __forceinline__ int tile_stage(const BinGeom &g, const float4 *__restrict__ pts,
                                          int npieces, int r0, int r1, int cx, float ox, float oy,
                                          float Lx, float Ly, float ksc, TileLds<THREADS, DEF> &L,
                                          const float4 *__restrict__ gq = nullptr)
{
    const int lane = threadIdx.x & (WAVE - 1), wid = threadIdx.x / WAVE;
    const float hx = 0.5f * g.cw, hy = 0.5f * g.ch;
    int n = 0;
    __syncthreads();   // the previous pass's readers are done with cand
    for (int i0 = r0; i0 < r1; i0 += THREADS) {
        const int i = i0 + threadIdx.x;
        bool keep = false;
        float4 rec = make_float4(0.f, 0.f, 0.f, 0.f), recg = rec;
        if (i < r1) {
            int lo = 0, hi = npieces;   // last piece with p_off <= i
            while (hi - lo > 1) {
                int mid = (lo + hi) >> 1;
                if (L.p_off[mid] <= i) lo = mid; else hi = mid;
            }
            const int src = L.p_beg[lo] + (i - L.p_off[lo]);
            const float4 c = pts[src];
            const float2 sh = piece_shift(g, cx, lo, Lx, Ly);
            const float x = (c.x + sh.x) - ox, y = (c.y + sh.y) - oy;
            const float gx = fmaxf(fabsf(x) - hx, 0.f), gy = fmaxf(fabsf(y) - hy, 0.f);
            keep = fmaf(gx, gx, gy * gy) <= g.rs2;
            if (DEF) {   // r'^2 = p^T G' p - 2 (G' c)^T p + c^T G' c, G' = ksc^2 G
                const float4 G = gq[src];
                keep = keep && fmaf(gx, gx, gy * gy) <= G.w * g.rs2_1;   // own reach
                const float k2 = ksc * ksc;
                const float g11 = k2 * G.x, g12x2 = k2 * G.y, g22 = k2 * G.z;
                const float g12 = 0.5f * g12x2;
                const float gcx = fmaf(g11, x, g12 * y), gcy = fmaf(g12, x, g22 * y);
                rec = make_float4(g11, g12x2, g22, -2.f * gcx);
                recg = make_float4(-2.f * gcy, fmaf(x, gcx, y * gcy), c.z, 0.f);
            } else {
                const float xs = ksc * x, ys = ksc * y;
                rec = make_float4(-2.f * xs, -2.f * ys, fmaf(xs, xs, ys * ys), c.z);
            }
        }
        const unsigned long long m = __ballot(keep);
        const int pre = __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                                  __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
        if (lane == 0) L.wsum[wid] = __popcll(m);
        __syncthreads();
        int base = n, tot = 0;
#pragma unroll
        for (int w = 0; w < TileCfg<THREADS>::WAVES; ++w) {
            base += (w < wid) ? L.wsum[w] : 0;
            tot += L.wsum[w];
        }
        if (keep) {
            L.cand[base + pre] = rec;
            if (DEF) L.cand_g[base + pre] = recg;
        }
        n += tot;
        __syncthreads();   // wsum is rewritten by the next round
    }
    const int npad = (n + 3) & ~3;
    if ((int)threadIdx.x < npad - n) {
        L.cand[n + threadIdx.x] = make_float4(0.f, 0.f, 0.f, 0.f);
        if (DEF) L.cand_g[n + threadIdx.x] = make_float4(0.f, 0.f, 0.f, 0.f);
    }
    __syncthreads();
    return npad;
}

// the spline term of one (sample, candidate) pair from r'^2 = 3 q^2 (see SQRT3)
__device__ __forceinline__ float spline_term(float r2)
{
    const float s = __builtin_amdgcn_sqrtf(fabsf(r2));   // |.|: rounding below 0
    const float t = __builtin_amdgcn_fmed3f(fmaf(s, -INV_SQRT3, 1.f), 0.f, 1.f);
    const float in = fmaf(-r2, t, 0.5f);
    return fminf(t * t * t, in);
}

// Summation order of a sample's terms.  Record j of a staged pass adds to
// class j mod 4 (a pass holds a multiple of four records: zero-weight
// padding), each class in record order, and the sample's sum is
// (c0 + c2) + (c1 + c3).  The same sum comes out whether one lane
// accumulates all four classes, two lanes split them ({0, 2}, {1, 3}) or
// four lanes take one each, since every split adds the same partial sums in
// the same tree; a tile's last, partly filled wave therefore spreads its
// samples over 2 or 4 lanes each (tile_eval_split) without changing a bit.
struct Acc4 {
    float c[4];
};

__device__ __forceinline__ float acc4_sum(const Acc4 &a)
{
    return (a.c[0] + a.c[2]) + (a.c[1] + a.c[3]);
}

template <int SPL, bool COUNT>
__device__ __forceinline__ void tile_eval(const float4 *__restrict__ cand, int n,
                                          const float (&px)[SPL], const float (&py)[SPL],
                                          const float (&pp)[SPL], Acc4 (&acc)[SPL],
                                          EvalCount &ec, const bool (&live)[SPL])
{
    auto one = [&](const float4 &c, int cls) {
#pragma unroll
        for (int s = 0; s < SPL; ++s) {
            const float r2 = fmaf(c.x, px[s], fmaf(c.y, py[s], c.z + pp[s]));
            acc[s].c[cls] = fmaf(c.w, spline_term(r2), acc[s].c[cls]);
            if (COUNT && live[s] && c.w > 0.f) {   // padding records have w = 0
                ec.in_support += (r2 < 3.f) ? 1u : 0u;
                ec.visited += 1u;
            }
        }
    };
    // n is a multiple of 4; eight records a trip keep the LDS address update
    // to one instruction per eight pairs
    int k = 0;
    for (; k + 8 <= n; k += 8) {
        const float4 c0 = cand[k], c1 = cand[k + 1], c2 = cand[k + 2], c3 = cand[k + 3];
        const float4 c4 = cand[k + 4], c5 = cand[k + 5], c6 = cand[k + 6], c7 = cand[k + 7];
        one(c0, 0);
        one(c1, 1);
        one(c2, 2);
        one(c3, 3);
        one(c4, 0);
        one(c5, 1);
        one(c6, 2);
        one(c7, 3);
    }
    if (k < n) {
        const float4 c0 = cand[k], c1 = cand[k + 1], c2 = cand[k + 2], c3 = cand[k + 3];
        one(c0, 0);
        one(c1, 1);
        one(c2, 2);
        one(c3, 3);
    }
}

// One sample on 2^ksh lanes (ksh = 1, 2): lane part r = lane mod 2^ksh takes
// the records of classes r (and r + 2 when ksh = 1); the partial sums are
// combined by tail_combine.
__device__ __forceinline__ void tile_eval_split(const float4 *__restrict__ cand, int n, int ksh,
                                                int r, float px, float py, float pp, Acc4 &acc)
{
    auto one = [&](const float4 &c, float &a) {
        const float r2 = fmaf(c.x, px, fmaf(c.y, py, c.z + pp));
        a = fmaf(c.w, spline_term(r2), a);
    };
    if (ksh == 1) {   // classes r, r + 2
        for (int k = 0; k < n; k += 4) {
            const float4 c0 = cand[k + r], c1 = cand[k + r + 2];
            one(c0, acc.c[0]);
            one(c1, acc.c[2]);
        }
    } else {          // class r
        int k = 0;
        for (; k + 8 <= n; k += 8) {
            const float4 c0 = cand[k + r], c1 = cand[k + 4 + r];
            one(c0, acc.c[0]);
            one(c1, acc.c[0]);
        }
        if (k < n) one(cand[k + r], acc.c[0]);
    }
}

// the sample's sum on lane part 0 of its group (every lane of the wave calls it)
__device__ __forceinline__ float tail_combine(const Acc4 &acc, int ksh)
{
    if (ksh == 0) return acc4_sum(acc);
    float e;
    if (ksh == 1) {
        e = acc.c[0] + acc.c[2];   // part r: classes r, r + 2
    } else {
        const float a = acc.c[0];   // part r: class r; pair r with r + 2
        e = a + __shfl_xor(a, 2, WAVE);
    }
    return e + __shfl_xor(e, 1, WAVE);
}

// deformed records: r'^2 = 3 |F^-1 (p - c)|^2 / h_w^2 as five FMAs on the
// sample's monomials (px^2, px py, py^2, px, py); records (G11, 2 G12, G22,
// -2 (Gc)x) and (-2 (Gc)y, c^T G c, w, 0) of G = (3 / h_w^2) F^-T F^-1,
// then the same spline
template <int SPL, bool COUNT>
__device__ __forceinline__ void tile_eval_def(const float4 *__restrict__ cand,
                                              const float4 *__restrict__ cand_g, int n,
                                              const float (&px)[SPL], const float (&py)[SPL],
                                              Acc4 (&acc)[SPL], EvalCount &ec,
                                              const bool (&live)[SPL])
{
    float xx[SPL], xy[SPL], yy[SPL];
#pragma unroll
    for (int s = 0; s < SPL; ++s) {
        xx[s] = px[s] * px[s];
        xy[s] = px[s] * py[s];
        yy[s] = py[s] * py[s];
    }
    auto one = [&](const float4 &c, const float4 &e, int cls) {
#pragma unroll
        for (int s = 0; s < SPL; ++s) {
            const float r2 = fmaf(c.x, xx[s], fmaf(c.y, xy[s], fmaf(c.z, yy[s],
                                  fmaf(c.w, px[s], fmaf(e.x, py[s], e.y)))));
            acc[s].c[cls] = fmaf(e.z, spline_term(r2), acc[s].c[cls]);
            if (COUNT && live[s] && e.z > 0.f) {
                ec.in_support += (r2 < 3.f) ? 1u : 0u;
                ec.visited += 1u;
            }
        }
    };
    for (int k = 0; k < n; k += 4) {   // n is a multiple of 4 (zero-weight padding)
        const float4 c0 = cand[k], e0 = cand_g[k], c1 = cand[k + 1], e1 = cand_g[k + 1];
        const float4 c2 = cand[k + 2], e2 = cand_g[k + 2], c3 = cand[k + 3], e3 = cand_g[k + 3];
        one(c0, e0, 0);
        one(c1, e1, 1);
        one(c2, e2, 2);
        one(c3, e3, 3);
    }
}

// The samples srec[beg, end) of local tile `tile` (sign tile / ncell) of
// the simulation described by (m, g): stage, evaluate, write omega to
// om[sample id].  Called by the whole workgroup of THREADS threads.
template <int THREADS, int SPL, bool COUNT, bool DEF = false>
__device__ __forceinline__ void bin_eval_tile(const McD &m, const BinGeom &g, int tile, int beg,
                                              int end, const SRec *__restrict__ srec,
                                              float *__restrict__ om,
                                              TileLds<THREADS, DEF> &L, EvalCount &ec)
{
    constexpr int CAP = TileCfg<THREADS>::CAP;
    const int si = tile >= g.ncell ? 1 : 0;
    const int c = tile - si * g.ncell;
    const int cx = c / g.ny, cy = c - cx * g.ny;
    const float ox = ((float)cx + 0.5f) * g.cw, oy = ((float)cy + 0.5f) * g.ch;
    const float4 *__restrict__ pts = m.set[si].field.pts;
    const float4 *__restrict__ gq = m.set[si].field.gq;
    const int total = tile_pieces<THREADS, DEF>(g, m.set[si].field.cell_start, cx, cy, m.Lx,
                                                m.Ly, L);
    const int npieces = 3 * (2 * g.cols + 1);
    const int npass = (total + CAP - 1) / CAP;
    const float ksc = SQRT3 * m.inv_hw;   // pair-loop units: r'^2 = 3 q^2
    int n = 0;
    if (npass == 1)
        n = tile_stage<THREADS, DEF>(g, pts, npieces, 0, total, cx, ox, oy, m.Lx, m.Ly, ksc, L,
                                     gq);
    const int lane = threadIdx.x & (WAVE - 1), wid = threadIdx.x / WAVE;
    for (int bb = beg; bb < end; bb += THREADS * SPL) {   // uniform trip count
        // a wave past the tile's last sample stops (no barrier follows when
        // the stencil is staged once), freeing its SIMD slot for other waves
        if (npass == 1 && bb + wid * SPL * WAVE >= end) break;
        // the tile's last wave with at most 32 samples (one sample per lane)
        // gives each sample 2 or 4 lanes (same sums, see Acc4); not in the
        // counting pass or with deformed records
        const int mw = end - (bb + wid * WAVE);
        const int ksh = (SPL == 1 && !COUNT && !DEF && mw > 0 && mw <= 32) ? (mw <= 16 ? 2 : 1) : 0;
        const int part = lane & ((1 << ksh) - 1);
        float px[SPL], py[SPL], pp[SPL];
        Acc4 acc[SPL];
        uint32_t sid[SPL];
        bool live[SPL];
#pragma unroll
        for (int s = 0; s < SPL; ++s) {
            const int idx = bb + (wid * SPL + s) * WAVE + (lane >> ksh);
            live[s] = idx < end;   // tail lanes take the cell centre, results dropped
            const SRec p = live[s] ? srec[idx] : SRec{0x80008000u, 0u};
            sid[s] = p.id;
            srec_offset(p.xy, g.cw, g.ch, px[s], py[s]);   // relative to the cell centre
            // deformed: the quadratic form carries ksc^2
            if (!DEF) {
                px[s] *= ksc;
                py[s] *= ksc;
            }
            pp[s] = fmaf(px[s], px[s], py[s] * py[s]);
#pragma unroll
            for (int c4 = 0; c4 < 4; ++c4) acc[s].c[c4] = 0.f;
        }
        for (int pass = 0; pass < npass; ++pass) {
            if (npass > 1)
                n = tile_stage<THREADS, DEF>(g, pts, npieces, pass * CAP,
                                             min(total, (pass + 1) * CAP), cx, ox, oy, m.Lx,
                                             m.Ly, ksc, L, gq);
            if constexpr (DEF)
                tile_eval_def<SPL, COUNT>(L.cand, L.cand_g, n, px, py, acc, ec, live);
            else if (ksh == 0)
                tile_eval<SPL, COUNT>(L.cand, n, px, py, pp, acc, ec, live);
            else
                tile_eval_split(L.cand, n, ksh, part, px[0], py[0], pp[0], acc[0]);
        }
#pragma unroll
        for (int s = 0; s < SPL; ++s) {
            const float sum = tail_combine(acc[s], ksh);
            if (live[s] && part == 0) om[sid[s]] = 2.f * m.sig2 * sum;
        }
    }
}

// Waves per SIMD the evaluation kernels are compiled for: 256-thread tiles
// with one sample per lane fit 64 VGPRs, eight waves per SIMD (the LDS of
// eight such tiles per CU); smaller tiles are held to seven by their LDS,
// and more samples per lane take the registers they need.
template <int THREADS, int SPL, bool DEF>
constexpr int EVAL_WAVES_PER_EU = (THREADS == 256 && SPL == 1 && !DEF) ? 8 : 1;

// host check: the stencil's columns fit a workgroup of `threads`
inline bool tile_threads_ok(const BinGeom &g, int threads)
{
    const int maxcol = threads < MAX_COL_EXT ? threads : MAX_COL_EXT;
    return 2 * g.cols + 1 <= maxcol;
}

// One wave: the estimate of (owner, sign) = pair from its samples at
// r[base..], om[base..] in sample order (deterministic butterfly sum).  With
// auxiliary targets, branch-1 samples also contribute at each target t with
// the displacement r + offset(t) (min image in the periodic consistent
// mode), and the disk samples of target t only there.
template <int MODE, bool PER>
__device__ __forceinline__ void bin_final_pair(const McD &d, const BinGeom &g, int pair,
                                               const float2 *__restrict__ r,
                                               const float *__restrict__ om)
{
    const int lane = threadIdx.x & (WAVE - 1);
    const int owner = pair >> 1, si = pair & 1;
    const SetD &S = d.set[si];
    float sx[5] = {0.f, 0.f, 0.f, 0.f, 0.f}, sy[5] = {0.f, 0.f, 0.f, 0.f, 0.f};
    const int n_s = S.count_dev ? *S.count_dev : S.n;
    if (n_s > 0) {
        const float Om = (float)(*S.total_dev);
        float cum = Om;
        if (MODE == VPMC_MODE_REFERENCE && PER) cum *= 9.f;
        const size_t base = (size_t)pair * g.NS;
        // the next sample's loads are issued before this sample's arithmetic
        float2 rn = make_float2(0.f, 0.f);
        float on = 0.f;
        if ((uint32_t)lane < g.NS) {
            rn = r[base + lane];
            on = om[base + lane];
        }
        for (uint32_t k = lane; k < g.NS; k += WAVE) {
            const float2 rr = rn;
            const float ol = on;
            if (k + WAVE < g.NS) {
                rn = r[base + k + WAVE];
                on = om[base + k + WAVE];
            }
            const float o = rr.x == rr.x ? ol : 0.f;   // NaN: no query (om not written)
            if ((int)k < d.N1) {
                mis_contrib<MODE, PER>(d, o, rr.x, rr.y, Om, cum, sx[0], sy[0]);
                if (d.aux) {
#pragma unroll
                    for (int t = 1; t < 5; ++t) {
                        const float2 of = aux_offset(t, d.aux_dx);
                        float rx = rr.x + of.x, ry = rr.y + of.y;
                        if (MODE == VPMC_MODE_CONSISTENT && PER) {
                            if (rx > 0.5f * d.Lx) rx -= d.Lx; else if (rx < -0.5f * d.Lx) rx += d.Lx;
                            if (ry > 0.5f * d.Ly) ry -= d.Ly; else if (ry < -0.5f * d.Ly) ry += d.Ly;
                        }
                        mis_contrib<MODE, PER>(d, o, rx, ry, Om, cum, sx[t], sy[t]);
                    }
                }
            } else {
                const int kb = (int)k - d.N1;
                const int t = kb < d.N2 ? 0 : 1 + (kb - d.N2) / d.N2;
#pragma unroll
                for (int tt = 0; tt < 5; ++tt)   // unrolled: the accumulators stay in registers
                    if (tt == t) mis_contrib<MODE, PER>(d, o, rr.x, rr.y, Om, cum, sx[tt], sy[tt]);
            }
        }
    }
    const int nt = d.aux ? 5 : 1;
#pragma unroll
    for (int t = 0; t < 5; ++t) {
        if (t >= nt) break;
        const float ax = wave_sum(sx[t]), ay = wave_sum(sy[t]);
        if (lane == 0) {
            const float2 v = make_float2(ax * d.inv_tot, ay * d.inv_tot);
            if (t == 0) d.vel_sign[(size_t)si * d.vel_stride + owner] = v;
            else d.vel_aux[((size_t)si * 4 + (t - 1)) * d.vel_stride + owner] = v;
        }
    }
}

}  // namespace vpmc
