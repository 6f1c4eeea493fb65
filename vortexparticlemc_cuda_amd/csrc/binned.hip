// binned.hip -- the estimator with samples binned by fine cell.
//
// A sample's query needs every particle of its sign within h_w.  Samples of
// one owner are spread over the whole box, so the direct kernel (one
// workgroup per owner) gathers a different neighbourhood per lane.  Here the
// step is split in four passes so that the 64 lanes of a wave share one
// neighbourhood:
//   gen     one thread per sample: Philox (or injected) draws -> query point,
//           MIS displacement, fine-cell ("tile") key.
//   sort    stable radix sort of the keys (sort.hip: no global atomics, whose
//           throughput would bound this step), ending in cell-sorted 16-byte
//           sample records and the tile start offsets.
//   eval    one workgroup per cell ("tile").  The workgroup stages the
//           cell's candidates in LDS once: the neighbour stencil (per column,
//           the rows within h_w of the cell rectangle: contiguous runs of
//           the sorted particle array) minus every particle farther than the
//           search radius from the rectangle, which no sample of the cell can
//           reach (exact: such terms are zero).  Records are stored relative
//           to the cell centre o as (-2c'x, -2c'y, |c'|^2, w), c' = c - o, so
//           that with p' = p - o (|p'| <= half a cell diagonal)
//             r^2 = |p'|^2 + |c'|^2 - 2 c'.p'        (3 VALU)
//           and the spline W/sig2 = 2 (t^3 - v^3), t = clamp(1 - q),
//           v = clamp(4^(1/3) (t - 1/2)), accumulates in 8 more VALU per
//           (sample, candidate) pair.  Every wave takes 64*SPL samples of the
//           cell at a time; omega goes back to the sample's id.
//   final   one wave per (owner, sign): MIS contributions of its N1+N2
//           samples in sample order, butterfly sum -> deterministic result.
// Sample ids are owner-major: sid = (owner*2 + sign)*(N1+N2) + k.
#include "mc_common.hpp"

namespace vpmc {

constexpr int BIN_THREADS = 256;

struct BinGeom {
    int ncell, nx, ny;
    float inv_cx, inv_cy;
    float cw, ch, rs2;        // cell size, squared search radius (field geometry)
    int ntile, per_xcd;       // 2 * ncell tiles; tiles per XCD in the launch order
    int cols;                 // stencil half-width (columns)
    int ext[MAX_COL_EXT];     // rows each side for column offset ox (index ox+cols), -1: none
    uint32_t NS;              // N1 + N2
    size_t S;                 // samples
};

struct BinBufs {
    uint32_t *key;            // tile key per sample (ntile: no query)
    float2 *q, *r;
    float *om;
    float4 *srec;             // cell-sorted samples: (x, y, sample id bits, 0)
    uint32_t *keys_sorted;
    int *start;               // ntile + 1 tile offsets into srec
    void *sort_ws;
    size_t sort_bytes;
};

template <int MODE, bool PER>
__global__ void __launch_bounds__(BIN_THREADS)
    k_bin_gen(McD d, BinGeom g, uint32_t *__restrict__ key, float2 *__restrict__ q,
              float2 *__restrict__ r, float *__restrict__ om)
{
    size_t sid = (size_t)blockIdx.x * BIN_THREADS + threadIdx.x;
    if (sid >= g.S) return;
    const uint32_t pair = (uint32_t)(sid / g.NS);
    const int k = (int)(sid - (size_t)pair * g.NS);
    const int owner = (int)(pair >> 1), si = (int)(pair & 1);
    const SetD &S = d.set[si];
    const int n_s = S.count_dev ? *S.count_dev : S.n;
    om[sid] = 0.f;
    float qx = 0.f, qy = 0.f, rx = 0.f, ry = 0.f;
    bool live = n_s > 0 &&
                gen_sample<MODE, PER>(d, S, n_s, owner, si, k, d.owner[owner], qx, qy, rx, ry);
    r[sid] = make_float2(rx, ry);
    q[sid] = make_float2(qx, qy);
    if (!live) {
        key[sid] = (uint32_t)g.ntile;   // sorts after every tile
        return;
    }
    int cx = min(max((int)(qx * g.inv_cx), 0), g.nx - 1);
    int cy = min(max((int)(qy * g.inv_cy), 0), g.ny - 1);
    key[sid] = (uint32_t)(si * g.ncell + cx * g.ny + cy);
}

constexpr int EVAL_CAP = 1024;                  // staged candidates per pass (16 KB)
constexpr int MAX_PIECES = 3 * MAX_COL_EXT;     // <= 3 runs per stencil column
constexpr int BIN_WAVES = BIN_THREADS / WAVE;
constexpr float CBRT4 = 1.5874010519681994f;    // 4^(1/3)

// exclusive block-wide prefix of v (one value per thread); *total gets the sum
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
    for (int w = 0; w < BIN_WAVES; ++w) {
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
struct TileLds {
    float4 cand[EVAL_CAP];
    int p_beg[MAX_PIECES];
    int p_off[MAX_PIECES + 1];
    float p_sx[MAX_PIECES], p_sy[MAX_PIECES];
    int wsum[BIN_WAVES];
    int n_stage;
};

__device__ __forceinline__ int tile_pieces(const BinGeom &g, const int *__restrict__ cell_start,
                                           int cx, int cy, float Lx, float Ly, TileLds &L)
{
    const int ncol = 2 * g.cols + 1;   // <= MAX_COL_EXT <= BIN_THREADS
    const int j = threadIdx.x;
    int len[3] = {0, 0, 0}, pb[3] = {0, 0, 0};
    float sx = 0.f, sy[3] = {-Ly, Ly, 0.f};
    if (j < ncol) {
        const int e = g.ext[j];
        if (e >= 0) {
            int colw = cx + j - g.cols;
            if (colw < 0) { colw += g.nx; sx = -Lx; }
            else if (colw >= g.nx) { colw -= g.nx; sx = Lx; }
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
    int off = block_excl_scan(len[0] + len[1] + len[2], L.wsum, &total);
    if (j < ncol) {
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            L.p_beg[3 * j + k] = pb[k];
            L.p_off[3 * j + k] = off;
            L.p_sx[3 * j + k] = sx;
            L.p_sy[3 * j + k] = sy[k];
            off += len[k];
        }
    }
    if (j == 0) L.p_off[3 * ncol] = total;
    __syncthreads();
    return total;
}

// stages the candidates with raw index in [r0, r1) (r1 - r0 <= EVAL_CAP) that
// can reach the cell: cand[0 .. n) in raw order, padded with zero weights to
// a multiple of 4; returns n (padded).  Starts and ends with a barrier.
__device__ __forceinline__ int tile_stage(const BinGeom &g, const float4 *__restrict__ pts,
                                          int npieces, int r0, int r1, float ox, float oy,
                                          TileLds &L)
{
    const int lane = threadIdx.x & (WAVE - 1), wid = threadIdx.x / WAVE;
    const float hx = 0.5f * g.cw, hy = 0.5f * g.ch;
    int n = 0;
    __syncthreads();   // the previous pass's readers are done with cand
    for (int i0 = r0; i0 < r1; i0 += BIN_THREADS) {
        const int i = i0 + threadIdx.x;
        bool keep = false;
        float4 rec = make_float4(0.f, 0.f, 0.f, 0.f);
        if (i < r1) {
            int lo = 0, hi = npieces;   // last piece with p_off <= i
            while (hi - lo > 1) {
                int mid = (lo + hi) >> 1;
                if (L.p_off[mid] <= i) lo = mid; else hi = mid;
            }
            const float4 c = pts[L.p_beg[lo] + (i - L.p_off[lo])];
            const float x = (c.x + L.p_sx[lo]) - ox, y = (c.y + L.p_sy[lo]) - oy;
            const float gx = fmaxf(fabsf(x) - hx, 0.f), gy = fmaxf(fabsf(y) - hy, 0.f);
            keep = fmaf(gx, gx, gy * gy) <= g.rs2;
            rec = make_float4(-2.f * x, -2.f * y, fmaf(x, x, y * y), c.z);
        }
        const unsigned long long m = __ballot(keep);
        const int pre = __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                                  __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
        if (lane == 0) L.wsum[wid] = __popcll(m);
        __syncthreads();
        int base = n, tot = 0;
#pragma unroll
        for (int w = 0; w < BIN_WAVES; ++w) {
            base += (w < wid) ? L.wsum[w] : 0;
            tot += L.wsum[w];
        }
        if (keep) L.cand[base + pre] = rec;
        n += tot;
        __syncthreads();   // wsum is rewritten by the next round
    }
    const int npad = (n + 3) & ~3;
    if ((int)threadIdx.x < npad - n) L.cand[n + threadIdx.x] = make_float4(0.f, 0.f, 0.f, 0.f);
    __syncthreads();
    return npad;
}

template <int SPL, bool COUNT>
__device__ __forceinline__ void tile_eval(const float4 *__restrict__ cand, int n,
                                          const float (&px)[SPL], const float (&py)[SPL],
                                          const float (&pp)[SPL], float inv_hw,
                                          float (&acc)[SPL], EvalCount &ec,
                                          const bool (&live)[SPL])
{
    auto one = [&](const float4 &c) {
#pragma unroll
        for (int s = 0; s < SPL; ++s) {
            float r2 = fmaf(c.x, px[s], fmaf(c.y, py[s], c.z + pp[s]));
            float q = __builtin_amdgcn_sqrtf(fabsf(r2));   // |.|: rounding below 0
            float t = __builtin_amdgcn_fmed3f(fmaf(-inv_hw, q, 1.f), 0.f, 1.f);
            float v = __builtin_amdgcn_fmed3f(fmaf(t, CBRT4, -0.5f * CBRT4), 0.f, 1.f);
            float t2 = t * t, v3 = v * v * v;
            acc[s] = fmaf(c.w, fmaf(t2, t, -v3), acc[s]);
            if (COUNT && live[s] && c.w > 0.f) {   // padding records have w = 0
                ec.in_support += (t > 0.f) ? 1u : 0u;
                ec.visited += 1u;
            }
        }
    };
    for (int k = 0; k < n; k += 4) {   // n is a multiple of 4 (zero-weight padding)
        const float4 c0 = cand[k], c1 = cand[k + 1], c2 = cand[k + 2], c3 = cand[k + 3];
        one(c0);
        one(c1);
        one(c2);
        one(c3);
    }
}

template <int SPL, bool COUNT>
__global__ void __launch_bounds__(BIN_THREADS)
    k_bin_eval(BinGeom g, float inv_hw, float sig2, float Lx, float Ly,
               const float4 *__restrict__ pts0, const int *__restrict__ cs0,
               const float4 *__restrict__ pts1, const int *__restrict__ cs1,
               const int *__restrict__ start, const float4 *__restrict__ srec,
               float *__restrict__ om, unsigned long long *__restrict__ eval_count)
{
    // workgroups are dealt round-robin to the 8 XCDs: give each XCD a
    // contiguous range of tiles so that neighbouring cells share its L2
    const int tile = (int)(blockIdx.x & 7) * g.per_xcd + (int)(blockIdx.x >> 3);
    if (tile >= g.ntile) return;
    const int beg = start[tile], end = start[tile + 1];
    if (beg >= end) return;   // uniform over the workgroup
    __shared__ TileLds L;
    EvalCount ec;
    const int si = tile >= g.ncell ? 1 : 0;
    const int c = tile - si * g.ncell;
    const int cx = c / g.ny, cy = c - cx * g.ny;
    const float ox = ((float)cx + 0.5f) * g.cw, oy = ((float)cy + 0.5f) * g.ch;
    const float4 *__restrict__ pts = si ? pts1 : pts0;
    const int total = tile_pieces(g, si ? cs1 : cs0, cx, cy, Lx, Ly, L);
    const int npieces = 3 * (2 * g.cols + 1);
    const int npass = (total + EVAL_CAP - 1) / EVAL_CAP;
    int n = 0;
    if (npass == 1) n = tile_stage(g, pts, npieces, 0, total, ox, oy, L);
    const int lane = threadIdx.x & (WAVE - 1), wid = threadIdx.x / WAVE;
    for (int bb = beg; bb < end; bb += BIN_THREADS * SPL) {   // uniform trip count
        float px[SPL], py[SPL], pp[SPL], acc[SPL];
        uint32_t sid[SPL];
        bool live[SPL];
#pragma unroll
        for (int s = 0; s < SPL; ++s) {
            const int idx = bb + (wid * SPL + s) * WAVE + lane;
            live[s] = idx < end;   // tail lanes take the cell centre, results dropped
            const float4 p = live[s] ? srec[idx] : make_float4(ox, oy, 0.f, 0.f);
            sid[s] = __float_as_uint(p.z);
            px[s] = p.x - ox;
            py[s] = p.y - oy;
            pp[s] = fmaf(px[s], px[s], py[s] * py[s]);
            acc[s] = 0.f;
        }
        for (int pass = 0; pass < npass; ++pass) {
            if (npass > 1)
                n = tile_stage(g, pts, npieces, pass * EVAL_CAP,
                               min(total, (pass + 1) * EVAL_CAP), ox, oy, L);
            tile_eval<SPL, COUNT>(L.cand, n, px, py, pp, inv_hw, acc, ec, live);
        }
#pragma unroll
        for (int s = 0; s < SPL; ++s)
            if (live[s]) om[sid[s]] = 2.f * sig2 * acc[s];
    }
    if (COUNT) {
        unsigned a = wave_sum(ec.in_support), b = wave_sum(ec.visited);
        if (lane == 0) {
            atomicAdd(eval_count, (unsigned long long)a);
            atomicAdd(eval_count + 1, (unsigned long long)b);
        }
    }
}

template <int MODE, bool PER>
__global__ void __launch_bounds__(BIN_THREADS)
    k_bin_final(McD d, BinGeom g, const float2 *__restrict__ r, const float *__restrict__ om)
{
    const int pair = blockIdx.x * (BIN_THREADS / WAVE) + threadIdx.x / WAVE;
    const int lane = threadIdx.x & (WAVE - 1);
    if (pair >= 2 * d.n_owner) return;
    const int owner = pair >> 1, si = pair & 1;
    const SetD &S = d.set[si];
    float sx = 0.f, sy = 0.f;
    const int n_s = S.count_dev ? *S.count_dev : S.n;
    if (n_s > 0) {
        const float Om = (float)(*S.total_dev);
        float cum = Om;
        if (MODE == VPMC_MODE_REFERENCE && PER) cum *= 9.f;
        const size_t base = (size_t)pair * g.NS;
        for (uint32_t k = lane; k < g.NS; k += WAVE) {
            float2 rr = r[base + k];
            mis_contrib<MODE, PER>(d, om[base + k], rr.x, rr.y, Om, cum, sx, sy);
        }
    }
    sx = wave_sum(sx);
    sy = wave_sum(sy);
    if (lane == 0) d.vel_sign[(size_t)si * d.n_owner + owner] = make_float2(sx * d.inv_tot, sy * d.inv_tot);
}

static int bin_geom(const McD &m, const vpmc_mc_desc *in, BinGeom *g)
{
    const FieldGeom &f = m.set[0].field.g;
    const FieldGeom &f1 = m.set[1].field.g;
    if (f.nx != f1.nx || f.ny != f1.ny) return VPMC_ERR_ARG;
    g->nx = f.nx;
    g->ny = f.ny;
    g->ncell = f.nx * f.ny;
    g->inv_cx = f.inv_cx;
    g->inv_cy = f.inv_cy;
    g->cw = f.cw;
    g->ch = f.ch;
    g->rs2 = f.rs2;
    g->ntile = 2 * g->ncell;
    g->per_xcd = (g->ntile + 7) / 8;
    g->cols = 1 + (int)floorf(f.rs / f.cw);
    if (2 * g->cols + 1 > MAX_COL_EXT || 2 * g->cols + 1 > f.nx) return VPMC_ERR_GRID;
    int emax = 0;
    for (int ox = -g->cols; ox <= g->cols; ++ox) {
        int a = ox < 0 ? -ox : ox;
        float gap = (a > 0 ? (float)(a - 1) : 0.f) * f.cw;
        float rem = f.rs2 - gap * gap;
        int e = rem < 0.f ? -1 : 1 + (int)floorf(sqrtf(rem) / f.ch);
        g->ext[ox + g->cols] = e;
        if (e > emax) emax = e;
    }
    if (2 * emax + 1 > f.ny) return VPMC_ERR_GRID;
    g->NS = (uint32_t)(in->N1 + in->N2);
    g->S = (size_t)in->n_owner * 2 * g->NS;
    if (g->S >= (size_t)1 << 31) return VPMC_ERR_ARG;   // int32 sample slots
    return VPMC_OK;
}

static size_t bin_ws_bytes(const BinGeom &g)
{
    size_t S = g.S > 0 ? g.S : 1, T = (size_t)g.ntile;
    return align256(S * 4) + 2 * align256(S * 8) + align256(S * 4) + align256(S * 16) +
           align256(S * 4) + align256((T + 1) * 4) + radix_workspace_bytes((int)S) + 4096;
}

static int carve_bin(const BinGeom &g, void *ws, size_t bytes, BinBufs *B)
{
    Carve cv(ws, bytes);
    size_t S = g.S > 0 ? g.S : 1, T = (size_t)g.ntile;
    B->key = cv.take<uint32_t>(S);
    B->q = cv.take<float2>(S);
    B->r = cv.take<float2>(S);
    B->om = cv.take<float>(S);
    B->srec = cv.take<float4>(S);
    B->keys_sorted = cv.take<uint32_t>(S);
    B->start = cv.take<int>(T + 1);
    B->sort_bytes = radix_workspace_bytes((int)S);
    B->sort_ws = cv.take<char>(B->sort_bytes);
    if (!B->key || !B->q || !B->r || !B->om || !B->srec || !B->keys_sorted || !B->start ||
        !B->sort_ws)
        return VPMC_ERR_WORKSPACE;
    return VPMC_OK;
}

template <int MODE, bool PER>
static void launch_gen_final(const McD &m, const BinGeom &g, const BinBufs &B, bool final,
                             hipStream_t s)
{
    if (!final) {
        unsigned gb = (unsigned)((g.S + BIN_THREADS - 1) / BIN_THREADS);
        hipLaunchKernelGGL((k_bin_gen<MODE, PER>), dim3(gb), dim3(BIN_THREADS), 0, s, m, g, B.key,
                           B.q, B.r, B.om);
    } else {
        unsigned gb = (unsigned)((2 * (size_t)m.n_owner + 3) / 4);
        hipLaunchKernelGGL((k_bin_final<MODE, PER>), dim3(gb), dim3(BIN_THREADS), 0, s, m, g,
                           (const float2 *)B.r, (const float *)B.om);
    }
}

static void launch_mode(int mode, int per, const McD &m, const BinGeom &g, const BinBufs &B,
                        bool final, hipStream_t s)
{
    if (mode == VPMC_MODE_REFERENCE) {
        if (per) launch_gen_final<VPMC_MODE_REFERENCE, true>(m, g, B, final, s);
        else launch_gen_final<VPMC_MODE_REFERENCE, false>(m, g, B, final, s);
    } else {
        if (per) launch_gen_final<VPMC_MODE_CONSISTENT, true>(m, g, B, final, s);
        else launch_gen_final<VPMC_MODE_CONSISTENT, false>(m, g, B, final, s);
    }
}

int mc_binned_workspace(const vpmc_mc_desc *d, size_t *bytes)
{
    McD m;
    VPMC_TRY(make_mcd(d, &m));
    BinGeom g;
    VPMC_TRY(bin_geom(m, d, &g));
    *bytes = bin_ws_bytes(g);
    return VPMC_OK;
}

int mc_binned_impl(const vpmc_mc_desc *d, void *ws, size_t ws_bytes, hipStream_t s)
{
    McD m;
    VPMC_TRY(make_mcd(d, &m));
    if (!d->periodic) return VPMC_ERR_ARG;   // the cell stencil wraps periodically
    if (m.n_owner == 0) return VPMC_OK;
    BinGeom g;
    VPMC_TRY(bin_geom(m, d, &g));
    BinBufs B;
    VPMC_TRY(carve_bin(g, ws, ws_bytes, &B));
    launch_mode(d->mode, d->periodic, m, g, B, false, s);
    VPMC_CHECK_LAUNCH();
    int key_bits = 1;
    while (key_bits < 32 && ((uint32_t)g.ntile >> key_bits) != 0u) ++key_bits;
    VPMC_TRY(radix_sort_to_records(B.key, (int)g.S, key_bits, (const float2 *)B.q, B.srec,
                                   B.keys_sorted, g.ntile, B.start, B.sort_ws, B.sort_bytes, s));
    const float4 *p0 = m.set[0].field.pts, *p1 = m.set[1].field.pts;
    const int *c0 = m.set[0].field.cell_start, *c1 = m.set[1].field.cell_start;
    VPMC_TRY(record_event(d->eval_events[0], s));
    const dim3 eval_grid((unsigned)(8 * g.per_xcd));
    if (d->eval_count)
        hipLaunchKernelGGL((k_bin_eval<2, true>), eval_grid, dim3(BIN_THREADS), 0, s, g, m.inv_hw,
                           m.sig2, m.Lx, m.Ly, p0, c0, p1, c1, (const int *)B.start,
                           (const float4 *)B.srec, B.om, d->eval_count);
    else
        hipLaunchKernelGGL((k_bin_eval<2, false>), eval_grid, dim3(BIN_THREADS), 0, s, g, m.inv_hw,
                           m.sig2, m.Lx, m.Ly, p0, c0, p1, c1, (const int *)B.start,
                           (const float4 *)B.srec, B.om, (unsigned long long *)nullptr);
    VPMC_CHECK_LAUNCH();
    VPMC_TRY(record_event(d->eval_events[1], s));
    launch_mode(d->mode, d->periodic, m, g, B, true, s);
    VPMC_CHECK_LAUNCH();
    return VPMC_OK;
}

}  // namespace vpmc

using namespace vpmc;

extern "C" {

size_t vpmc_mc_binned_workspace_bytes(const vpmc_mc_desc *d)
{
    size_t b = 0;
    return mc_binned_workspace(d, &b) == VPMC_OK ? b : 0;
}

int vpmc_mc_estimate_binned(const vpmc_mc_desc *d, void *workspace, size_t workspace_bytes,
                            void *stream)
{
    return mc_binned_impl(d, workspace, workspace_bytes, as_stream(stream));
}

}  // extern "C"
