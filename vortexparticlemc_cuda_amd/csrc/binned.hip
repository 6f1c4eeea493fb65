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
// The four passes run once per owner chunk (at most chunk_samples sample
// slots), so the workspace is that of one chunk.  A sample's omega depends
// only on its point and the field, its draws on the global owner index, and
// each owner's reduction on its own samples, so the chunk size does not
// change any result.
#include "binned_core.hpp"

namespace vpmc {

struct BinBufs {
    uint32_t *key;            // tile key per sample (ntile: no query)
    SRec *rec0;               // samples in sample order: (query point, sample id)
    float2 *r;
    float *om;
    SRec *srec;               // cell-sorted samples: (offsets in the cell, sample id)
    uint32_t *keys_sorted;
    int *start;               // ntile + 1 tile offsets into srec
    void *sort_ws;
    size_t sort_bytes;
    int *coff;                // tile chunks: ntile + 1 offsets
    int4 *ctab;               // chunk -> (tile, first sample, end, 0)
    void *scan_ws;
    size_t scan_bytes;
};

// samples per lane and workgroup size of the evaluation launch (spl 0: 2,
// threads 0: 256, as launch_eval picks them)
inline int eval_spl(int spl) { return spl == 1 ? 1 : (spl == 4 ? 4 : 2); }
inline int eval_threads_of(int threads) { return threads == 0 ? 256 : threads; }
// the smallest chunk any launch shape uses (64-thread tiles, 1 sample per
// lane): sizes the chunk table for every shape
constexpr int TILE_CHUNK_MIN = TILE_CHUNK_BATCHES * WAVE;

// chunks of tile t (tile_nch); nch[ntile] = 0 (the scan's total slot)
__global__ void k_tile_nch(int ntile, const int *__restrict__ start, int chunk,
                           int *__restrict__ nch)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t > ntile) return;
    nch[t] = t < ntile ? tile_nch(start[t + 1] - start[t], chunk) : 0;
}

// the chunk table: tile t writes its chunks' entries (most tiles have one);
// chunks of `chunk` samples, the last one to the tile's end
__global__ void k_chunk_fill(int ntile, const int *__restrict__ coff,
                             const int *__restrict__ start, int chunk, int4 *__restrict__ ctab)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ntile) return;
    const int b = start[t], e = start[t + 1], c0 = coff[t], c1 = coff[t + 1];
    for (int c = c0, p = b; c < c1; ++c, p += chunk)
        ctab[c] = make_int4(t, p, c + 1 < c1 ? p + chunk : e, 0);
}

int tile_chunks_build(int *coff, int ntile, const int *start, int chunk, int4 *ctab,
                      void *scan_ws, size_t scan_bytes, hipStream_t s)
{
    VPMC_TRY(scan_i32(coff, coff, ntile + 1, false, nullptr, scan_ws, scan_bytes, s));
    hipLaunchKernelGGL(k_chunk_fill, dim3((unsigned)((ntile + 255) / 256)), dim3(256), 0, s, ntile,
                       (const int *)coff, start, chunk, ctab);
    VPMC_CHECK_LAUNCH();
    return VPMC_OK;
}

template <int MODE, bool PER>
__global__ void __launch_bounds__(BIN_THREADS) __attribute__((amdgpu_waves_per_eu(8)))
    k_bin_gen(McD d, BinGeom g, uint32_t *__restrict__ key, SRec *__restrict__ rec0,
              float2 *__restrict__ r)
{
    // two samples a lane, 64 apart (each wave's stores stay 64 consecutive
    // samples wide), their loads overlapped (bin_gen_pair)
    const size_t a = (size_t)blockIdx.x * (2 * BIN_THREADS) + (threadIdx.x / WAVE) * (2 * WAVE) +
                     (threadIdx.x & (WAVE - 1));
    if (a >= g.S) return;
    const size_t ids[2] = {a, a + WAVE};
    bin_gen_pair<MODE, PER>(d, g, ids, ids, ids[1] < g.S ? 2 : 1, 0u, (uint32_t)g.ntile, key,
                            rec0, r);
}

template <int THREADS, int SPL, bool COUNT, bool DEF = false>
__global__ void __launch_bounds__(THREADS) __attribute__((amdgpu_waves_per_eu(EVAL_WAVES_PER_EU<THREADS, SPL, DEF>)))
    k_bin_eval(McD m, BinGeom g, TileChunks tc, const SRec *__restrict__ srec,
               float *__restrict__ om,
               unsigned long long *__restrict__ eval_count)
{
    // one workgroup per tile chunk; workgroups are dealt round-robin to the
    // 8 XCDs, and each XCD takes a contiguous range of chunks so that
    // neighbouring cells share its L2
    int tile, beg, end;
    if (!chunk_range(tc, g.ntile, tile, beg, end)) return;   // uniform
    __shared__ TileLds<THREADS, DEF> L;
    EvalCount ec;
    bin_eval_tile<THREADS, SPL, COUNT, DEF>(m, g, tile, beg, end, srec, om, L, ec);
    if (COUNT) {
        unsigned a = wave_sum(ec.in_support), b = wave_sum(ec.visited);
        if ((threadIdx.x & (WAVE - 1)) == 0) {
            atomicAdd(eval_count, (unsigned long long)a);
            atomicAdd(eval_count + 1, (unsigned long long)b);
        }
    }
}

template <int MODE, bool PER>
__global__ void __launch_bounds__(BIN_THREADS)
    k_bin_final(McD d, BinGeom g, const float2 *__restrict__ r, const float *__restrict__ om)
{
    // wave-uniform, stated so that the set's fields load through the scalar unit
    const int pair = __builtin_amdgcn_readfirstlane(blockIdx.x * (BIN_THREADS / WAVE) +
                                                    threadIdx.x / WAVE);
    if (pair >= 2 * d.n_owner) return;
    bin_final_pair<MODE, PER>(d, g, pair, r, om);
}

// eval launch for runtime workgroup-size and samples-per-lane choices
template <int THREADS, bool COUNT, bool DEF>
static void launch_eval_t(int spl, const McD &m, const BinGeom &g, const TileChunks &tc,
                          const SRec *srec, float *om, unsigned long long *cnt, hipStream_t s)
{
    const size_t cmax = tile_chunks_max(g.ntile, g.S, tc.chunk);
    const dim3 grid((unsigned)(8 * ((cmax + 7) / 8))), block(THREADS);
    if (spl == 1)
        hipLaunchKernelGGL((k_bin_eval<THREADS, 1, COUNT, DEF>), grid, block, 0, s, m, g, tc,
                           srec, om, cnt);
    else if (spl == 4)
        hipLaunchKernelGGL((k_bin_eval<THREADS, 4, COUNT, DEF>), grid, block, 0, s, m, g, tc,
                           srec, om, cnt);
    else
        hipLaunchKernelGGL((k_bin_eval<THREADS, 2, COUNT, DEF>), grid, block, 0, s, m, g, tc,
                           srec, om, cnt);
}

template <bool COUNT, bool DEF>
static void launch_eval_d(int threads, int spl, const McD &m, const BinGeom &g,
                          const TileChunks &tc, const SRec *srec, float *om,
                          unsigned long long *cnt, hipStream_t s)
{
    if (threads == 64) launch_eval_t<64, COUNT, DEF>(spl, m, g, tc, srec, om, cnt, s);
    else if (threads == 128) launch_eval_t<128, COUNT, DEF>(spl, m, g, tc, srec, om, cnt, s);
    else launch_eval_t<256, COUNT, DEF>(spl, m, g, tc, srec, om, cnt, s);
}

// the tiles' chunks (TILE_CHUNK_BATCHES batches of the launch shape each,
// TILE_CHUNK_DEF times that for deformed tiles) from the sort's tile offsets
static int build_chunks(int threads, int spl, bool def, const BinGeom &g, const BinBufs &B,
                        TileChunks *tc, hipStream_t s)
{
    threads = eval_threads_of(threads);
    if (threads != 64 && threads != 128 && threads != 256) return VPMC_ERR_ARG;
    *tc = TileChunks{B.coff, B.ctab,
                     TILE_CHUNK_BATCHES * threads * eval_spl(spl) * (def ? TILE_CHUNK_DEF : 1)};
    hipLaunchKernelGGL(k_tile_nch, dim3((unsigned)((g.ntile + 1 + 255) / 256)), dim3(256), 0, s,
                       g.ntile, (const int *)B.start, tc->chunk, B.coff);
    return tile_chunks_build(B.coff, g.ntile, B.start, tc->chunk, B.ctab, B.scan_ws,
                             B.scan_bytes, s);
}

// deformed sets (field.gq != NULL) take the quadratic-form records
template <bool COUNT>
static int launch_eval(int threads, int spl, const McD &m, const BinGeom &g, const BinBufs &B,
                       const TileChunks &tc, unsigned long long *cnt, hipStream_t s)
{
    threads = eval_threads_of(threads);
    if (threads != 64 && threads != 128 && threads != 256) return VPMC_ERR_ARG;
    if (!tile_threads_ok(g, threads)) return VPMC_ERR_GRID;
    if (m.set[0].field.gq)
        launch_eval_d<COUNT, true>(threads, spl, m, g, tc, B.srec, B.om, cnt, s);
    else launch_eval_d<COUNT, false>(threads, spl, m, g, tc, B.srec, B.om, cnt, s);
    return VPMC_OK;
}

int bin_geom(const McD &m, int N1, int N2, int n_owner, BinGeom *g)
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
    g->rs2_1 = f.rs1 * f.rs1;
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
    g->NS = (uint32_t)samples_per_pair(N1, N2, m.aux);
    g->S = (size_t)n_owner * 2 * g->NS;
    if (g->S >= (size_t)1 << 31) return VPMC_ERR_ARG;   // int32 sample slots
    set_ns_division(g);
    return VPMC_OK;
}

// Owners per chunk of the binned estimate: balanced chunks of at most
// chunk_samples (0: 2^30) sample slots, each below the sort's int32 limit.
constexpr uint64_t BIN_CHUNK_DEFAULT = (uint64_t)1 << 30;

static int bin_chunk_owners(const vpmc_mc_desc *d, int *per_chunk)
{
    const uint64_t per_owner = 2 * (uint64_t)samples_per_pair(d->N1, d->N2, d->aux);
    uint64_t cap = d->chunk_samples ? d->chunk_samples : BIN_CHUNK_DEFAULT;
    if (cap > ((uint64_t)1 << 31) - 1) cap = ((uint64_t)1 << 31) - 1;
    if (per_owner > cap) return VPMC_ERR_ARG;   // one owner's samples exceed a chunk
    const int n = d->n_owner;
    if (n <= 0) {
        *per_chunk = 1;
        return VPMC_OK;
    }
    const uint64_t most = cap / per_owner;
    const uint64_t nchunk = ((uint64_t)n + most - 1) / most;
    *per_chunk = (int)(((uint64_t)n + nchunk - 1) / nchunk);
    return VPMC_OK;
}

static size_t bin_ws_bytes(const BinGeom &g)
{
    size_t S = g.S > 0 ? g.S : 1, T = (size_t)g.ntile;
    return align256(S * 4) + align256(S * sizeof(SRec)) + align256(S * 8) + align256(S * 4) +
           align256(S * sizeof(SRec)) +
           align256(S * 4) + align256((T + 1) * 4) + radix_workspace_bytes((int)S, sizeof(SRec)) +
           align256((T + 1) * 4) + align256(tile_chunks_max(g.ntile, S, TILE_CHUNK_MIN) * 16) +
           scan_workspace_bytes(g.ntile + 1, 4) + 4096;
}

static int carve_bin(const BinGeom &g, void *ws, size_t bytes, BinBufs *B)
{
    Carve cv(ws, bytes);
    size_t S = g.S > 0 ? g.S : 1, T = (size_t)g.ntile;
    B->key = cv.take<uint32_t>(S);
    B->rec0 = cv.take<SRec>(S);
    B->r = cv.take<float2>(S);
    B->om = cv.take<float>(S);
    B->srec = cv.take<SRec>(S);
    B->keys_sorted = cv.take<uint32_t>(S);
    B->start = cv.take<int>(T + 1);
    B->sort_bytes = radix_workspace_bytes((int)S, sizeof(SRec));
    B->sort_ws = cv.take<char>(B->sort_bytes);
    B->coff = cv.take<int>(T + 1);
    B->ctab = cv.take<int4>(tile_chunks_max(g.ntile, S, TILE_CHUNK_MIN));
    B->scan_bytes = scan_workspace_bytes(g.ntile + 1, 4);
    B->scan_ws = cv.take<char>(B->scan_bytes);
    if (!B->key || !B->rec0 || !B->r || !B->om || !B->srec || !B->keys_sorted || !B->start ||
        !B->sort_ws || !B->coff || !B->ctab || !B->scan_ws)
        return VPMC_ERR_WORKSPACE;
    return VPMC_OK;
}

template <int MODE, bool PER>
static void launch_gen_final(const McD &m, const BinGeom &g, const BinBufs &B, bool final,
                             hipStream_t s)
{
    if (!final) {
        unsigned gb = (unsigned)((g.S + 2 * BIN_THREADS - 1) / (2 * BIN_THREADS));
        hipLaunchKernelGGL((k_bin_gen<MODE, PER>), dim3(gb), dim3(BIN_THREADS), 0, s, m, g, B.key,
                           B.rec0, B.r);
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
    int per_chunk;
    VPMC_TRY(bin_chunk_owners(d, &per_chunk));
    BinGeom g;
    VPMC_TRY(bin_geom(m, d->N1, d->N2, d->n_owner < per_chunk ? d->n_owner : per_chunk, &g));
    *bytes = bin_ws_bytes(g);
    return VPMC_OK;
}

// Owners [o0, o0 + cnt) of the call: the chunk's descriptor reads owner
// o0 + i as its owner i, keys its draws by the global owner index and
// writes rows o0 + i of vel_sign / vel_aux (stride n_owner of the call).
static McD chunk_mcd(const McD &m, int o0, int cnt)
{
    McD c = m;
    c.owner = m.owner + o0;
    c.pbase = m.pbase + o0;
    c.n_owner = cnt;
    c.vel_sign = m.vel_sign + o0;
    c.vel_aux = m.vel_aux ? m.vel_aux + o0 : nullptr;
    return c;
}

int mc_binned_impl(const vpmc_mc_desc *d, void *ws, size_t ws_bytes, hipStream_t s)
{
    McD m;
    VPMC_TRY(make_mcd(d, &m));
    if (!d->periodic) return VPMC_ERR_ARG;   // the cell stencil wraps periodically
    if (m.n_owner == 0) return VPMC_OK;
    int per_chunk;
    VPMC_TRY(bin_chunk_owners(d, &per_chunk));
    BinGeom g0;
    VPMC_TRY(bin_geom(m, d->N1, d->N2, per_chunk < m.n_owner ? per_chunk : m.n_owner, &g0));
    BinBufs B;
    VPMC_TRY(carve_bin(g0, ws, ws_bytes, &B));
    for (int o0 = 0; o0 < m.n_owner; o0 += per_chunk) {
        const int cnt = per_chunk < m.n_owner - o0 ? per_chunk : m.n_owner - o0;
        const McD mc = chunk_mcd(m, o0, cnt);
        BinGeom g = g0;
        g.S = (size_t)cnt * 2 * g.NS;
        launch_mode(d->mode, d->periodic, mc, g, B, false, s);
        VPMC_CHECK_LAUNCH();
        VPMC_TRY(radix_sort_samples(B.key, (int)g.S, key_bits_for((uint32_t)g.ntile),
                                    B.rec0, B.srec, B.keys_sorted, g.ntile, B.start,
                                    B.sort_ws, B.sort_bytes, s));
        TileChunks tc;
        VPMC_TRY(build_chunks(d->eval_threads, d->spl, mc.set[0].field.gq != nullptr, g, B, &tc,
                              s));
        if (o0 == 0) VPMC_TRY(record_event(d->eval_events[0], s));
        if (d->eval_count)
            VPMC_TRY(launch_eval<true>(d->eval_threads, d->spl, mc, g, B, tc, d->eval_count, s));
        else
            VPMC_TRY(launch_eval<false>(d->eval_threads, d->spl, mc, g, B, tc, nullptr, s));
        VPMC_CHECK_LAUNCH();
        if (o0 + cnt >= m.n_owner) VPMC_TRY(record_event(d->eval_events[1], s));
        launch_mode(d->mode, d->periodic, mc, g, B, true, s);
        VPMC_CHECK_LAUNCH();
    }
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
