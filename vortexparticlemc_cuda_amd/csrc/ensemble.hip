// ensemble.hip -- many independent periodic simulations advanced together.
//
// The reference runs one Simulation per process.  Here an ensemble of
// simulations of uneven sizes shares every launch of a step:
//   fields   one stable radix sort of all (simulation, sign, subset entry)
//            records by global cell key -> every simulation's two fields;
//   gen      one thread per sample of every simulation (per-simulation
//            sample ranges padded to 256, so a block never straddles two);
//   sort     one stable radix sort of all sample tile keys;
//   eval     one workgroup per tile of every simulation;
//   final    one wave per (owner, sign) of every simulation;
//   advect   one thread per particle; then the step counter advances.
// Per-simulation parameters live in device tables (McD, BinGeom) in the
// caller's workspace; a block finds its simulation by a binary search of an
// offset table with block-uniform operands.  The device code of every stage
// is the single-simulation code (binned_core.hpp, advect_one), and every
// order it depends on (field order, sample order in a tile, summation order)
// is the single-simulation order, so each simulation's trajectory equals
// vpmc_step's on that simulation alone (sim = its index) bit for bit.
// The step counter is device state, so a captured step replays correctly.
#include <vector>

#include "binned_core.hpp"

namespace vpmc {

// largest k with off[k] <= x, for off[0] = 0 <= x < off[n] (off non-decreasing)
template <class T>
__device__ __forceinline__ int seg_of(const T *__restrict__ off, int n, T x)
{
    int lo = 0, hi = n;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (off[mid] <= x) lo = mid; else hi = mid;
    }
    return lo;
}

struct EnsTabs {
    McD *sim;          // n_sim
    BinGeom *geom;     // n_sim
    int *poff;         // n_sim + 1 particle offsets
    int *soff;         // n_sim + 1 sample offsets (padded)
    int *toff;         // n_sim + 1 tile offsets
    int *eoff;         // 2 n_sim + 1 field entry offsets, (sim, sign)-major
    int *coff;         // 2 n_sim + 1 field cell offsets
};

// ---- kernels ----------------------------------------------------------------

// entry -> global particle index (-1 beyond the subset's count); init only
__global__ void k_ens_entries(EnsTabs t, int n_sim, int E, const int *__restrict__ sub_pool,
                              const int *__restrict__ counts, int *__restrict__ ent_p,
                              int *__restrict__ ent_sg)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= E) return;
    const int sg = seg_of(t.eoff, 2 * n_sim, e);
    const int s = sg >> 1, j = e - t.eoff[sg];
    const int n_s = t.poff[s + 1] - t.poff[s];
    // the sign's subset of simulation s starts at sub_pool[2 poff[s] + g n_s]
    const int *sub = sub_pool + 2 * (size_t)t.poff[s] + (size_t)(sg & 1) * n_s;
    ent_p[e] = j < counts[sg] ? t.poff[s] + sub[j] : -1;
    ent_sg[e] = sg;
}

// field keys and records: key = global cell, record = (x, y, |w|, 0);
// set_xy = the entry's position (branch-1 sampling)
__global__ void k_ens_field_keys(EnsTabs t, int E, uint32_t dead, const int *__restrict__ ent_p,
                                 const int *__restrict__ ent_sg, const float2 *__restrict__ pos,
                                 const float *__restrict__ omega, uint32_t *__restrict__ key,
                                 float4 *__restrict__ rec, float2 *__restrict__ set_xy,
                                 const float4 *__restrict__ jac, float4 *__restrict__ set_jac)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= E) return;
    const int p = ent_p[e];
    if (p < 0) {
        key[e] = dead;
        rec[e] = make_float4(0.f, 0.f, 0.f, 0.f);
        return;
    }
    const int sg = ent_sg[e];
    const BinGeom &g = t.geom[sg >> 1];
    const float2 x = pos[p];
    // the single-simulation field key (field.hip, k_field_keys)
    const int ix = min(max((int)floorf(x.x * g.inv_cx), 0), g.nx - 1);
    const int iy = min(max((int)floorf(x.y * g.inv_cy), 0), g.ny - 1);
    key[e] = (uint32_t)(t.coff[sg] + ix * g.ny + iy);
    // deformed: the record's 4th word carries its entry (for its quadratic form)
    rec[e] = make_float4(x.x, x.y, fabsf(omega[p]), jac ? __uint_as_float((uint32_t)e) : 0.f);
    set_xy[e] = x;
    if (jac) set_jac[e] = jac[p];
}

// deformed fields: the quadratic form of each sorted record's particle
__global__ void k_ens_field_gq(int E, const float4 *__restrict__ pts,
                               const int *__restrict__ ent_p, const float4 *__restrict__ jac,
                               float4 *__restrict__ gq)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= E) return;
    const int p = ent_p[__float_as_uint(pts[k].w)];
    gq[k] = quad_form(p >= 0 ? jac[p] : make_float4(1.f, 0.f, 0.f, 1.f));
}

// F of every particle from its simulation's five estimates, before the move
__global__ void k_ens_update_jac(EnsTabs t, int n_sim, int N, float dt, float4 *__restrict__ jac)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const int s = seg_of(t.poff, n_sim, i);
    const McD &d = t.sim[s];
    jac[i] = jacobian_from5(d.n_owner, i - t.poff[s], (const float2 *)d.vel_aux,
                            0.5f * dt / d.aux_dx);
}

template <int MODE>
__global__ void __launch_bounds__(BIN_THREADS / 2) __attribute__((amdgpu_waves_per_eu(8)))
    k_ens_gen(EnsTabs t, int n_sim, uint32_t *__restrict__ key,
              SRec *__restrict__ rec0, float2 *__restrict__ r)
{
    // BIN_THREADS samples a block of BIN_THREADS / 2 lanes: two a lane, 64
    // apart, their loads overlapped (bin_gen_pair); a simulation's sample range
    // is padded to BIN_THREADS, so the block stays in one simulation
    const int first = blockIdx.x * BIN_THREADS;
    const int s = seg_of(t.soff, n_sim, first);   // block-uniform
    const McD &d = t.sim[s];
    const BinGeom &g = t.geom[s];
    const uint32_t dk = (uint32_t)(t.toff[s] + g.ntile);   // the simulation's dead tile
    const size_t out0 = (size_t)first + (threadIdx.x / WAVE) * (2 * WAVE) +
                        (threadIdx.x & (WAVE - 1));
    const size_t out1 = out0 + WAVE;
    const size_t sid0 = out0 - (size_t)t.soff[s], sid1 = out1 - (size_t)t.soff[s];
    const size_t outs[2] = {out0, out1}, sids[2] = {sid0, sid1};
    const int count = sid1 < g.S ? 2 : (sid0 < g.S ? 1 : 0);
    if (count > 0)
        bin_gen_pair<MODE, true>(d, g, sids, outs, count, (uint32_t)t.toff[s], dk, key, rec0, r);
    for (int q = count; q < 2; ++q) {   // padding of the simulation's sample range
        key[outs[q]] = dk;
        rec0[outs[q]] = SRec{0u, (uint32_t)outs[q]};
        r[outs[q]] = make_float2(__builtin_nanf(""), __builtin_nanf(""));   // no query
    }
}

// chunks of tile k of the ensemble's tiles (0 for an empty tile and for a
// simulation's dead tile, whose samples have no query); nch[ntile] = 0
__global__ void k_ens_tile_nch(EnsTabs t, int n_sim, int ntile, const int *__restrict__ start,
                               int chunk, int *__restrict__ nch)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k > ntile) return;
    int c = 0;
    if (k < ntile) {
        const int s = seg_of(t.toff, n_sim, k);
        if (k - t.toff[s] < t.geom[s].ntile) c = tile_nch(start[k + 1] - start[k], chunk);
    }
    nch[k] = c;
}

template <int THREADS, int SPL, bool DEF>
__global__ void __launch_bounds__(THREADS) __attribute__((amdgpu_waves_per_eu(EVAL_WAVES_PER_EU<THREADS, SPL, DEF>)))
    k_ens_eval(EnsTabs t, int n_sim, int ntile, TileChunks tc, const SRec *__restrict__ srec,
               float *__restrict__ om)
{
    int tile, beg, end;   // one workgroup per tile chunk (chunk_range)
    if (!chunk_range(tc, ntile, tile, beg, end)) return;   // uniform
    const int s = seg_of(t.toff, n_sim, tile);
    if (tile - t.toff[s] >= t.geom[s].ntile) return;   // the dead tile: no query
    __shared__ TileLds<THREADS, DEF> L;
    EvalCount ec;
    bin_eval_tile<THREADS, SPL, false, DEF>(t.sim[s], t.geom[s], tile - t.toff[s], beg, end,
                                            srec, om, L, ec);
}

template <int MODE>
__global__ void __launch_bounds__(BIN_THREADS)
    k_ens_final(EnsTabs t, int n_sim, int npair, const float2 *__restrict__ r,
                const float *__restrict__ om)
{
    // wave-uniform, stated so that the simulation's tables load through the
    // scalar unit (the Horner coefficients of K_s among them)
    const int pair = __builtin_amdgcn_readfirstlane(blockIdx.x * (BIN_THREADS / WAVE) +
                                                    threadIdx.x / WAVE);
    if (pair >= npair) return;
    const int s = seg_of(t.poff, n_sim, pair >> 1);
    const size_t so = (size_t)t.soff[s];
    bin_final_pair<MODE, true>(t.sim[s], t.geom[s], pair - 2 * t.poff[s], r + so, om + so);
}

__global__ void k_ens_advect(EnsTabs t, int n_sim, int N, float dt, float2 *__restrict__ pos,
                             float2 *__restrict__ vel)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const int s = seg_of(t.poff, n_sim, i);
    const McD &d = t.sim[s];
    const int j = i - t.poff[s], n = d.n_owner;
    const float2 *vs = d.vel_sign;
    pos[i] = advect_one(vs[j], vs[(size_t)n + j], pos[i], vel ? vel + i : nullptr, dt, d.Lx, d.Ly,
                        1);
}

__global__ void k_ens_step_add(McD *__restrict__ sims, int n_sim, uint32_t add, int set,
                               uint32_t value)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < n_sim) sims[s].step = set ? value : sims[s].step + add;
}

// ---- host layout ----------------------------------------------------------------

struct EnsLayout {
    int n_sim = 0, N = 0, E = 0, C = 0, T = 0, per_xcd = 0;
    int local_bits = 1, nblk = 0;   // the sample sort: bits of a local tile key, blocks
    bool def = false;               // kernel deformation (desc.jac)
    size_t S = 0;
    std::vector<int4> blk;          // the sample sort's block table (host copy)
    std::vector<uint32_t> blk_base;
    std::vector<int> poff, soff, toff, eoff, coff, nx, ny, glog2, goff;
    // workspace
    EnsTabs t{};
    int *sub_pool = nullptr, *counts = nullptr, *guides = nullptr;
    double *cdf_pool = nullptr, *totals = nullptr;
    int *ent_p = nullptr, *ent_sg = nullptr;
    uint32_t *fkey = nullptr, *fkeys_sorted = nullptr;
    float4 *frec = nullptr, *fpts = nullptr;
    float2 *set_xy = nullptr;
    int *fstart = nullptr;
    void *fsort_ws = nullptr;
    size_t fsort_bytes = 0;
    uint32_t *skey = nullptr, *skeys_sorted = nullptr;
    float2 *r = nullptr, *vel_sign = nullptr;
    SRec *rec0 = nullptr;   // samples in sample order (query point, id)
    float *om = nullptr;
    SRec *srec = nullptr;
    int *tstart = nullptr;
    int *tcoff = nullptr;                     // tile chunks: offsets, table
    int4 *ctab = nullptr;
    void *cscan_ws = nullptr;
    size_t cscan_bytes = 0;
    int chunk = 0;                            // samples per tile chunk
    int4 *sblk = nullptr;
    float4 *fgq = nullptr, *set_jac = nullptr;   // deformed fields / sets
    float2 *vel_aux = nullptr;
    uint32_t *sblk_base = nullptr;
    void *ssort_ws = nullptr, *init_ws = nullptr;
    size_t ssort_bytes = 0, init_bytes = 0;
};

static int ens_plan(const vpmc_ensemble_desc *e, EnsLayout *L)
{
    if (!e || e->n_sim <= 0 || !e->h_offset || !e->h_h_w || !e->h_L || e->N1 < 0 || e->N2 < 0 ||
        e->N1 + e->N2 <= 0)
        return VPMC_ERR_ARG;
    if (e->mode != VPMC_MODE_REFERENCE && e->mode != VPMC_MODE_CONSISTENT) return VPMC_ERR_ARG;
    const int ns = e->n_sim;
    L->n_sim = ns;
    L->poff.assign(e->h_offset, e->h_offset + ns + 1);
    if (L->poff[0] != 0) return VPMC_ERR_ARG;
    L->soff.assign(ns + 1, 0);
    L->toff.assign(ns + 1, 0);
    L->eoff.assign(2 * ns + 1, 0);
    L->coff.assign(2 * ns + 1, 0);
    L->nx.assign(ns, 0);
    L->ny.assign(ns, 0);
    L->glog2.assign(ns, 0);
    L->goff.assign(2 * ns + 1, 0);
    L->def = e->jac != nullptr;
    const size_t NS = (size_t)samples_per_pair(e->N1, e->N2, L->def ? 1 : 0);
    for (int s = 0; s < ns; ++s) {
        const int n = L->poff[s + 1] - L->poff[s];
        if (n <= 0 || !(e->h_h_w[s] > 0.f) || !(e->h_L[s] > 0.f)) return VPMC_ERR_ARG;
        VPMC_TRY(vpmc_field_grid(e->h_h_w[s], e->h_L[s], e->h_L[s], e->cells_per_hw, &L->nx[s],
                                 &L->ny[s]));
        const size_t Ss = (size_t)n * 2 * NS;
        size_t so = (size_t)L->soff[s] + ((Ss + BIN_THREADS - 1) / BIN_THREADS) * BIN_THREADS;
        if (so >= ((size_t)1 << 31)) return VPMC_ERR_ARG;
        L->soff[s + 1] = (int)so;
        const int cells = L->nx[s] * L->ny[s];
        // 2 * cells tiles, then the simulation's dead tile (samples without a
        // query), so that every key of simulation s lies in [toff[s], toff[s+1])
        L->toff[s + 1] = L->toff[s] + 2 * cells + 1;
        const int lb = key_bits_for((uint32_t)(2 * cells));
        if (lb > L->local_bits) L->local_bits = lb;
        for (int g = 0; g < 2; ++g) {
            L->eoff[2 * s + g + 1] = L->eoff[2 * s + g] + n;
            L->coff[2 * s + g + 1] = L->coff[2 * s + g] + cells;
        }
        int gl = 0;
        while ((1 << gl) < n) ++gl;
        L->glog2[s] = gl;
        for (int g = 0; g < 2; ++g)
            L->goff[2 * s + g + 1] = L->goff[2 * s + g] + (1 << gl) + 1;
    }
    L->N = L->poff[ns];
    L->S = (size_t)L->soff[ns];
    L->T = L->toff[ns];
    L->E = L->eoff[2 * ns];
    L->C = L->coff[2 * ns];
    L->per_xcd = (L->T + 7) / 8;
    {   // the evaluation's launch shape (as ens_estimate picks it) sets the chunk
        const int th = e->eval_threads ? e->eval_threads : 256;
        const int spl = e->spl == 1 ? 1 : (e->spl == 4 ? 4 : 2);
        L->chunk = TILE_CHUNK_BATCHES * th * spl * (L->def ? TILE_CHUNK_DEF : 1);
    }
    std::vector<uint32_t> base(ns);
    for (int s = 0; s < ns; ++s) base[s] = (uint32_t)L->toff[s];
    VPMC_TRY(radix_seg_blocks(L->soff, base, &L->blk, &L->blk_base));
    L->nblk = (int)L->blk.size();
    return VPMC_OK;
}

static size_t subsets_ws_max(const EnsLayout &L)
{
    size_t m = 0;
    for (int s = 0; s < L.n_sim; ++s) {
        size_t b = vpmc_subsets_workspace_bytes(L.poff[s + 1] - L.poff[s]);
        if (b > m) m = b;
    }
    return m;
}

// carve (or, with ws == NULL, measure) the workspace
static size_t ens_carve(EnsLayout &L, void *ws, size_t bytes)
{
    Carve cv(ws ? ws : (void *)nullptr, ws ? bytes : ~(size_t)0);
    const int ns = L.n_sim;
    const size_t N = (size_t)L.N, E = (size_t)L.E, S = L.S > 0 ? L.S : 1;
    L.t.sim = cv.take<McD>(ns);
    L.t.geom = cv.take<BinGeom>(ns);
    L.t.poff = cv.take<int>(ns + 1);
    L.t.soff = cv.take<int>(ns + 1);
    L.t.toff = cv.take<int>(ns + 1);
    L.t.eoff = cv.take<int>(2 * ns + 1);
    L.t.coff = cv.take<int>(2 * ns + 1);
    L.sub_pool = cv.take<int>(2 * N);
    L.cdf_pool = cv.take<double>(2 * N);
    L.counts = cv.take<int>(2 * ns);
    L.totals = cv.take<double>(2 * ns);
    L.guides = cv.take<int>((size_t)L.goff[2 * ns]);
    L.ent_p = cv.take<int>(E);
    L.ent_sg = cv.take<int>(E);
    L.fkey = cv.take<uint32_t>(E);
    L.frec = cv.take<float4>(E);
    L.fpts = cv.take<float4>(E);
    L.fkeys_sorted = cv.take<uint32_t>(E);
    L.set_xy = cv.take<float2>(E);
    L.fstart = cv.take<int>((size_t)L.C + 1);
    L.fsort_bytes = radix_workspace_bytes(L.E);
    L.fsort_ws = cv.take<char>(L.fsort_bytes);
    L.skey = cv.take<uint32_t>(S);
    L.rec0 = cv.take<SRec>(S);
    L.r = cv.take<float2>(S);
    L.om = cv.take<float>(S);
    L.srec = cv.take<SRec>(S);
    L.skeys_sorted = cv.take<uint32_t>(S);
    L.tstart = cv.take<int>((size_t)L.T + 1);
    L.tcoff = cv.take<int>((size_t)L.T + 1);
    L.ctab = cv.take<int4>(tile_chunks_max(L.T, S, L.chunk));
    L.cscan_bytes = scan_workspace_bytes(L.T + 1, 4);
    L.cscan_ws = cv.take<char>(L.cscan_bytes);
    L.sblk = cv.take<int4>((size_t)(L.nblk > 0 ? L.nblk : 1));
    L.sblk_base = cv.take<uint32_t>((size_t)(L.nblk > 0 ? L.nblk : 1));
    L.ssort_bytes = radix_seg_workspace_bytes((int)S, L.nblk);
    L.ssort_ws = cv.take<char>(L.ssort_bytes);
    L.vel_sign = cv.take<float2>(2 * N);
    L.fgq = cv.take<float4>(L.def ? E : 1);
    L.set_jac = cv.take<float4>(L.def ? E : 1);
    L.vel_aux = cv.take<float2>(L.def ? 8 * N : 1);
    L.init_bytes = subsets_ws_max(L);
    L.init_ws = cv.take<char>(L.init_bytes);
    return cv.used + 4096;
}

static int ens_layout(const vpmc_ensemble_desc *e, void *ws, size_t bytes, EnsLayout *L)
{
    VPMC_TRY(ens_plan(e, L));
    if (!ws) return VPMC_ERR_ARG;
    size_t need = ens_carve(*L, nullptr, 0);
    if (bytes < need) return VPMC_ERR_WORKSPACE;
    ens_carve(*L, ws, bytes);
    if (!e->pos_xy || !e->omega) return VPMC_ERR_ARG;
    return VPMC_OK;
}

// the single-simulation descriptor of simulation s inside the ensemble
static vpmc_mc_desc ens_sim_desc(const vpmc_ensemble_desc *e, const EnsLayout &L, int s)
{
    vpmc_mc_desc d{};
    const int n = L.poff[s + 1] - L.poff[s];
    d.mode = e->mode;
    d.N1 = e->N1;
    d.N2 = e->N2;
    d.h_w = e->h_h_w[s];
    d.Lx = d.Ly = e->h_L[s];
    d.periodic = 1;
    d.seed = e->seed;
    d.step = 0;
    d.sim = (uint32_t)s;
    d.n_owner = n;
    d.owner_xy = e->pos_xy + 2 * (size_t)L.poff[s];
    d.vel_sign = (float *)(L.vel_sign + 2 * (size_t)L.poff[s]);
    d.spl = e->spl;
    for (int g = 0; g < 2; ++g) {
        vpmc_sign_set &S = d.set[g];
        const int sg = 2 * s + g;
        S.n = n;
        S.count_dev = L.counts + sg;
        S.cdf = L.cdf_pool + 2 * (size_t)L.poff[s] + (size_t)g * n;
        S.total_dev = L.totals + sg;
        S.guide = L.guides + L.goff[sg];
        S.guide_log2 = L.glog2[s];
        S.set_xy = (const float *)(L.set_xy + L.eoff[sg]);
        S.field.nx = L.nx[s];
        S.field.ny = L.ny[s];
        S.field.Lx = S.field.Ly = e->h_L[s];
        S.field.n = n;
        S.field.pts = (const float *)L.fpts;            // global positions ...
        S.field.cell_start = L.fstart + L.coff[sg];      // ... indexed by global offsets
        S.field.order = nullptr;
        if (L.def) {   // the CUDA step's F (fields' forms, the sets' jitter)
            S.field.gq = (const float *)L.fgq;
            S.field.gq_bound = 0.f;
            S.set_jac = (const float *)(L.set_jac + L.eoff[sg]);
        }
    }
    if (L.def) {   // the four aux targets at dst -/+ dx, dx = h_w / 2 (SimParam::dx)
        d.aux = 1;
        d.aux_dx = 0.5f * e->h_h_w[s];
        d.vel_aux = (float *)(L.vel_aux + 8 * (size_t)L.poff[s]);
    }
    return d;
}

static int ens_fields(const vpmc_ensemble_desc *e, EnsLayout &L, hipStream_t st)
{
    const unsigned gb = (unsigned)((L.E + 255) / 256);
    const float4 *jac = L.def ? (const float4 *)e->jac : nullptr;
    hipLaunchKernelGGL(k_ens_field_keys, dim3(gb), dim3(256), 0, st, L.t, L.E, (uint32_t)L.C,
                       (const int *)L.ent_p, (const int *)L.ent_sg, (const float2 *)e->pos_xy,
                       e->omega, L.fkey, L.frec, L.set_xy, jac, L.set_jac);
    VPMC_CHECK_LAUNCH();
    VPMC_TRY(radix_sort_to_records(L.fkey, L.E, key_bits_for((uint32_t)L.C), nullptr, L.frec,
                                   L.fpts, L.fkeys_sorted, L.C, L.fstart, L.fsort_ws,
                                   L.fsort_bytes, st));
    if (L.def) {
        hipLaunchKernelGGL(k_ens_field_gq, dim3(gb), dim3(256), 0, st, L.E,
                           (const float4 *)L.fpts, (const int *)L.ent_p, jac, L.fgq);
        VPMC_CHECK_LAUNCH();
    }
    return VPMC_OK;
}

template <int MODE>
static int ens_estimate(const vpmc_ensemble_desc *e, EnsLayout &L, hipStream_t st)
{
    const int ns = L.n_sim;
    hipLaunchKernelGGL((k_ens_gen<MODE>), dim3((unsigned)(L.S / BIN_THREADS)), dim3(BIN_THREADS / 2),
                       0,
                       st, L.t, ns, L.skey, L.rec0, L.r);
    VPMC_CHECK_LAUNCH();
    // segmented by simulation: samples are grouped by simulation already
    VPMC_TRY(radix_sort_segmented(L.skey, (int)L.S, L.local_bits, L.rec0, L.srec,
                                  L.skeys_sorted, L.T, L.tstart, L.sblk, L.sblk_base, L.nblk,
                                  L.ssort_ws, L.ssort_bytes, st));
    const int th = e->eval_threads ? e->eval_threads : 256;
    // the tiles' chunks from the tile offsets
    hipLaunchKernelGGL(k_ens_tile_nch, dim3((unsigned)((L.T + 1 + 255) / 256)), dim3(256), 0, st,
                       L.t, ns, L.T, (const int *)L.tstart, L.chunk, L.tcoff);
    VPMC_TRY(tile_chunks_build(L.tcoff, L.T, L.tstart, L.chunk, L.ctab, L.cscan_ws,
                               L.cscan_bytes, st));
    const TileChunks tc{L.tcoff, L.ctab, L.chunk};
    const size_t cmax = tile_chunks_max(L.T, L.S, L.chunk);
    const unsigned egrid = (unsigned)(8 * ((cmax + 7) / 8));
    VPMC_TRY(record_event(e->eval_events[0], st));
#define VPMC_ENS_EVAL(TH, SP, DF)                                                              \
    hipLaunchKernelGGL((k_ens_eval<TH, SP, DF>), dim3(egrid), dim3(TH), 0, st, L.t, ns, L.T,   \
                       tc, (const SRec *)L.srec, L.om)
#define VPMC_ENS_EVAL_SPL(TH, DF)                        \
    do {                                                 \
        if (e->spl == 1) VPMC_ENS_EVAL(TH, 1, DF);       \
        else if (e->spl == 4) VPMC_ENS_EVAL(TH, 4, DF);  \
        else VPMC_ENS_EVAL(TH, 2, DF);                   \
    } while (0)
#define VPMC_ENS_EVAL_TH(DF)                               \
    do {                                                   \
        if (th == 64) VPMC_ENS_EVAL_SPL(64, DF);           \
        else if (th == 128) VPMC_ENS_EVAL_SPL(128, DF);    \
        else VPMC_ENS_EVAL_SPL(256, DF);                   \
    } while (0)
    if (L.def) VPMC_ENS_EVAL_TH(true);
    else VPMC_ENS_EVAL_TH(false);
#undef VPMC_ENS_EVAL_TH
#undef VPMC_ENS_EVAL_SPL
#undef VPMC_ENS_EVAL
    VPMC_CHECK_LAUNCH();
    VPMC_TRY(record_event(e->eval_events[1], st));
    const int npair = 2 * L.N;
    hipLaunchKernelGGL((k_ens_final<MODE>), dim3((unsigned)((npair + 3) / 4)), dim3(BIN_THREADS), 0,
                       st, L.t, ns, npair, (const float2 *)L.r, (const float *)L.om);
    VPMC_CHECK_LAUNCH();
    return VPMC_OK;
}

}  // namespace vpmc

using namespace vpmc;

extern "C" {

size_t vpmc_ensemble_workspace_bytes(const vpmc_ensemble_desc *e)
{
    EnsLayout L;
    if (ens_plan(e, &L) != VPMC_OK) return 0;
    return ens_carve(L, nullptr, 0);
}

int vpmc_ensemble_init(const vpmc_ensemble_desc *e, void *workspace, size_t workspace_bytes,
                       void *stream)
{
    EnsLayout L;
    VPMC_TRY(ens_layout(e, workspace, workspace_bytes, &L));
    hipStream_t st = as_stream(stream);
    const int ns = L.n_sim;
    // subsets, FP64 CDFs and guide tables, one simulation at a time (once)
    for (int s = 0; s < ns; ++s) {
        const int n = L.poff[s + 1] - L.poff[s];
        int *sub = L.sub_pool + 2 * (size_t)L.poff[s];
        double *cdf = L.cdf_pool + 2 * (size_t)L.poff[s];
        VPMC_TRY(vpmc_build_subsets(n, e->omega + L.poff[s], sub, sub + n, cdf, cdf + n,
                                    L.counts + 2 * s, L.totals + 2 * s, L.init_ws, L.init_bytes,
                                    stream));
        for (int g = 0; g < 2; ++g)
            VPMC_TRY(vpmc_build_guide(cdf + (size_t)g * n, L.counts + 2 * s + g, L.glog2[s],
                                      L.guides + L.goff[2 * s + g], stream));
    }
    // per-simulation parameter tables
    std::vector<McD> sims(ns);
    std::vector<BinGeom> geoms(ns);
    for (int s = 0; s < ns; ++s) {
        vpmc_mc_desc d = ens_sim_desc(e, L, s);
        VPMC_TRY(make_mcd(&d, &sims[s]));
        VPMC_TRY(bin_geom(sims[s], e->N1, e->N2, d.n_owner, &geoms[s]));
        const int th = e->eval_threads ? e->eval_threads : 256;
        if (th != 64 && th != 128 && th != 256) return VPMC_ERR_ARG;
        if (!tile_threads_ok(geoms[s], th)) return VPMC_ERR_GRID;
    }
    struct Up { void *dst; const void *src; size_t bytes; } ups[] = {
        {L.t.sim, sims.data(), sizeof(McD) * ns},
        {L.t.geom, geoms.data(), sizeof(BinGeom) * ns},
        {L.t.poff, L.poff.data(), sizeof(int) * (ns + 1)},
        {L.t.soff, L.soff.data(), sizeof(int) * (ns + 1)},
        {L.t.toff, L.toff.data(), sizeof(int) * (ns + 1)},
        {L.t.eoff, L.eoff.data(), sizeof(int) * (2 * ns + 1)},
        {L.t.coff, L.coff.data(), sizeof(int) * (2 * ns + 1)},
        {L.sblk, L.blk.data(), sizeof(int4) * L.blk.size()},
        {L.sblk_base, L.blk_base.data(), sizeof(uint32_t) * L.blk_base.size()},
    };
    hipError_t err = hipStreamSynchronize(st);
    for (const Up &u : ups)
        if (err == hipSuccess) err = hipMemcpy(u.dst, u.src, u.bytes, hipMemcpyHostToDevice);
    if (err != hipSuccess) { set_hip_error(err); return VPMC_ERR_HIP; }
    hipLaunchKernelGGL(k_ens_entries, dim3((unsigned)((L.E + 255) / 256)), dim3(256), 0, st, L.t,
                       ns, L.E, (const int *)L.sub_pool, (const int *)L.counts, L.ent_p, L.ent_sg);
    VPMC_CHECK_LAUNCH();
    err = hipStreamSynchronize(st);
    if (err != hipSuccess) { set_hip_error(err); return VPMC_ERR_HIP; }
    return VPMC_OK;
}

int vpmc_ensemble_set_step(const vpmc_ensemble_desc *e, void *workspace, size_t workspace_bytes,
                           uint32_t step, void *stream)
{
    EnsLayout L;
    VPMC_TRY(ens_layout(e, workspace, workspace_bytes, &L));
    hipLaunchKernelGGL(k_ens_step_add, dim3((unsigned)((L.n_sim + 63) / 64)), dim3(64), 0,
                       as_stream(stream), L.t.sim, L.n_sim, 0u, 1, step);
    VPMC_CHECK_LAUNCH();
    return VPMC_OK;
}

int vpmc_ensemble_step(const vpmc_ensemble_desc *e, void *workspace, size_t workspace_bytes,
                       void *stream)
{
    EnsLayout L;
    VPMC_TRY(ens_layout(e, workspace, workspace_bytes, &L));
    hipStream_t st = as_stream(stream);
    VPMC_TRY(ens_fields(e, L, st));
    if (e->mode == VPMC_MODE_REFERENCE) VPMC_TRY(ens_estimate<VPMC_MODE_REFERENCE>(e, L, st));
    else VPMC_TRY(ens_estimate<VPMC_MODE_CONSISTENT>(e, L, st));
    if (L.def) {   // F from the five estimates, before the move (vpmc_step with jac)
        hipLaunchKernelGGL(k_ens_update_jac, dim3((unsigned)((L.N + 255) / 256)), dim3(256), 0,
                           st, L.t, L.n_sim, L.N, e->dt, (float4 *)e->jac);
        VPMC_CHECK_LAUNCH();
    }
    hipLaunchKernelGGL(k_ens_advect, dim3((unsigned)((L.N + 255) / 256)), dim3(256), 0, st, L.t,
                       L.n_sim, L.N, e->dt, (float2 *)e->pos_xy, (float2 *)e->vel_xy);
    hipLaunchKernelGGL(k_ens_step_add, dim3((unsigned)((L.n_sim + 63) / 64)), dim3(64), 0, st,
                       L.t.sim, L.n_sim, 1u, 0, 0u);
    VPMC_CHECK_LAUNCH();
    return VPMC_OK;
}

int vpmc_ensemble_graph_create(const vpmc_ensemble_desc *e, void *workspace,
                               size_t workspace_bytes, int steps, void *stream, void **graph_exec)
{
    if (!graph_exec || !stream || steps <= 0) return VPMC_ERR_ARG;
    *graph_exec = nullptr;
    hipStream_t st = as_stream(stream);
    hipError_t err = hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal);
    if (err != hipSuccess) { set_hip_error(err); return VPMC_ERR_HIP; }
    int rc = VPMC_OK;
    for (int k = 0; k < steps && rc == VPMC_OK; ++k)
        rc = vpmc_ensemble_step(e, workspace, workspace_bytes, stream);
    hipGraph_t graph = nullptr;
    err = hipStreamEndCapture(st, &graph);
    if (rc != VPMC_OK) {
        if (graph) (void)hipGraphDestroy(graph);
        return rc;
    }
    if (err != hipSuccess) { set_hip_error(err); return VPMC_ERR_HIP; }
    hipGraphExec_t exec = nullptr;
    err = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (err != hipSuccess) { set_hip_error(err); return VPMC_ERR_HIP; }
    *graph_exec = (void *)exec;
    return VPMC_OK;
}

int vpmc_graph_launch(void *graph_exec, void *stream)
{
    if (!graph_exec) return VPMC_ERR_ARG;
    hipError_t err = hipGraphLaunch((hipGraphExec_t)graph_exec, as_stream(stream));
    if (err != hipSuccess) { set_hip_error(err); return VPMC_ERR_HIP; }
    return VPMC_OK;
}

int vpmc_graph_destroy(void *graph_exec)
{
    if (!graph_exec) return VPMC_OK;
    hipError_t err = hipGraphExecDestroy((hipGraphExec_t)graph_exec);
    if (err != hipSuccess) { set_hip_error(err); return VPMC_ERR_HIP; }
    return VPMC_OK;
}

}  // extern "C"
