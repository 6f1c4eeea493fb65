// binned.hip -- the estimator with samples binned by fine cell.
//
// A sample's query needs every particle of its sign within h_w.  Samples of
// one owner are spread over the whole box, so the direct kernel (one
// workgroup per owner) gathers a different neighbourhood per lane.  Here the
// step is split in four passes so that the 64 lanes of a wave share one
// neighbourhood:
//   gen     one thread per sample: Philox (or injected) draws -> query point,
//           MIS displacement, fine-cell key; atomic per-cell count (slot).
//   scatter samples into cell order (exclusive scan of the counts).
//   eval    one workgroup per cell: every wave takes 64*SPL samples of the
//           cell and walks the cell's neighbour stencil (the cells within
//           h_w of the cell rectangle, one contiguous run of the sorted
//           particle array per column).  Candidate records are read at
//           wave-uniform addresses (scalar loads), the per-sample work is
//           pure VALU.  omega is written back to the sample's id.
//   final   one wave per (owner, sign): MIS contributions of its N1+N2
//           samples in sample order, butterfly sum -> deterministic result.
// Sample ids are owner-major: sid = (owner*2 + sign)*(N1+N2) + k.
#include "mc_common.hpp"

namespace vpmc {

constexpr int BIN_THREADS = 256;

struct BinGeom {
    int ncell, nx, ny;
    float inv_cx, inv_cy;
    int cols;                 // stencil half-width (columns)
    int ext[MAX_COL_EXT];     // rows each side for column offset ox (index ox+cols), -1: none
    uint32_t NS;              // N1 + N2
    size_t S;                 // samples
};

struct BinBufs {
    int *key, *slot;
    float2 *q, *r;
    float *om;
    float2 *sq;
    uint32_t *ssid;
    int *count, *start;
    void *scan_ws;
    size_t scan_bytes;
};

template <int MODE, bool PER>
__global__ void __launch_bounds__(BIN_THREADS)
    k_bin_gen(McD d, BinGeom g, int *__restrict__ key, int *__restrict__ slot,
              float2 *__restrict__ q, float2 *__restrict__ r, float *__restrict__ om,
              int *__restrict__ count)
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
    if (!live) {
        key[sid] = -1;
        return;
    }
    q[sid] = make_float2(qx, qy);
    int cx = min(max((int)(qx * g.inv_cx), 0), g.nx - 1);
    int cy = min(max((int)(qy * g.inv_cy), 0), g.ny - 1);
    int kk = si * g.ncell + cx * g.ny + cy;
    key[sid] = kk;
    slot[sid] = atomicAdd(count + kk, 1);
}

__global__ void __launch_bounds__(BIN_THREADS)
    k_bin_scatter(BinGeom g, const int *__restrict__ key, const int *__restrict__ slot,
                  const float2 *__restrict__ q, const int *__restrict__ start,
                  float2 *__restrict__ sq, uint32_t *__restrict__ ssid)
{
    size_t sid = (size_t)blockIdx.x * BIN_THREADS + threadIdx.x;
    if (sid >= g.S) return;
    int kk = key[sid];
    if (kk < 0) return;
    int dst = start[kk] + slot[sid];
    sq[dst] = q[sid];
    ssid[dst] = (uint32_t)sid;
}

// one contiguous run [b, e) of candidates seen through a uniform image shift
template <int SPL, bool COUNT>
__device__ __forceinline__ void bin_run(const float4 *__restrict__ pts, int b, int e,
                                        const float (&qx)[SPL], const float (&qy)[SPL],
                                        float sx, float sy, float inv_hw, float (&a1)[SPL],
                                        float (&a2)[SPL], EvalCount &ec,
                                        const bool (&live)[SPL])
{
    float px[SPL], py[SPL];
#pragma unroll
    for (int s = 0; s < SPL; ++s) {
        px[s] = qx[s] - sx;
        py[s] = qy[s] - sy;
    }
    // q >= 0, so 1-q <= 1 and 1/2-q <= 1/2: a max with 0 is the whole clamp
    auto one = [&](const float4 &c) {
#pragma unroll
        for (int s = 0; s < SPL; ++s) {
            float dx = c.x - px[s], dy = c.y - py[s];
            float r2 = fmaf(dx, dx, dy * dy);
            float qq = __builtin_amdgcn_sqrtf(r2) * inv_hw;
            float t = fmaxf(1.f - qq, 0.f);
            float u = fmaxf(0.5f - qq, 0.f);
            a1[s] = fmaf(c.z * t, t * t, a1[s]);   // record = (x, y, |w|, 0)
            a2[s] = fmaf(c.z * u, u * u, a2[s]);
            if (COUNT && live[s]) {
                ec.in_support += (qq < 1.f) ? 1u : 0u;
                ec.visited += 1u;
            }
        }
    };
    int k = b;
    for (; k + 4 <= e; k += 4) {   // four wave-uniform records per batch of loads
        const float4 c0 = pts[k], c1 = pts[k + 1], c2 = pts[k + 2], c3 = pts[k + 3];
        one(c0);
        one(c1);
        one(c2);
        one(c3);
    }
    for (; k < e; ++k) one(pts[k]);
}

template <int SPL, bool COUNT>
__global__ void __launch_bounds__(BIN_THREADS)
    k_bin_eval(BinGeom g, float inv_hw, float sig2, float Lx, float Ly,
               const float4 *__restrict__ pts0, const int *__restrict__ cs0,
               const float4 *__restrict__ pts1, const int *__restrict__ cs1,
               const int *__restrict__ start, const float2 *__restrict__ sq,
               const uint32_t *__restrict__ ssid, float *__restrict__ om,
               unsigned long long *__restrict__ eval_count)
{
    const int tile = blockIdx.x;
    const int beg = start[tile], end = start[tile + 1];
    EvalCount ec;
    if (beg < end) {
        const int si = tile >= g.ncell ? 1 : 0;
        const int c = tile - si * g.ncell;
        const int cx = c / g.ny, cy = c - (c / g.ny) * g.ny;
        const float4 *__restrict__ pts = si ? pts1 : pts0;
        const int *__restrict__ cell_start = si ? cs1 : cs0;
        const int lane = threadIdx.x & (WAVE - 1), wid = threadIdx.x / WAVE;
        constexpr int CHUNK = WAVE * SPL;
        for (int base = beg + wid * CHUNK; base < end; base += (BIN_THREADS / WAVE) * CHUNK) {
            float qx[SPL], qy[SPL], a1[SPL], a2[SPL];
            bool live[SPL];
#pragma unroll
            for (int s = 0; s < SPL; ++s) {
                int idx = base + s * WAVE + lane;
                live[s] = idx < end;   // tail lanes repeat a valid point, results dropped
                float2 p = live[s] ? sq[idx] : sq[beg];
                qx[s] = p.x;
                qy[s] = p.y;
                a1[s] = 0.f;
                a2[s] = 0.f;
            }
            for (int ox = -g.cols; ox <= g.cols; ++ox) {
                const int e = g.ext[ox + g.cols];
                if (e < 0) continue;
                int colw = cx + ox;
                float sx = 0.f;
                if (colw < 0) { colw += g.nx; sx = -Lx; }
                else if (colw >= g.nx) { colw -= g.nx; sx = Lx; }
                const int *csc = cell_start + (size_t)colw * g.ny;
                int r0 = cy - e, r1 = cy + e;
                if (r0 < 0) {   // rows ny+r0 .. ny-1, particle image y - Ly
                    bin_run<SPL, COUNT>(pts, csc[g.ny + r0], csc[g.ny], qx, qy, sx, -Ly, inv_hw,
                                        a1, a2, ec, live);
                    r0 = 0;
                }
                if (r1 >= g.ny) {   // rows 0 .. r1-ny, particle image y + Ly
                    bin_run<SPL, COUNT>(pts, csc[0], csc[r1 - g.ny + 1], qx, qy, sx, Ly, inv_hw,
                                        a1, a2, ec, live);
                    r1 = g.ny - 1;
                }
                bin_run<SPL, COUNT>(pts, csc[r0], csc[r1 + 1], qx, qy, sx, 0.f, inv_hw, a1, a2,
                                    ec, live);
            }
#pragma unroll
            for (int s = 0; s < SPL; ++s) {
                int idx = base + s * WAVE + lane;
                if (live[s]) om[ssid[idx]] = sig2 * (2.f * a1[s] - 8.f * a2[s]);
            }
        }
    }
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
    size_t S = g.S > 0 ? g.S : 1, T = (size_t)2 * g.ncell;
    return 2 * align256(S * 4) + 2 * align256(S * 8) + align256(S * 4) + align256(S * 8) +
           align256(S * 4) + align256(T * 4) + align256((T + 1) * 4) +
           scan_workspace_bytes((int)T, sizeof(int)) + 4096;
}

static int carve_bin(const BinGeom &g, void *ws, size_t bytes, BinBufs *B)
{
    Carve cv(ws, bytes);
    size_t S = g.S > 0 ? g.S : 1, T = (size_t)2 * g.ncell;
    B->key = cv.take<int>(S);
    B->slot = cv.take<int>(S);
    B->q = cv.take<float2>(S);
    B->r = cv.take<float2>(S);
    B->om = cv.take<float>(S);
    B->sq = cv.take<float2>(S);
    B->ssid = cv.take<uint32_t>(S);
    B->count = cv.take<int>(T);
    B->start = cv.take<int>(T + 1);
    B->scan_bytes = scan_workspace_bytes((int)T, sizeof(int));
    B->scan_ws = cv.take<char>(B->scan_bytes);
    if (!B->key || !B->slot || !B->q || !B->r || !B->om || !B->sq || !B->ssid || !B->count ||
        !B->start || !B->scan_ws)
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
                           B.slot, B.q, B.r, B.om, B.count);
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
    const int T = 2 * g.ncell;
    VPMC_TRY(fill_i32(B.count, 0, (size_t)T, s));
    launch_mode(d->mode, d->periodic, m, g, B, false, s);
    VPMC_CHECK_LAUNCH();
    VPMC_TRY(scan_i32(B.count, B.start, T, false, B.start + T, B.scan_ws, B.scan_bytes, s));
    unsigned gb = (unsigned)((g.S + BIN_THREADS - 1) / BIN_THREADS);
    hipLaunchKernelGGL(k_bin_scatter, dim3(gb), dim3(BIN_THREADS), 0, s, g, (const int *)B.key,
                       (const int *)B.slot, (const float2 *)B.q, (const int *)B.start, B.sq, B.ssid);
    VPMC_CHECK_LAUNCH();
    const float4 *p0 = m.set[0].field.pts, *p1 = m.set[1].field.pts;
    const int *c0 = m.set[0].field.cell_start, *c1 = m.set[1].field.cell_start;
    VPMC_TRY(record_event(d->eval_events[0], s));
    if (d->eval_count)
        hipLaunchKernelGGL((k_bin_eval<2, true>), dim3(T), dim3(BIN_THREADS), 0, s, g, m.inv_hw,
                           m.sig2, m.Lx, m.Ly, p0, c0, p1, c1, (const int *)B.start,
                           (const float2 *)B.sq, (const uint32_t *)B.ssid, B.om, d->eval_count);
    else
        hipLaunchKernelGGL((k_bin_eval<2, false>), dim3(T), dim3(BIN_THREADS), 0, s, g, m.inv_hw,
                           m.sig2, m.Lx, m.Ly, p0, c0, p1, c1, (const int *)B.start,
                           (const float2 *)B.sq, (const uint32_t *)B.ssid, B.om,
                           (unsigned long long *)nullptr);
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
