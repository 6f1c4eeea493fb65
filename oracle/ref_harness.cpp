// ref_harness.cpp -- C entry points over the reference's own CPU path.
//
// TEST INFRASTRUCTURE ONLY.  Linked with the reference's src/cpu sources by
// oracle/build.py into oracle/_ref/libvpmc_ref.so; used by tests to pin the
// oracle and by bench.py as the "reference" CPU baseline.  Nothing here
// restates the algorithm: each function builds the reference's own types and
// calls its functions.
#include <omp.h>
#include <cstdint>
#include <random>
#include <span>
#include <vector>

#include "monte_carlo_bs.hpp"
#include "spatial_hash.hpp"
#include "utilities.hpp"
#include "vortex_particle_mc.hpp"

extern "C" {

float ref_spline(float r, float h_w) { return cubicSplinePDF(r, h_w); }

void ref_fill_cells(int rshx, int rshy, float Lx, float Ly, int n, const int *sub,
                    int n_total, const float *px, const float *py, int *cell_num, int *cell_acc,
                    int *vp_cell)
{
    int cells = rshx * rshy;
    FillCells_cpu(rshx, rshy, Lx, Ly, n, std::span<const int>(sub, n),
                  std::span<const float>(px, n_total), std::span<const float>(py, n_total),
                  std::span<int>(cell_num, cells), std::span<int>(cell_acc, cells),
                  std::span<int>(vp_cell, n));
}

void ref_counting_sort(int n, int cells, int *cell_num, const int *cell_acc, const int *vp_cell,
                       int *vp_map)
{
    CountingSort_cpu(n, std::span<int>(cell_num, cells), std::span<const int>(cell_acc, cells),
                     std::span<const int>(vp_cell, n), std::span<int>(vp_map, n));
}

// the OpenMP fill (its histogram and scan are deterministic)
void ref_fill_cells_omp(int rshx, int rshy, float Lx, float Ly, int n, const int *sub,
                        int n_total, const float *px, const float *py, int *cell_num,
                        int *cell_acc, int *vp_cell)
{
    int cells = rshx * rshy;
    FillCells_omp(rshx, rshy, Lx, Ly, n, std::span<const int>(sub, n),
                  std::span<const float>(px, n_total), std::span<const float>(py, n_total),
                  std::span<int>(cell_num, cells), std::span<int>(cell_acc, cells),
                  std::span<int>(vp_cell, n));
}

struct RefSet {
    std::vector<float2> pos;
    std::vector<float4> F;
    std::vector<float> w;
    std::vector<int> sub_pos, sub_neg, vp_pos, vp_neg, acc_pos, acc_neg;
    std::vector<float> cdf_pos, cdf_neg;
};

static VorticityView make_view(const RefSet &S, bool positive, int rshx, int rshy, float Lx,
                               float Ly)
{
    VorticityView v{};
    v.pos = S.pos.data();
    v.F = S.F.data();
    v.w = S.w.data();
    v.N = (int)S.pos.size();
    v.vp_sort = positive ? S.vp_pos.data() : S.vp_neg.data();
    v.cell_acc = positive ? S.acc_pos.data() : S.acc_neg.data();
    v.rshx = rshx;
    v.rshy = rshy;
    v.Lx = Lx;
    v.Ly = Ly;
    v.InvLx = 1.f / Lx;
    v.InvLy = 1.f / Ly;
    v.cnt_view = positive ? (int)S.sub_pos.size() : (int)S.sub_neg.size();
    v.sub = positive ? S.sub_pos.data() : S.sub_neg.data();
    return v;
}

// Builds the reference's sign subsets (FP32 CDF, src/cpu/build_particle_sets.cpp
// semantics are in the Simulation class, so the sequential sums are done here
// the same way) and its per-sign hash, then answers queries / MC estimates.
static RefSet *build_set(int n, const float *px, const float *py, const float *w, int rshx,
                         int rshy, float Lx, float Ly, float *cum_out)
{
    RefSet *S = new RefSet;
    S->pos.resize(n);
    S->F.resize(n);
    S->w.assign(w, w + n);
    for (int i = 0; i < n; ++i) {
        S->pos[i] = float2{px[i], py[i]};
        S->F[i] = float4{1.f, 0.f, 0.f, 1.f};
    }
    float cp = 0.f, cn = 0.f;
    for (int i = 0; i < n; ++i) {
        if (w[i] > 0) { S->sub_pos.push_back(i); cp += w[i]; S->cdf_pos.push_back(cp); }
        else if (w[i] < 0) { S->sub_neg.push_back(i); cn -= w[i]; S->cdf_neg.push_back(cn); }
    }
    if (cp > 0.f) { float inv = 1.f / cp; for (float &v : S->cdf_pos) v *= inv; }
    if (cn > 0.f) { float inv = 1.f / cn; for (float &v : S->cdf_neg) v *= inv; }
    cum_out[0] = cp;
    cum_out[1] = cn;
    std::vector<float> X(px, px + n), Y(py, py + n);
    int cells = rshx * rshy;
    for (int s = 0; s < 2; ++s) {
        auto &sub = s == 0 ? S->sub_pos : S->sub_neg;
        auto &vp = s == 0 ? S->vp_pos : S->vp_neg;
        auto &acc = s == 0 ? S->acc_pos : S->acc_neg;
        int m = (int)sub.size();
        std::vector<int> num(cells), cell(m);
        acc.assign(cells, 0);
        vp.assign(m, 0);
        FillCells_cpu(rshx, rshy, Lx, Ly, m, sub, X, Y, num, acc, cell);
        CountingSort_cpu(m, num, acc, cell, vp);
    }
    return S;
}

// evalVorticityAbs_cpu at npts points against the positive (sign>0), negative
// (sign<0) or all-particle set; returns the FP32 CDFs and totals as a side
// product so tests can hand the oracle identical inputs.
void ref_eval_vorticity(int n, const float *px, const float *py, const float *w, int rshx,
                        int rshy, float Lx, float Ly, int npts, const float *qx,
                        const float *qy, int sign, int use_signed, int exclude_center,
                        float h_w, int periodic, float *out)
{
    float cum[2];
    RefSet *S = build_set(n, px, py, w, rshx, rshy, Lx, Ly, cum);
    VorticityView v = make_view(*S, sign > 0, rshx, rshy, Lx, Ly);
    for (int i = 0; i < npts; ++i)
        out[i] = evalVorticityAbs_cpu(v, float2{qx[i], qy[i]}, sign, use_signed != 0,
                                      exclude_center != 0, h_w, periodic != 0);
    delete S;
}

void ref_subsets_f32(int n, const float *w, int *cnt, float *cdf_pos, float *cdf_neg,
                     float *cum)
{
    float cp = 0.f, cn = 0.f;
    int a = 0, b = 0;
    for (int i = 0; i < n; ++i) {
        if (w[i] > 0) { cp += w[i]; cdf_pos[a++] = cp; }
        else if (w[i] < 0) { cn -= w[i]; cdf_neg[b++] = cn; }
    }
    if (cp > 0.f) { float inv = 1.f / cp; for (int j = 0; j < a; ++j) cdf_pos[j] *= inv; }
    if (cn > 0.f) { float inv = 1.f / cn; for (int j = 0; j < b; ++j) cdf_neg[j] *= inv; }
    cnt[0] = a; cnt[1] = b; cum[0] = cp; cum[1] = cn;
}

// The reference's step velocity for owners [o0, o1): monte_carlo_bs_cpu for
// sign +1 then -1 per owner with one std::mt19937 shared in that order, as
// src/cpu/simulation_step.cpp:28-57 does.  aux targets are the owner itself
// (their results are discarded; aux only changes how many draws are used).
void ref_mc_velocities(int n, const float *px, const float *py, const float *w, int rshx,
                       int rshy, float Lx, float Ly, float h_w, int periodic, int N1, int N2,
                       int aux, uint32_t seed, int o0, int o1, float *vel)
{
    float cum[2];
    RefSet *S = build_set(n, px, py, w, rshx, rshy, Lx, Ly, cum);
    ParticleSet P;
    P.resize(n);
    for (int i = 0; i < n; ++i) P.pos[i] = S->pos[i];
    P.omega = S->w;
    MCBSCtx C;
    C.sub_pos = &S->sub_pos;
    C.sub_neg = &S->sub_neg;
    C.cdf_pos = &S->cdf_pos;
    C.cdf_neg = &S->cdf_neg;
    C.h_w = h_w;
    C.periodic = periodic != 0;
    C.Lx = Lx; C.Ly = Ly; C.InvLx = 1.f / Lx; C.InvLy = 1.f / Ly;
    C.cum_pos = cum[0];
    C.cum_neg = cum[1];
    C.pos_view = make_view(*S, true, rshx, rshy, Lx, Ly);
    C.neg_view = make_view(*S, false, rshx, rshy, Lx, Ly);
    std::mt19937 rng(seed);
    for (int i = o0; i < o1; ++i) {
        float2 d = P.pos[i];
        MCBSEstimate vp = monte_carlo_bs_cpu(d, +1, aux, d, d, d, d, P, C, N1, N2, rng);
        MCBSEstimate vn = monte_carlo_bs_cpu(d, -1, aux, d, d, d, d, P, C, N1, N2, rng);
        vel[2 * (i - o0)] = vp.main.x - vn.main.x;
        vel[2 * (i - o0) + 1] = vp.main.y - vn.main.y;
    }
    delete S;
}

// ref_mc_velocities over nthreads OpenMP threads (owners split dynamically),
// one std::mt19937 per thread seeded seed + 1000003 * thread: the same
// estimator on other draws, for timing the reference's CPU path on all cores
void ref_mc_velocities_mt(int n, const float *px, const float *py, const float *w, int rshx,
                          int rshy, float Lx, float Ly, float h_w, int periodic, int N1, int N2,
                          uint32_t seed, int o0, int o1, int nthreads, float *vel)
{
    float cum[2];
    RefSet *S = build_set(n, px, py, w, rshx, rshy, Lx, Ly, cum);
    ParticleSet P;
    P.resize(n);
    for (int i = 0; i < n; ++i) P.pos[i] = S->pos[i];
    P.omega = S->w;
    MCBSCtx C;
    C.sub_pos = &S->sub_pos;
    C.sub_neg = &S->sub_neg;
    C.cdf_pos = &S->cdf_pos;
    C.cdf_neg = &S->cdf_neg;
    C.h_w = h_w;
    C.periodic = periodic != 0;
    C.Lx = Lx; C.Ly = Ly; C.InvLx = 1.f / Lx; C.InvLy = 1.f / Ly;
    C.cum_pos = cum[0];
    C.cum_neg = cum[1];
    C.pos_view = make_view(*S, true, rshx, rshy, Lx, Ly);
    C.neg_view = make_view(*S, false, rshx, rshy, Lx, Ly);
#pragma omp parallel num_threads(nthreads)
    {
        std::mt19937 rng(seed + 1000003u * (uint32_t)omp_get_thread_num());
#pragma omp for schedule(dynamic, 8)
        for (int i = o0; i < o1; ++i) {
            float2 d = P.pos[i];
            MCBSEstimate vp = monte_carlo_bs_cpu(d, +1, 0, d, d, d, d, P, C, N1, N2, rng);
            MCBSEstimate vn = monte_carlo_bs_cpu(d, -1, 0, d, d, d, d, P, C, N1, N2, rng);
            vel[2 * (i - o0)] = vp.main.x - vn.main.x;
            vel[2 * (i - o0) + 1] = vp.main.y - vn.main.y;
        }
    }
    delete S;
}

// The reference's estimates with real auxiliary targets (dst -/+ dx along x,
// dst +/- dx along y, the CUDA step's placement, src/cuda/monte_carlo_bs.cu:166-169)
// for owners [o0, o1): sign +1 then -1 per owner on one std::mt19937.
// out[((i - o0) * 2 + sign_index) * 10 + 2 * target + c], targets main, L, R, T, B.
void ref_mc_aux(int n, const float *px, const float *py, const float *w, int rshx, int rshy,
                float Lx, float Ly, float h_w, int periodic, int N1, int N2, float dx,
                uint32_t seed, int o0, int o1, float *out)
{
    float cum[2];
    RefSet *S = build_set(n, px, py, w, rshx, rshy, Lx, Ly, cum);
    ParticleSet P;
    P.resize(n);
    for (int i = 0; i < n; ++i) P.pos[i] = S->pos[i];
    P.omega = S->w;
    MCBSCtx C;
    C.sub_pos = &S->sub_pos;
    C.sub_neg = &S->sub_neg;
    C.cdf_pos = &S->cdf_pos;
    C.cdf_neg = &S->cdf_neg;
    C.h_w = h_w;
    C.periodic = periodic != 0;
    C.Lx = Lx; C.Ly = Ly; C.InvLx = 1.f / Lx; C.InvLy = 1.f / Ly;
    C.cum_pos = cum[0];
    C.cum_neg = cum[1];
    C.pos_view = make_view(*S, true, rshx, rshy, Lx, Ly);
    C.neg_view = make_view(*S, false, rshx, rshy, Lx, Ly);
    std::mt19937 rng(seed);
    for (int i = o0; i < o1; ++i) {
        float2 d = P.pos[i];
        float2 dl{d.x - dx, d.y}, dr{d.x + dx, d.y}, dt{d.x, d.y + dx}, db{d.x, d.y - dx};
        for (int sg = 0; sg < 2; ++sg) {
            MCBSEstimate e = monte_carlo_bs_cpu(d, sg == 0 ? +1 : -1, 1, dl, dr, dt, db, P, C, N1,
                                                N2, rng);
            float2 v[5] = {e.main, e.left, e.right, e.top, e.bottom};
            float *o = out + ((size_t)(i - o0) * 2 + sg) * 10;
            for (int t = 0; t < 5; ++t) { o[2 * t] = v[t].x; o[2 * t + 1] = v[t].y; }
        }
    }
    delete S;
}

// The draws a std::mt19937(seed) + uniform_real_distribution<float>(0,1)
// produce, in order: the injected-uniform stream for replay parity.
void ref_uniform_stream(uint32_t seed, int64_t count, float *out)
{
    std::mt19937 rng(seed);
    std::uniform_real_distribution<float> U(0.f, 1.f);
    for (int64_t i = 0; i < count; ++i) out[i] = U(rng);
}

}  // extern "C"
