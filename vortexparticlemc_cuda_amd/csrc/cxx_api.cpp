// cxx_api.cpp -- the reference's C++ API (vortex_particle_mc.hpp,
// spatial_hash.hpp, spatial_hasher.hpp, monte_carlo_bs.hpp, cg_rbf.hpp,
// cg_rbf_init.cuh, utilities.hpp) implemented on the C ABI of vpmc.h.
// Host code only: every computation is a libvpmc kernel reached through
// vpmc_* calls.  C ABI status codes become exceptions here, where the
// reference printed, threw or asserted.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <iostream>
#include <numeric>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

#include "analytical_flow.hpp"
#include "cg_rbf.hpp"
#include "cg_rbf_init.cuh"
#include "monte_carlo_bs.hpp"
#include "spatial_hash.hpp"
#include "utilities.hpp"
#include "vortex_particle_mc.hpp"
#include "vpmc.h"

namespace {

void vcheck(int st, const char *what)
{
    if (st != VPMC_OK) {
        std::string m = std::string(what) + ": " + vpmc_status_string(st);
        if (st == VPMC_ERR_HIP)
            m += std::string(" (") + hipGetErrorString((hipError_t)vpmc_last_hip_error()) + ")";
        throw std::runtime_error(m);
    }
}

// growable device scratch owned by the C++ layer (the C ABI never allocates)
struct Scratch {
    void *p = nullptr;
    size_t bytes = 0;
    void *get(size_t b)
    {
        if (b > bytes) {
            if (p) (void)hipFree(p);
            p = nullptr;
            vpmc_detail::hip_check(hipMalloc(&p, b), "scratch hipMalloc");
            bytes = b;
        }
        return p;
    }
    Scratch() = default;
    Scratch(const Scratch &) = delete;
    ~Scratch()
    {
        if (p) (void)hipFree(p);
    }
};

Scratch &free_fn_scratch()
{
    static thread_local Scratch s;
    return s;
}

// CountingSort_cuda's reference signature carries no cell count; the count of
// the grid last filled into a cell_num buffer is remembered here.
struct CellsMemo {
    const int *cell_num = nullptr;
    int cells = 0;
};
CellsMemo &cells_memo()
{
    static thread_local CellsMemo m;
    return m;
}

int log2_ceil(int n)
{
    int m = 0;
    while ((1 << m) < n) ++m;
    return m;
}

}  // namespace

// ---------------------------------------------------------------------------
// spatial_hash.hpp

void FillCells_cuda(int rshx, int rshy, float Lx, float Ly, int N, const int *sub,
                    const float2 *pos, int *cell_num, int *cell_acc, int *vp_cell)
{
    size_t b = vpmc_hash_workspace_bytes(N, rshx * rshy);
    void *ws = free_fn_scratch().get(b);
    vcheck(vpmc_fill_cells(rshx, rshy, Lx, Ly, N, sub, (const float *)pos, cell_num, cell_acc,
                           vp_cell, ws, b, nullptr),
           "FillCells_cuda");
    cells_memo() = {cell_num, rshx * rshy};
    vpmc_detail::hip_check(hipDeviceSynchronize(), "FillCells_cuda");
}

void CountingSort_cuda(int N, int *cell_num, const int *cell_acc, const int *vp_cell, int *vp_map)
{
    if (N <= 0) return;
    int cells = 0;
    if (cells_memo().cell_num == cell_num) {
        cells = cells_memo().cells;
    } else {   // any count above the largest key bounds every cell correctly
        std::vector<int> keys(N);
        vpmc_detail::hip_check(hipMemcpy(keys.data(), vp_cell, N * sizeof(int),
                                         hipMemcpyDeviceToHost),
                               "CountingSort_cuda");
        cells = *std::max_element(keys.begin(), keys.end()) + 1;
    }
    size_t b = vpmc_hash_workspace_bytes(N, cells);
    void *ws = free_fn_scratch().get(b);
    vcheck(vpmc_counting_sort(N, cells, cell_num, cell_acc, vp_cell, vp_map, ws, b, nullptr),
           "CountingSort_cuda");
    vpmc_detail::hip_check(hipDeviceSynchronize(), "CountingSort_cuda");
}

// the host-span variants: the same kernels on copies of the spans
void FillCells_cpu(int rshx, int rshy, float Lx, float Ly, int N, std::span<const int> sub,
                   std::span<const float> px, std::span<const float> py,
                   std::span<int> cell_num, std::span<int> cell_acc, std::span<int> vp_cell)
{
    const size_t cells = (size_t)rshx * rshy;
    if (N < 0 || rshx <= 0 || rshy <= 0 || sub.size() < (size_t)N || px.size() != py.size() ||
        cell_num.size() < cells || cell_acc.size() < cells || vp_cell.size() < (size_t)N)
        throw std::invalid_argument("FillCells_cpu: span sizes");
    for (int i = 0; i < N; ++i)
        if (sub[i] < 0 || (size_t)sub[i] >= px.size())
            throw std::invalid_argument("FillCells_cpu: sub index outside the positions");
    std::vector<float2> xy(px.size());
    for (size_t j = 0; j < px.size(); ++j) xy[j] = make_float2(px[j], py[j]);
    dvec<float2> dxy(xy);
    dvec<int> dsub(std::vector<int>(sub.begin(), sub.begin() + N)), num(cells), acc(cells),
        key(std::max(N, 1));
    FillCells_cuda(rshx, rshy, Lx, Ly, N, raw_ptr(dsub), raw_ptr(dxy), raw_ptr(num), raw_ptr(acc),
                   raw_ptr(key));
    const std::vector<int> hn = num.to_host(), ha = acc.to_host(), hk = key.to_host();
    std::copy(hn.begin(), hn.end(), cell_num.begin());
    std::copy(ha.begin(), ha.end(), cell_acc.begin());
    std::copy(hk.begin(), hk.begin() + N, vp_cell.begin());
}

void FillCells_omp(int rshx, int rshy, float Lx, float Ly, int N, std::span<const int> sub,
                   std::span<const float> px, std::span<const float> py,
                   std::span<int> cell_num, std::span<int> cell_acc, std::span<int> vp_cell)
{
    FillCells_cpu(rshx, rshy, Lx, Ly, N, sub, px, py, cell_num, cell_acc, vp_cell);
}

void CountingSort_cpu(int N, std::span<int> cell_num, std::span<const int> cell_acc,
                      std::span<const int> vp_cell, std::span<int> vp_map)
{
    const size_t cells = cell_num.size();
    if (N < 0 || cell_acc.size() < cells || vp_cell.size() < (size_t)N ||
        vp_map.size() < (size_t)N)
        throw std::invalid_argument("CountingSort_cpu: span sizes");
    if (N == 0) return;
    for (int i = 0; i < N; ++i)
        if (vp_cell[i] < 0 || (size_t)vp_cell[i] >= cells)
            throw std::invalid_argument("CountingSort_cpu: cell key outside cell_num");
    dvec<int> num(std::vector<int>(cell_num.begin(), cell_num.end())),
        acc(std::vector<int>(cell_acc.begin(), cell_acc.begin() + cells)),
        key(std::vector<int>(vp_cell.begin(), vp_cell.begin() + N)), map(N);
    cells_memo() = {raw_ptr(num), (int)cells};
    CountingSort_cuda(N, raw_ptr(num), raw_ptr(acc), raw_ptr(key), raw_ptr(map));
    const std::vector<int> hn = num.to_host(), hm = map.to_host();
    std::copy(hn.begin(), hn.end(), cell_num.begin());
    std::copy(hm.begin(), hm.end(), vp_map.begin());
}

void CountingSort_omp(int N, std::span<int> cell_num, std::span<const int> cell_acc,
                      std::span<const int> vp_cell, std::span<int> vp_map)
{
    CountingSort_cpu(N, cell_num, cell_acc, vp_cell, vp_map);
}

// ---------------------------------------------------------------------------
// utilities.hpp -- src/utilities.cu:12-96 (host-built tables)

void build_cubic_rcdf(float h_w, int length, dvec<float> &cdf)
{
    std::vector<float> h(length);
    const float h_a = h_w / length;
    float acc = 0.f;
    for (int i = 0; i < length; ++i) {
        float x = (0.5f + i) * h_a;
        acc += cubicSplinePDF(x, h_w) * x;
        h[i] = acc;
    }
    float total = h.back();
    for (float &v : h) v /= total;
    cdf.assign_host(h.data(), h.size());
}

void cdf_to_icdf(const dvec<float> &cdf, float h_w, dvec<float> &icdf)
{
    std::vector<float> c = cdf.to_host();
    const int L = (int)c.size();
    const float dh = 1.f / L;
    std::vector<float> out(L);
    auto lookup = [&](float u) {
        int p = 0;
        while (p < L && u > c[p]) ++p;
        if (p == 0) return (1 - (c[0] - u) / c[0]) * 0.5f;
        if (p == L) return (1 - (1 - u) / (1 - c[L - 1])) * (L - 0.5f);
        float lam = (c[p] - u) / (c[p] - c[p - 1]);
        return (1 - lam) * (p + 0.5f) + lam * (p - 0.5f);
    };
    for (int i = 0; i < L; ++i) out[i] = lookup((i + 0.5f) * dh) * dh * h_w;
    icdf.assign_host(out.data(), out.size());
}

// ---------------------------------------------------------------------------
// monte_carlo_bs_cpu on the GPU (src/cpu/monte_carlo_bs.cpp:28-128)
//
// One call of the reference takes from its generator N1 * b1 + N2 * 2 nt
// draws (b1 = 3 periodic, 1 otherwise; nt = 5 targets with auxiliary, else 1):
// per branch-1 sample the CDF draw then the two image draws, per branch-2
// sample two draws for each target's disk in the order dst, left, right,
// top, bottom.  Branch-1 samples are shared by the targets (the same point and
// vorticity enter mis_contrib at each), so the call equals nt owners of the
// reference-mode estimator with the branch-1 draws repeated and each owner's
// own disk draws: a replay stream with two draws per branch-2 sample.

namespace {

struct ReplaySet {   // one sign set as the estimator reads it
    int n = 0, nx = 0, ny = 0, glog2 = 0;
    const double *cdf = nullptr;
    dvec<double> cdf_wide;   // an FP32 CDF widened to FP64 (exact)
    dvec<int> count;
    dvec<double> total;
    dvec<int> guide, cell_start;
    dvec<float4> pts;
    dvec<float2> set_xy;
};

size_t mcbs_call_draws(int N1, int N2, bool periodic, bool aux)
{
    return (size_t)N1 * (periodic ? 3 : 1) + (size_t)N2 * 2 * (aux ? 5 : 1);
}

void mcbs_draw(std::mt19937 &rng, size_t len, float *out)
{
    std::uniform_real_distribution<float> U(0.f, 1.f);
    for (size_t k = 0; k < len; ++k) out[k] = U(rng);
}

}  // namespace

static void replay_set_build(ReplaySet &S, const MCBSCtx &C, int sign, const ParticleSet &P,
                             hipStream_t st)
{
    const bool pos = sign > 0;
    const dvec<int> *sub = pos ? C.sub_pos : C.sub_neg;
    // the FP64 CDF of the added members, else the reference's FP32 one
    const dvec<double> *cdf64 = pos ? C.cdf_pos64 : C.cdf_neg64;
    const dvec<float> *cdf32 = pos ? C.cdf_pos : C.cdf_neg;
    if (!sub || (!cdf64 && !cdf32))
        throw std::invalid_argument("monte_carlo_bs_cpu: MCBSCtx needs the sign's sub and cdf");
    const size_t cdf_size = cdf64 ? cdf64->size() : cdf32->size();
    int n = pos ? C.cnt_pos : C.cnt_neg;
    if (n < 0) n = (int)cdf_size;
    if ((size_t)n > cdf_size || (size_t)n > sub->size())
        throw std::invalid_argument("monte_carlo_bs_cpu: subset count beyond its buffers");
    if (!(C.h_w > 0.f) || !(C.Lx > 0.f) || !(C.Ly > 0.f))
        throw std::invalid_argument("monte_carlo_bs_cpu: h_w, Lx, Ly must be positive");
    vcheck(vpmc_field_grid(C.h_w, C.Lx, C.Ly, 8, &S.nx, &S.ny), "monte_carlo_bs_cpu");
    S.n = n;
    double cum;
    if (cdf64) {
        S.cdf = raw_ptr(*cdf64);
        cum = pos ? C.cum_pos64 : C.cum_neg64;
    } else {   // every float is a double: the same comparisons as the reference's
        const std::vector<float> h = cdf32->to_host();
        const std::vector<double> wide(h.begin(), h.begin() + n);
        if (n > 0) S.cdf_wide.assign_host(wide.data(), wide.size());
        else S.cdf_wide.resize(1);
        S.cdf = raw_ptr(S.cdf_wide);
        cum = pos ? C.cum_pos : C.cum_neg;
    }
    S.count.assign_host(&n, 1);
    S.total.assign_host(&cum, 1);
    int gl = 0;
    while ((1 << gl) < std::max(n, 1)) ++gl;
    S.glog2 = gl;
    S.guide.resize((size_t(1) << gl) + 1);
    S.pts.resize(std::max(n, 1));
    S.set_xy.resize(std::max(n, 1));
    S.cell_start.resize((size_t)S.nx * S.ny + 1);
    const size_t b = vpmc_field_workspace_bytes(n, S.nx, S.ny);
    vcheck(vpmc_field_build(S.nx, S.ny, C.Lx, C.Ly, n, raw_ptr(*sub), (const float *)P.d_p(),
                            P.d_omega(), 1, (float *)raw_ptr(S.pts), raw_ptr(S.cell_start),
                            nullptr, (float *)raw_ptr(S.set_xy), free_fn_scratch().get(b), b, st),
           "monte_carlo_bs_cpu");
    if (n > 0)
        vcheck(vpmc_build_guide(S.cdf, raw_ptr(S.count), gl, raw_ptr(S.guide), st),
               "monte_carlo_bs_cpu");
}

// Groups of nt targets (owner g * nt + t); u[s] holds, per group, the draws of
// the monte_carlo_bs_cpu call for sign slot s (0: set[0], 1: set[1]).  Returns
// vel_sign: [slot][owner].
static std::vector<float2> replay_estimate(const std::vector<float2> &targets, int nt,
                                           const ReplaySet *S, const std::vector<float> *u,
                                           int N1, int N2, bool periodic, float h_w, float Lx,
                                           float Ly, hipStream_t st)
{
    const int n_owner = (int)targets.size();
    const size_t G = (size_t)n_owner / nt;
    const size_t n1 = (size_t)N1 * (periodic ? 3 : 1), per = n1 + 2 * (size_t)N2;
    const size_t len = n1 + 2 * (size_t)N2 * nt;
    std::vector<float> U((size_t)n_owner * 2 * per);
    for (size_t g = 0; g < G; ++g)
        for (int t = 0; t < nt; ++t)
            for (int s = 0; s < 2; ++s) {
                const float *src = u[s].data() + g * len;
                float *dst = U.data() + ((g * nt + t) * 2 + s) * per;
                std::copy(src, src + n1, dst);
                for (int k = 0; k < N2; ++k) {
                    dst[n1 + 2 * (size_t)k] = src[n1 + (size_t)k * 2 * nt + 2 * t];
                    dst[n1 + 2 * (size_t)k + 1] = src[n1 + (size_t)k * 2 * nt + 2 * t + 1];
                }
            }
    dvec<float> dU(U);
    dvec<float2> dT(targets), vs(2 * (size_t)n_owner);
    vpmc_mc_desc d{};
    d.mode = VPMC_MODE_REFERENCE;
    d.N1 = N1;
    d.N2 = N2;
    d.h_w = h_w;
    d.Lx = Lx;
    d.Ly = Ly;
    d.periodic = periodic ? 1 : 0;
    d.uniforms = raw_ptr(dU);
    d.n_owner = n_owner;
    d.owner_xy = (const float *)raw_ptr(dT);
    for (int s = 0; s < 2; ++s) {
        const ReplaySet &R = S[s];
        vpmc_sign_set &q = d.set[s];
        q.n = R.n;
        q.count_dev = raw_ptr(R.count);
        q.cdf = R.cdf;
        q.total_dev = raw_ptr(R.total);
        q.guide = raw_ptr(R.guide);
        q.guide_log2 = R.glog2;
        q.set_xy = (const float *)raw_ptr(R.set_xy);
        q.field = vpmc_field{R.nx, R.ny, Lx, Ly, R.n, (const float *)raw_ptr(R.pts),
                             raw_ptr(R.cell_start), nullptr};
    }
    d.vel_sign = (float *)raw_ptr(vs);
    d.algo = VPMC_ALGO_DIRECT;
    vcheck(vpmc_mc_estimate(&d, st), "monte_carlo_bs_cpu");
    vpmc_detail::hip_check(hipStreamSynchronize(st), "monte_carlo_bs_cpu");
    return vs.to_host();
}

// ---------------------------------------------------------------------------
// Simulation

struct Simulation::Impl {
    int nx = 0, ny = 0, glog2 = 0;
    dvec<int> counts;         // [2] device subset sizes
    dvec<double> totals;      // [2] device sum |w|
    dvec<int> guide[2];
    dvec<float4> pts[2];
    dvec<int> cell_start[2], order[2];
    dvec<float2> set_xy[2];
    dvec<float2> vel_sign;
    dvec<float4> gq[2], set_jac[2];   // kernel deformation: forms and F per set
    dvec<float2> vel_aux;             // [sign][L, R, T, B][owner]
    Scratch ws;
    Impl() = default;
    Impl(const Impl &o)   // device buffers deep-copied; scratch is not state
        : nx(o.nx), ny(o.ny), glog2(o.glog2), counts(o.counts), totals(o.totals),
          guide{o.guide[0], o.guide[1]}, pts{o.pts[0], o.pts[1]},
          cell_start{o.cell_start[0], o.cell_start[1]}, order{o.order[0], o.order[1]},
          set_xy{o.set_xy[0], o.set_xy[1]}, vel_sign(o.vel_sign), gq{o.gq[0], o.gq[1]},
          set_jac{o.set_jac[0], o.set_jac[1]}, vel_aux(o.vel_aux)
    {
    }
};

Simulation::Simulation(SimParam p) : P_(std::move(p)), impl_(new Impl) {}
Simulation::~Simulation() { delete impl_; }

Simulation::Simulation(const Simulation &o) : P_(o.P_), impl_(new Impl) { *this = o; }

Simulation &Simulation::operator=(const Simulation &o)
{
    if (this == &o) return *this;
    particles_ = o.particles_;
    P_ = o.P_;
    rng = o.rng;
    mode = o.mode;
    seed = o.seed;
    sim_id = o.sim_id;
    step_index = o.step_index;
    cells_per_hw = o.cells_per_hw;
    kernel_deformation = o.kernel_deformation;
    deformation_bound = o.deformation_bound;
    stream = o.stream;
    cg_iterations = o.cg_iterations;
    cg_rel_residual = o.cg_rel_residual;
    hash_pos_ = o.hash_pos_;
    hash_neg_ = o.hash_neg_;
    hash_all_ = o.hash_all_;
    sub_all_ = o.sub_all_;
    sub_pos_ = o.sub_pos_;
    sub_neg_ = o.sub_neg_;
    cdf_pos_ = o.cdf_pos_;
    cdf_neg_ = o.cdf_neg_;
    cdf_pos_f32_ = o.cdf_pos_f32_;
    cdf_neg_f32_ = o.cdf_neg_f32_;
    cnt_pos_ = o.cnt_pos_;
    cnt_neg_ = o.cnt_neg_;
    cum_pos_ = o.cum_pos_;
    cum_neg_ = o.cum_neg_;
    cdf_kernel = o.cdf_kernel;
    icdf_kernel = o.icdf_kernel;
    icdf_kernel_ = o.icdf_kernel_;
    icdf_L_ = o.icdf_L_;
    if (impl_ != o.impl_) {
        Impl *n = new Impl(*o.impl_);
        delete impl_;
        impl_ = n;
    }
    // views onto this copy's own hash buffers
    pos_view = hash_pos_.vp_sort.size() ? makePosView() : VorticityView{};
    neg_view = hash_neg_.vp_sort.size() ? makeNegView() : VorticityView{};
    return *this;
}

void Simulation::init(int num_of_p)
{
    // src/vortex_particle_mc.cu:77-120
    if (!(num_of_p > 0 && num_of_p <= P_.N_max))
        throw std::invalid_argument("Simulation::init: 0 < num_of_p <= N_max required");
    std::vector<float2> host_pos(num_of_p);
    std::vector<float> host_omega(num_of_p);
    std::mt19937 gen(42);
    std::uniform_real_distribution<float> uni(0.f, 1.f);
    ensure_icdf_kernel(20);
    for (int i = 0; i < num_of_p; ++i) {
        float2 p{uni(gen), uni(gen)};
        host_pos[i] = p;
        host_omega[i] = taylor_green_vorticity(p);
    }
    particles_.resize(P_.N_max);
    particles_.set_n_current(num_of_p);
    vpmc_detail::hip_check(hipMemcpy(particles_.d_p(), host_pos.data(),
                                     num_of_p * sizeof(float2), hipMemcpyHostToDevice),
                           "init");
    std::vector<float4> F(P_.N_max, make_float4(1.f, 0.f, 0.f, 1.f));
    particles_.F.assign_host(F.data(), F.size());
    std::vector<float> one(P_.N_max, 1.f);
    particles_.jac.assign_host(one.data(), one.size());
    prepare_rhs_b();
    conjugate_gradient_rbf(*this);
}

void Simulation::init_with(const std::vector<float2> &pos, const std::vector<float> &b, float tol,
                           int max_iter)
{
    if (pos.size() != b.size() || pos.empty())
        throw std::invalid_argument("Simulation::init_with: sizes");
    int n = (int)pos.size();
    if (P_.N_max < n) P_.N_max = n;
    particles_.resize(n);
    particles_.pos.assign_host(pos.data(), n);
    particles_.omega_field.assign_host(b.data(), n);
    std::vector<float4> F(n, make_float4(1.f, 0.f, 0.f, 1.f));
    particles_.F.assign_host(F.data(), n);
    conjugate_gradient_rbf(*this, tol, max_iter);
}

void Simulation::set_particles(const std::vector<float2> &pos, const std::vector<float> &omega)
{
    if (pos.size() != omega.size() || pos.empty())
        throw std::invalid_argument("Simulation::set_particles: sizes");
    int n = (int)pos.size();
    if (P_.N_max < n) P_.N_max = n;
    particles_.resize(n);
    particles_.pos.assign_host(pos.data(), n);
    particles_.omega.assign_host(omega.data(), n);
    std::vector<float4> F(n, make_float4(1.f, 0.f, 0.f, 1.f));
    particles_.F.assign_host(F.data(), n);
    build_subsets_cuda();
    do_spatial_hashing();
}

void Simulation::reinitialize(int nx_l, int ny_l, float tol, int max_iter)
{
    // README.md:319-321, 416-420: positions to a uniform grid, F to I,
    // strengths re-fitted to the current field
    if (nx_l <= 0) nx_l = P_.ResolutionX;
    if (ny_l <= 0) ny_l = P_.ResolutionY;
    const int N = particles_.N_cur;
    const int n_new = nx_l * ny_l;
    const float *jac = nullptr;
    if (kernel_deformation && particles_.F.size() >= (size_t)N)
        jac = (const float *)raw_ptr(particles_.F);
    dvec<float2> pos_out(n_new);
    dvec<float> w_out(n_new);
    size_t b = vpmc_reinit_workspace_bytes(N, nx_l, ny_l, P_.h_w, P_.Lx, P_.Ly, cells_per_hw,
                                           jac != nullptr);
    Impl &I = *impl_;
    int it = 0;
    double rr = 0.0;
    int st = vpmc_reinit(N, (const float *)particles_.d_p(), particles_.d_omega(), jac, P_.h_w,
                         P_.Lx, P_.Ly, cells_per_hw, nx_l, ny_l, tol, max_iter,
                         (float *)raw_ptr(pos_out), raw_ptr(w_out), &it, &rr, I.ws.get(b), b,
                         stream);
    if (st != VPMC_OK && st != VPMC_ERR_NOT_CONVERGED) vcheck(st, "reinitialize");
    cg_iterations = it;
    cg_rel_residual = rr;
    set_particles(pos_out.to_host(), w_out.to_host());   // F = I, subsets, fields
}

void Simulation::prepare_rhs_b()
{
    // src/vortex_particle_mc.cu:237-260: b_i = omega_TG(x_i)
    int N = particles_.N_cur;
    std::vector<float2> p(N);
    vpmc_detail::hip_check(hipMemcpy(p.data(), particles_.d_p(), N * sizeof(float2),
                                     hipMemcpyDeviceToHost),
                           "prepare_rhs_b");
    std::vector<float> b(N);
    for (int i = 0; i < N; ++i) b[i] = taylor_green_vorticity(p[i]);
    if (particles_.omega_field.size() < (size_t)N) particles_.omega_field.resize(N);
    vpmc_detail::hip_check(hipMemcpy(particles_.d_omega_field(), b.data(), N * sizeof(float),
                                     hipMemcpyHostToDevice),
                           "prepare_rhs_b");
}

void Simulation::build_subsets_cpu()
{
    // src/cpu/build_particle_sets.cpp:6-42: the same subsets (ascending index),
    // the CDFs and totals in FP64
    build_subsets_cuda();
}

void Simulation::build_subsets_cuda()
{
    const int N = particles_.N_cur;
    sub_pos_.resize(N);
    sub_neg_.resize(N);
    cdf_pos_.resize(N);
    cdf_neg_.resize(N);
    impl_->counts.resize(2);
    impl_->totals.resize(2);
    size_t b = vpmc_subsets_workspace_bytes(N);
    vcheck(vpmc_build_subsets(N, particles_.d_omega(), raw_ptr(sub_pos_), raw_ptr(sub_neg_),
                              raw_ptr(cdf_pos_), raw_ptr(cdf_neg_), raw_ptr(impl_->counts),
                              raw_ptr(impl_->totals), impl_->ws.get(b), b, stream),
           "build_subsets");
    impl_->glog2 = log2_ceil(std::max(N, 1));
    for (int s = 0; s < 2; ++s) {
        impl_->guide[s].resize((size_t(1) << impl_->glog2) + 1);
        vcheck(vpmc_build_guide(s ? raw_ptr(cdf_neg_) : raw_ptr(cdf_pos_),
                                raw_ptr(impl_->counts) + s, impl_->glog2,
                                raw_ptr(impl_->guide[s]), stream),
               "build_guide");
    }
    std::vector<int> c = impl_->counts.to_host();
    std::vector<double> t = impl_->totals.to_host();
    cnt_pos_ = c[0];
    cnt_neg_ = c[1];
    cum_pos_ = t[0];
    cum_neg_ = t[1];
    // the reference's FP32 CDF vectors (MCBSCtx::cdf_pos / cdf_neg), one entry
    // per subset member
    for (int s = 0; s < 2; ++s) {
        const std::vector<double> h = (s ? cdf_neg_ : cdf_pos_).to_host();
        const std::vector<float> f(h.begin(), h.begin() + c[s]);
        (s ? cdf_neg_f32_ : cdf_pos_f32_).assign_host(f.data(), f.size());
    }
}

void Simulation::do_spatial_hashing()
{
    // reference grid (src/vortex_particle_mc.cu:123-210) for the views ...
    int rshx = P_.ResolutionX / int(P_.hwr) / 2;
    int rshy = P_.ResolutionY / int(P_.hwr) / 2;
    hash_pos_.resize_cells(rshx, rshy);
    hash_neg_.resize_cells(rshx, rshy);
    hash_pos_.resize_particles(std::max(cnt_pos_, 1));
    hash_neg_.resize_particles(std::max(cnt_neg_, 1));
    FillCells_cuda(rshx, rshy, P_.Lx, P_.Ly, cnt_pos_, raw_ptr(sub_pos_), particles_.d_p(),
                   raw_ptr(hash_pos_.cell_num), raw_ptr(hash_pos_.cell_acc),
                   raw_ptr(hash_pos_.vp_cell));
    CountingSort_cuda(cnt_pos_, raw_ptr(hash_pos_.cell_num), raw_ptr(hash_pos_.cell_acc),
                      raw_ptr(hash_pos_.vp_cell), raw_ptr(hash_pos_.vp_sort));
    FillCells_cuda(rshx, rshy, P_.Lx, P_.Ly, cnt_neg_, raw_ptr(sub_neg_), particles_.d_p(),
                   raw_ptr(hash_neg_.cell_num), raw_ptr(hash_neg_.cell_acc),
                   raw_ptr(hash_neg_.vp_cell));
    CountingSort_cuda(cnt_neg_, raw_ptr(hash_neg_.cell_num), raw_ptr(hash_neg_.cell_acc),
                      raw_ptr(hash_neg_.vp_cell), raw_ptr(hash_neg_.vp_sort));
    pos_view = makePosView();
    neg_view = makeNegView();
    // ... and the estimator's fine grid
    Impl &I = *impl_;
    vcheck(vpmc_field_grid(P_.h_w, P_.Lx, P_.Ly, cells_per_hw, &I.nx, &I.ny), "field_grid");
    const int cnt[2] = {cnt_pos_, cnt_neg_};
    size_t b = vpmc_field_workspace_bytes(std::max(cnt_pos_, cnt_neg_), I.nx, I.ny);
    if (kernel_deformation && particles_.F.size() < (size_t)particles_.N_cur) {
        std::vector<float4> F(particles_.N_cur, make_float4(1.f, 0.f, 0.f, 1.f));
        particles_.F.assign_host(F.data(), F.size());
    }
    for (int s = 0; s < 2; ++s) {
        I.pts[s].resize(std::max(cnt[s], 1));
        I.order[s].resize(std::max(cnt[s], 1));
        I.set_xy[s].resize(std::max(cnt[s], 1));
        I.cell_start[s].resize((size_t)I.nx * I.ny + 1);
        if (kernel_deformation) {   // F in the query (include/vorticity_query.cuh:52-66)
            I.gq[s].resize(std::max(cnt[s], 1));
            I.set_jac[s].resize(std::max(cnt[s], 1));
            vcheck(vpmc_field_build_jac(
                       I.nx, I.ny, P_.Lx, P_.Ly, cnt[s], s ? raw_ptr(sub_neg_) : raw_ptr(sub_pos_),
                       (const float *)particles_.d_p(), particles_.d_omega(), 1,
                       (const float *)raw_ptr(particles_.F), (float *)raw_ptr(I.pts[s]),
                       (float *)raw_ptr(I.gq[s]), raw_ptr(I.cell_start[s]), raw_ptr(I.order[s]),
                       (float *)raw_ptr(I.set_xy[s]), (float *)raw_ptr(I.set_jac[s]),
                       I.ws.get(b), b, stream),
                   "field_build");
            continue;
        }
        vcheck(vpmc_field_build(I.nx, I.ny, P_.Lx, P_.Ly, cnt[s],
                                s ? raw_ptr(sub_neg_) : raw_ptr(sub_pos_),
                                (const float *)particles_.d_p(), particles_.d_omega(), 1,
                                (float *)raw_ptr(I.pts[s]), raw_ptr(I.cell_start[s]),
                                raw_ptr(I.order[s]), (float *)raw_ptr(I.set_xy[s]), I.ws.get(b),
                                b, stream),
               "field_build");
    }
}

void Simulation::build_global_hash()
{
    // src/vortex_particle_mc.cu:212-235
    const int N = particles_.N_cur;
    std::vector<int> iota(N);
    std::iota(iota.begin(), iota.end(), 0);
    sub_all_.assign_host(iota.data(), N);
    const int rshx = P_.ResolutionX / int(P_.hwr) / 2;
    const int rshy = P_.ResolutionY / int(P_.hwr) / 2;
    hash_all_.resize_cells(rshx, rshy);
    hash_all_.resize_particles(N);
    FillCells_cuda(rshx, rshy, P_.Lx, P_.Ly, N, raw_ptr(sub_all_), particles_.d_p(),
                   raw_ptr(hash_all_.cell_num), raw_ptr(hash_all_.cell_acc),
                   raw_ptr(hash_all_.vp_cell));
    CountingSort_cuda(N, raw_ptr(hash_all_.cell_num), raw_ptr(hash_all_.cell_acc),
                      raw_ptr(hash_all_.vp_cell), raw_ptr(hash_all_.vp_sort));
}

static vpmc_mc_desc make_desc(Simulation &S, int N1, int N2, int cnt_pos, int cnt_neg,
                              const int *counts, const double *totals, dvec<int> *guide,
                              int glog2, const double *cdf_pos, const double *cdf_neg,
                              dvec<float4> *pts, dvec<int> *cell_start, dvec<int> *order,
                              dvec<float2> *set_xy, int nx, int ny, float2 *vel_sign,
                              bool periodic)
{
    vpmc_mc_desc d{};
    d.mode = (int)S.mode;
    d.N1 = N1;
    d.N2 = N2;
    d.h_w = S.P_.h_w;
    d.Lx = S.P_.Lx;
    d.Ly = S.P_.Ly;
    d.periodic = periodic ? 1 : 0;
    d.seed = S.seed;
    d.step = S.step_index;
    d.sim = S.sim_id;
    d.n_owner = S.particles_.N_cur;
    d.owner_xy = (const float *)S.particles_.d_p();
    const int cnt[2] = {cnt_pos, cnt_neg};
    const double *cdf[2] = {cdf_pos, cdf_neg};
    for (int s = 0; s < 2; ++s) {
        vpmc_sign_set &q = d.set[s];
        q.n = cnt[s];
        q.count_dev = counts + s;
        q.cdf = cdf[s];
        q.total_dev = totals + s;
        q.guide = raw_ptr(guide[s]);
        q.guide_log2 = glog2;
        q.set_xy = (const float *)raw_ptr(set_xy[s]);
        q.field = vpmc_field{nx, ny, S.P_.Lx, S.P_.Ly, cnt[s], (const float *)raw_ptr(pts[s]),
                             raw_ptr(cell_start[s]), raw_ptr(order[s])};
    }
    d.vel_sign = (float *)vel_sign;
    // the binned estimator's cell stencil wraps periodically; a bounded box
    // takes the direct estimator (one workgroup per owner and sign)
    d.algo = periodic ? VPMC_ALGO_BINNED : VPMC_ALGO_DIRECT;
    return d;
}

// the estimate of descriptor d with the algorithm make_desc chose
static void run_estimate(const vpmc_mc_desc &d, Scratch &ws, hipStream_t stream,
                         const char *what)
{
    if (d.algo == VPMC_ALGO_BINNED) {
        const size_t wb = vpmc_mc_binned_workspace_bytes(&d);
        vcheck(vpmc_mc_estimate_binned(&d, ws.get(wb), wb, stream), what);
    } else {
        vcheck(vpmc_mc_estimate(&d, stream), what);
    }
}

// the CUDA step's F in the fields and the jitter (src/cuda/monte_carlo_bs.cu:210-212)
static void attach_deformation(vpmc_mc_desc &d, Simulation::Impl &I, float bound)
{
    for (int s = 0; s < 2; ++s) {
        d.set[s].field.gq = (const float *)raw_ptr(I.gq[s]);
        d.set[s].field.gq_bound = bound;
        d.set[s].set_jac = (const float *)raw_ptr(I.set_jac[s]);
    }
}

void Simulation::estimate_velocity(bool periodic)
{
    Impl &I = *impl_;
    const int N = particles_.N_cur;
    do_spatial_hashing();
    I.vel_sign.resize(2 * (size_t)N);
    vpmc_mc_desc d = make_desc(*this, P_.N1, P_.N2, cnt_pos_, cnt_neg_, raw_ptr(I.counts),
                               raw_ptr(I.totals), I.guide, I.glog2, raw_ptr(cdf_pos_),
                               raw_ptr(cdf_neg_), I.pts, I.cell_start, I.order, I.set_xy, I.nx,
                               I.ny, raw_ptr(I.vel_sign), periodic);
    if (kernel_deformation) {
        if (!periodic)
            throw std::invalid_argument(
                "estimate_velocity: kernel deformation needs a periodic box (binned estimator)");
        attach_deformation(d, I, deformation_bound);
    }
    run_estimate(d, I.ws, stream, "monte_carlo_bs");
    // vel = v+ - v- (dt = 0 leaves the positions unchanged)
    vcheck(vpmc_advect(N, (float *)particles_.d_p(), (const float *)raw_ptr(I.vel_sign),
                       (float *)particles_.d_v(), 0.f, P_.Lx, P_.Ly, periodic ? 1 : 0, stream),
           "combine");
    vpmc_detail::hip_check(hipStreamSynchronize(stream), "estimate_velocity");
}

void Simulation::step_cpu(bool periodic)
{
    const int N = particles_.N_cur;
    if (N <= 0) return;
    if (cdf_pos_.size() != (size_t)N) build_subsets_cuda();
    const MCBSCtx C = make_mcbs_ctx(*this, periodic);
    ReplaySet S[2];
    replay_set_build(S[0], C, +1, particles_, stream);
    replay_set_build(S[1], C, -1, particles_, stream);
    // targets: particle i's group is (pos, left, right, top, bottom)
    dvec<float2> *arr[5] = {&particles_.pos, &particles_.pos_left, &particles_.pos_right,
                            &particles_.pos_top, &particles_.pos_bottom};
    std::vector<float2> hpos[5];
    for (int t = 0; t < 5; ++t) hpos[t] = arr[t]->to_host();
    // each call's draws (they do not depend on the estimates) in the
    // reference's order: per particle the +1 call, then the -1 call; replayed
    // in chunks of particles that keep the replay stream near 64M draws
    std::mt19937 g = rng;
    const size_t len = mcbs_call_draws(P_.N1, P_.N2, periodic, true);
    const size_t per = (size_t)P_.N1 * (periodic ? 3 : 1) + 2 * (size_t)P_.N2;
    const int chunk = (int)std::max<size_t>(1, ((size_t)1 << 26) / (10 * per));
    std::vector<float2> vp((size_t)5 * N), vn((size_t)5 * N);   // [particle][target]
    for (int i0 = 0; i0 < N; i0 += chunk) {
        const int m = std::min(chunk, N - i0);
        std::vector<float> u[2];
        u[0].resize((size_t)m * len);
        u[1].resize((size_t)m * len);
        for (int i = 0; i < m; ++i)
            for (int s = 0; s < 2; ++s) mcbs_draw(g, len, u[s].data() + (size_t)i * len);
        std::vector<float2> tg((size_t)5 * m);
        for (int i = 0; i < m; ++i)
            for (int t = 0; t < 5; ++t) tg[(size_t)i * 5 + t] = hpos[t][i0 + i];
        const std::vector<float2> v = replay_estimate(tg, 5, S, u, P_.N1, P_.N2, periodic,
                                                      P_.h_w, P_.Lx, P_.Ly, stream);
        std::copy(v.begin(), v.begin() + 5 * (size_t)m, vp.begin() + 5 * (size_t)i0);
        std::copy(v.begin() + 5 * (size_t)m, v.end(), vn.begin() + 5 * (size_t)i0);
    }
    // x += (v+ - v-) dt per target array, wrapped; vel = v+ - v- of the particles
    dvec<float2> vs(2 * (size_t)N), scratch_vel(N);
    for (int t = 0; t < 5; ++t) {
        std::vector<float2> h(2 * (size_t)N);
        for (int i = 0; i < N; ++i) {
            h[i] = vp[(size_t)i * 5 + t];
            h[(size_t)N + i] = vn[(size_t)i * 5 + t];
        }
        vs.assign_host(h.data(), h.size());
        vcheck(vpmc_advect(N, (float *)raw_ptr(*arr[t]), (const float *)raw_ptr(vs),
                           t == 0 ? (float *)particles_.d_v() : (float *)raw_ptr(scratch_vel),
                           P_.dt, P_.Lx, P_.Ly, periodic ? 1 : 0, stream),
               "step_cpu");
        vpmc_detail::hip_check(hipStreamSynchronize(stream), "step_cpu");
    }
}

void Simulation::step_cuda(bool periodic)
{
    // src/cuda/simulation_step.cu:9-90 with aux = 0: rebuild the sign fields
    // from the current positions, estimate both signs, forward Euler + wrap
    Impl &I = *impl_;
    const int N = particles_.N_cur;
    if (I.nx == 0) do_spatial_hashing();
    I.vel_sign.resize(2 * (size_t)N);
    vpmc_mc_desc d = make_desc(*this, P_.N1, P_.N2, cnt_pos_, cnt_neg_, raw_ptr(I.counts),
                               raw_ptr(I.totals), I.guide, I.glog2, raw_ptr(cdf_pos_),
                               raw_ptr(cdf_neg_), I.pts, I.cell_start, I.order, I.set_xy, I.nx,
                               I.ny, raw_ptr(I.vel_sign), periodic);
    if (kernel_deformation) {
        if (!periodic)
            throw std::invalid_argument(
                "step_cuda: kernel deformation needs a periodic box (binned estimator)");
        attach_deformation(d, I, deformation_bound);
        // aux = 1: estimates at dst -/+ dx, F = update_jacobian_from5 of them
        // before the move (src/cuda/simulation_step.cu:51-73)
        I.vel_aux.resize(8 * (size_t)N);
        d.aux = 1;
        d.aux_dx = P_.dx;
        d.vel_aux = (float *)raw_ptr(I.vel_aux);
        d.jac = (float *)raw_ptr(particles_.F);
    }
    size_t b = vpmc_step_workspace_bytes(&d);
    vcheck(vpmc_step(&d, N, (float *)particles_.d_p(), particles_.d_omega(), raw_ptr(sub_pos_),
                     raw_ptr(sub_neg_), (float *)particles_.d_v(), P_.dt, I.ws.get(b), b, stream),
           "step_cuda");
    ++step_index;
}

void Simulation::step_adjoint(uint32_t step, float2 *lambda, float *grad_w, float4 *jac_bar)
{
    // the adjoint of step_cuda(true) at Philox step `step`, from the state the
    // forward step started from (the current positions, and particles.F with
    // kernel deformation)
    Impl &I = *impl_;
    const int N = particles_.N_cur;
    if (I.nx == 0) do_spatial_hashing();
    I.vel_sign.resize(2 * (size_t)N);
    vpmc_mc_desc d = make_desc(*this, P_.N1, P_.N2, cnt_pos_, cnt_neg_, raw_ptr(I.counts),
                               raw_ptr(I.totals), I.guide, I.glog2, raw_ptr(cdf_pos_),
                               raw_ptr(cdf_neg_), I.pts, I.cell_start, I.order, I.set_xy, I.nx,
                               I.ny, raw_ptr(I.vel_sign), true);
    d.step = step;
    if (kernel_deformation) {   // F is state: dL/dF' in jac_bar becomes dL/dF
        if (!jac_bar)
            throw std::invalid_argument("step_adjoint: a deformed step needs jac_bar (dL/dF)");
        attach_deformation(d, I, deformation_bound);
        I.vel_aux.resize(8 * (size_t)N);
        d.aux = 1;
        d.aux_dx = P_.dx;
        d.vel_aux = (float *)raw_ptr(I.vel_aux);
        d.jac = (float *)raw_ptr(particles_.F);
        const size_t b = vpmc_step_adjoint_deformed_workspace_bytes(&d);
        if (b == 0)
            throw std::invalid_argument(
                "step_adjoint: the adjoint differentiates the consistent periodic estimator");
        vcheck(vpmc_step_adjoint_deformed(&d, N, (const float *)particles_.d_p(),
                                          particles_.d_omega(), raw_ptr(sub_pos_),
                                          raw_ptr(sub_neg_), P_.dt, (float *)lambda,
                                          (float *)jac_bar, grad_w, I.ws.get(b), b, stream),
               "step_adjoint");
        return;
    }
    const size_t b = vpmc_step_adjoint_workspace_bytes(&d);
    if (b == 0)
        throw std::invalid_argument(
            "step_adjoint: the adjoint differentiates the consistent periodic estimator");
    vcheck(vpmc_step_adjoint(&d, N, (const float *)particles_.d_p(), particles_.d_omega(),
                             raw_ptr(sub_pos_), raw_ptr(sub_neg_), P_.dt, (float *)lambda, grad_w,
                             I.ws.get(b), b, stream),
           "step_adjoint");
}

void Simulation::step_f64(uint32_t step, double2 *pos, const double *w, double4 *jac64)
{
    Impl &I = *impl_;
    const int N = particles_.N_cur;
    if (I.nx == 0) do_spatial_hashing();
    I.vel_sign.resize(2 * (size_t)N);
    vpmc_mc_desc d = make_desc(*this, P_.N1, P_.N2, cnt_pos_, cnt_neg_, raw_ptr(I.counts),
                               raw_ptr(I.totals), I.guide, I.glog2, raw_ptr(cdf_pos_),
                               raw_ptr(cdf_neg_), I.pts, I.cell_start, I.order, I.set_xy, I.nx,
                               I.ny, raw_ptr(I.vel_sign), true);
    d.step = step;
    if (kernel_deformation) {   // the FP64 F advances in jac64
        if (!jac64)
            throw std::invalid_argument("step_f64: a deformed step needs jac64 (F in FP64)");
        attach_deformation(d, I, deformation_bound);
        I.vel_aux.resize(8 * (size_t)N);   // the descriptor's aux buffer (FP64 sums stay apart)
        d.aux = 1;
        d.aux_dx = P_.dx;
        d.vel_aux = (float *)raw_ptr(I.vel_aux);
    }
    const size_t b = vpmc_step_f64_workspace_bytes(&d);
    if (b == 0)
        throw std::invalid_argument("step_f64: the consistent periodic estimator only");
    if (kernel_deformation)
        vcheck(vpmc_step_f64_deformed(&d, N, (double *)pos, w, particles_.d_omega(),
                                      raw_ptr(sub_pos_), raw_ptr(sub_neg_), (double)P_.dt,
                                      (double *)jac64, I.ws.get(b), b, stream),
               "step_f64");
    else
        vcheck(vpmc_step_f64(&d, N, (double *)pos, w, particles_.d_omega(), raw_ptr(sub_pos_),
                             raw_ptr(sub_neg_), (double)P_.dt, I.ws.get(b), b, stream),
               "step_f64");
}

void Simulation::compute_velocity_naive(bool periodic, float eps_core)
{
    // src/cuda/simulation_step.cu:92-108 on the FP64 direct-sum kernel
    const int N = particles_.N_cur;
    std::vector<float2> p = particles_.pos.to_host();
    std::vector<float> w = particles_.omega.to_host();
    std::vector<double> p64(2 * (size_t)N), g64(N);
    for (int i = 0; i < N; ++i) {
        p64[2 * i] = p[i].x;
        p64[2 * i + 1] = p[i].y;
        g64[i] = w[i];
    }
    dvec<double> dp, dg, du(2 * (size_t)N);
    dp.assign_host(p64.data(), p64.size());
    dg.assign_host(g64.data(), g64.size());
    vcheck(vpmc_direct_sum_eps(N, raw_ptr(dp), 1, N, raw_ptr(dp), raw_ptr(dg), P_.Lx,
                               periodic ? 1 : 0, eps_core, raw_ptr(du), stream),
           "compute_velocity_naive");
    std::vector<double> u = du.to_host();
    std::vector<float2> v(N);
    for (int i = 0; i < N; ++i) v[i] = make_float2((float)u[2 * i], (float)u[2 * i + 1]);
    if (particles_.vel.size() < (size_t)N) particles_.vel.resize(N);
    vpmc_detail::hip_check(hipMemcpy(particles_.d_v(), v.data(), N * sizeof(float2),
                                     hipMemcpyHostToDevice),
                           "compute_velocity_naive");
}

void Simulation::step_cuda_naive(bool periodic)
{
    const int N = particles_.N_cur;
    compute_velocity_naive(periodic, 0.5f * P_.h_w);
    Impl &I = *impl_;
    I.vel_sign.resize(2 * (size_t)N);   // [v, 0]: advect forms v - 0
    vpmc_detail::hip_check(hipMemcpy(raw_ptr(I.vel_sign), particles_.d_v(), N * sizeof(float2),
                                     hipMemcpyDeviceToDevice),
                           "step_cuda_naive");
    vpmc_detail::hip_check(hipMemset(raw_ptr(I.vel_sign) + N, 0, N * sizeof(float2)),
                           "step_cuda_naive");
    vcheck(vpmc_advect(N, (float *)particles_.d_p(), (const float *)raw_ptr(I.vel_sign),
                       (float *)particles_.d_v(), P_.dt, P_.Lx, P_.Ly, periodic ? 1 : 0, stream),
           "step_cuda_naive");
    vpmc_detail::hip_check(hipStreamSynchronize(stream), "step_cuda_naive");
}

void Simulation::test_vorticity_grid(int res)
{
    // src/vortex_particle_mc.cu:332-456: hashed vs naive positive-set query on
    // a res x res grid; prints the max abs and mean relative error
    std::vector<float2> q((size_t)res * res);
    for (int i = 0; i < res; ++i)
        for (int j = 0; j < res; ++j) q[(size_t)i * res + j] = make_float2((i + 0.5f) / res,
                                                                         (j + 0.5f) / res);
    dvec<float2> dq(q);
    dvec<float> a(q.size()), b(q.size());
    VorticityView h = makePosView(), nv = makePosViewNaive();
    auto run = [&](const VorticityView &v, dvec<float> &out) {
        vpmc_view cv{(const float *)v.pos, v.w, v.N, v.vp_sort, v.cell_acc, v.rshx, v.rshy,
                     v.Lx, v.Ly, v.cnt_view, v.sub};
        vcheck(vpmc_query_view(&cv, (int)q.size(), (const float *)raw_ptr(dq), 1, 0, 0, P_.h_w,
                               1, raw_ptr(out), stream),
               "test_vorticity_grid");
    };
    run(h, a);
    run(nv, b);
    std::vector<float> A = a.to_host(), B = b.to_host();
    double maxabs = 0, rel = 0;
    int cnt = 0;
    for (size_t i = 0; i < A.size(); ++i) {
        maxabs = std::max(maxabs, (double)std::fabs(A[i] - B[i]));
        if (std::fabs(B[i]) > 1e-6f) {
            rel += std::fabs(A[i] - B[i]) / std::fabs(B[i]);
            ++cnt;
        }
    }
    std::cout << "[test_vorticity_grid] max_abs_err = " << maxabs
              << "  mean_rel_err = " << (cnt ? rel / cnt : 0.0) << std::endl;
}

void Simulation::ensure_icdf_kernel(int L)
{
    if (icdf_L_ == L && icdf_kernel_.size() == (size_t)L) return;
    dvec<float> cdf;
    build_cubic_rcdf(P_.h_w, L, cdf);
    cdf_to_icdf(cdf, P_.h_w, icdf_kernel_);
    icdf_L_ = L;
}

VorticityView Simulation::makePosView(const float *w_ptr) const
{
    const SpatialHash &H = hash_pos_;
    return {particles_.d_p(), particles_.d_F(), w_ptr ? w_ptr : particles_.d_omega(),
            particles_.N_cur, raw_ptr(H.vp_sort), raw_ptr(H.cell_acc), H.rshx, H.rshy, P_.Lx,
            P_.Ly, 1.0f / P_.Lx, 1.0f / P_.Ly, cnt_pos_, raw_ptr(sub_pos_)};
}

VorticityView Simulation::makeNegView(const float *w_ptr) const
{
    const SpatialHash &H = hash_neg_;
    return {particles_.d_p(), particles_.d_F(), w_ptr ? w_ptr : particles_.d_omega(),
            particles_.N_cur, raw_ptr(H.vp_sort), raw_ptr(H.cell_acc), H.rshx, H.rshy, P_.Lx,
            P_.Ly, 1.0f / P_.Lx, 1.0f / P_.Ly, cnt_neg_, raw_ptr(sub_neg_)};
}

VorticityView Simulation::makePosViewNaive(const float *w_ptr) const
{
    return {particles_.d_p(), particles_.d_F(), w_ptr ? w_ptr : particles_.d_omega(),
            particles_.N_cur, nullptr, nullptr, 0, 0, P_.Lx, P_.Ly, 1.0f / P_.Lx, 1.0f / P_.Ly,
            cnt_pos_, raw_ptr(sub_pos_)};
}

VorticityView Simulation::makeNegViewNaive(const float *w_ptr) const
{
    return {particles_.d_p(), particles_.d_F(), w_ptr ? w_ptr : particles_.d_omega(),
            particles_.N_cur, nullptr, nullptr, 0, 0, P_.Lx, P_.Ly, 1.0f / P_.Lx, 1.0f / P_.Ly,
            cnt_neg_, raw_ptr(sub_neg_)};
}

VorticityView Simulation::makeAllView(const float *w_ptr) const
{
    const SpatialHash &H = hash_all_;
    return {particles_.d_p(), particles_.d_F(), w_ptr ? w_ptr : particles_.d_omega(),
            particles_.N_cur, raw_ptr(H.vp_sort), raw_ptr(H.cell_acc), H.rshx, H.rshy, P_.Lx,
            P_.Ly, 1.0f / P_.Lx, 1.0f / P_.Ly, particles_.N_cur, raw_ptr(sub_all_)};
}

// ---------------------------------------------------------------------------
// monte_carlo_bs.hpp

void monte_carlo_bs_gpu(Simulation &sim, int N1, int N2, std::vector<MCBSEstimate> &est)
{
    int s1 = sim.P_.N1, s2 = sim.P_.N2;
    sim.P_.N1 = N1;
    sim.P_.N2 = N2;
    try {
        sim.estimate_velocity(sim.P_.periodic);
    } catch (...) {
        sim.P_.N1 = s1;
        sim.P_.N2 = s2;
        throw;
    }
    sim.P_.N1 = s1;
    sim.P_.N2 = s2;
    std::vector<float2> v = sim.particles_.vel.to_host();
    est.assign(sim.particles_.N_cur, MCBSEstimate{});
    for (int i = 0; i < sim.particles_.N_cur; ++i) est[i].main = v[i];
}

MCBSCtx make_mcbs_ctx(const Simulation &sim, bool periodic)
{
    MCBSCtx C;
    C.sub_pos = &sim.sub_pos_;
    C.sub_neg = &sim.sub_neg_;
    C.cdf_pos = &sim.cdf_pos_f32_;
    C.cdf_neg = &sim.cdf_neg_f32_;
    C.cdf_pos64 = &sim.cdf_pos_;
    C.cdf_neg64 = &sim.cdf_neg_;
    C.cum_pos64 = sim.cum_pos_;
    C.cum_neg64 = sim.cum_neg_;
    C.icdf_kernel = &sim.icdf_kernel_;
    C.kernel_L = sim.icdf_L_;
    C.h_w = sim.P_.h_w;
    C.periodic = periodic;
    C.Lx = sim.P_.Lx;
    C.Ly = sim.P_.Ly;
    C.InvLx = 1.f / sim.P_.Lx;
    C.InvLy = 1.f / sim.P_.Ly;
    C.cum_pos = (float)sim.cum_pos_;
    C.cum_neg = (float)sim.cum_neg_;
    C.pos_view = sim.makePosView();
    C.neg_view = sim.makeNegView();
    C.cnt_pos = sim.cnt_pos_;
    C.cnt_neg = sim.cnt_neg_;
    return C;
}

MCBSEstimate monte_carlo_bs_cpu(const float2 &dst, int sign, int auxiliary,
                                const float2 &dst_left, const float2 &dst_right,
                                const float2 &dst_top, const float2 &dst_bottom,
                                const ParticleSet &P, const MCBSCtx &ctx, int N1, int N2,
                                std::mt19937 &rng)
{
    if (N1 < 0 || N2 < 0 || N1 + N2 <= 0)
        throw std::invalid_argument("monte_carlo_bs_cpu: N1, N2 >= 0 with N1 + N2 > 0");
    ReplaySet S[2];
    replay_set_build(S[0], ctx, sign, P, nullptr);
    const int nt = auxiliary ? 5 : 1;
    std::vector<float> u[2];
    u[0].resize(mcbs_call_draws(N1, N2, ctx.periodic, auxiliary != 0));
    mcbs_draw(rng, u[0].size(), u[0].data());
    u[1] = u[0];
    // the set occupies slot 0; slot 1 (the other sign) is given no particles
    S[1].cdf = S[0].cdf;
    S[1].nx = S[0].nx;
    S[1].ny = S[0].ny;
    S[1].glog2 = 0;
    const int zero = 0;
    const double one = 1.0;
    S[1].count.assign_host(&zero, 1);
    S[1].total.assign_host(&one, 1);
    S[1].guide.resize(2);
    S[1].pts.resize(1);
    S[1].set_xy.resize(1);
    S[1].cell_start.resize((size_t)S[1].nx * S[1].ny + 1);
    const std::vector<float2> tg = {dst, dst_left, dst_right, dst_top, dst_bottom};
    const std::vector<float2> v =
        replay_estimate(std::vector<float2>(tg.begin(), tg.begin() + nt), nt, S, u, N1, N2,
                        ctx.periodic, ctx.h_w, ctx.Lx, ctx.Ly, nullptr);
    MCBSEstimate e;
    e.main = v[0];
    if (auxiliary) {
        e.left = v[1];
        e.right = v[2];
        e.top = v[3];
        e.bottom = v[4];
    }
    return e;
}

float sample_kernel_radius(std::mt19937 &gen, const float *icdf, int L, float h)
{
    std::uniform_real_distribution<float> U(0.f, 1.f);
    const float pos = U(gen) * float(L - 1);   // position in the table
    const int lo = int(pos), hi = std::min(lo + 1, L - 1);
    const float f = pos - float(lo);
    return (icdf[lo] * (1.f - f) + icdf[hi] * f) * h;
}

int sample_vorticity_idx_cdf(std::mt19937 &gen, const float *cdf, int N)
{
    std::uniform_real_distribution<float> U(0.f, 1.f);
    const float u = U(gen);
    const int j = int(std::upper_bound(cdf, cdf + N, u) - cdf);
    return j < N ? j : N - 1;
}

float2 sample_disk_biot_savart(std::mt19937 &gen, const float2 &center, float R)
{
    std::uniform_real_distribution<float> U(0.f, 1.f);
    const float phi = 2.f * float(M_PI) * U(gen);
    const float r = R * std::sqrt(U(gen));
    return make_float2(center.x + r * std::cos(phi), center.y + r * std::sin(phi));
}

float evalVorticityAbs_cpu(const VorticityView &v, const float2 &p, int sign, bool use_signed,
                           bool exclude_center, float h_w, bool periodic)
{
    const vpmc_view q{(const float *)v.pos, v.w, v.N, v.vp_sort, v.cell_acc, v.rshx, v.rshy,
                      v.Lx, v.Ly, v.cnt_view, v.sub};
    dvec<float2> dp(std::vector<float2>{p});
    dvec<float> out(1);
    vcheck(vpmc_query_view(&q, 1, (const float *)raw_ptr(dp), sign, use_signed ? 1 : 0,
                           exclude_center ? 1 : 0, h_w, periodic ? 1 : 0, raw_ptr(out), nullptr),
           "evalVorticityAbs_cpu");
    vpmc_detail::hip_check(hipDeviceSynchronize(), "evalVorticityAbs_cpu");
    return out.to_host()[0];
}

// ---------------------------------------------------------------------------
// cg_rbf_init.cuh / cg_rbf.hpp

void Ax_cuda(const Simulation &sim, const float *x, float *y, hipStream_t stream)
{
    const int N = sim.particles_.N_cur;
    int nx, ny;
    vcheck(vpmc_field_grid(sim.P_.h_w, sim.P_.Lx, sim.P_.Ly, sim.cells_per_hw, &nx, &ny),
           "Ax_cuda");
    size_t b = vpmc_cg_workspace_bytes(N, nx, ny);
    vcheck(vpmc_rbf_matvec_particles(N, (const float *)sim.particles_.d_p(), x, y, sim.P_.h_w,
                                     sim.P_.Lx, sim.P_.Ly, sim.cells_per_hw,
                                     free_fn_scratch().get(b), b, stream),
           "Ax_cuda");
}

void conjugate_gradient_rbf(Simulation &sim, float tol, int max_iter)
{
    // src/cuda/cg_rbf_init.cu:53-116, then build_subsets + hashing (:113-115)
    const int N = sim.particles_.N_cur;
    if (N == 0) throw std::runtime_error("CG init: N_cur == 0 (particles not generated)");
    int nx, ny;
    vcheck(vpmc_field_grid(sim.P_.h_w, sim.P_.Lx, sim.P_.Ly, sim.cells_per_hw, &nx, &ny),
           "conjugate_gradient_rbf");
    size_t b = vpmc_cg_workspace_bytes(N, nx, ny);
    int it = 0;
    double rr = 0.0;
    int st = vpmc_cg_rbf(N, (const float *)sim.particles_.d_p(), sim.particles_.d_omega_field(),
                         sim.particles_.d_omega(), sim.P_.h_w, sim.P_.Lx, sim.P_.Ly,
                         sim.cells_per_hw, tol, max_iter, 1, &it, &rr, free_fn_scratch().get(b), b,
                         sim.stream);
    sim.cg_iterations = it;
    sim.cg_rel_residual = rr;
    if (st == VPMC_ERR_CG_BREAKDOWN || st == VPMC_ERR_NOT_CONVERGED)
        std::cerr << "[conjugate_gradient_rbf] " << vpmc_status_string(st) << " after " << it
                  << " iterations, relative residual " << rr << std::endl;
    else
        vcheck(st, "conjugate_gradient_rbf");
    sim.build_subsets_cuda();
    sim.do_spatial_hashing();
}

void build_global_hash_all(Simulation &sim) { sim.build_global_hash(); }

void solve_cg_rbf_cpu(Simulation &sim, float h_w, float rel_tol, float alpha_tol, int max_iter)
{
    if (h_w != sim.P_.h_w)
        throw std::invalid_argument("solve_cg_rbf_cpu: h_w must equal the simulation's h_w");
    (void)alpha_tol;
    conjugate_gradient_rbf(sim, rel_tol, max_iter);
}
