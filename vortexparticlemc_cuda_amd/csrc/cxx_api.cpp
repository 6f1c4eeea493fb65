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
};

Simulation::Simulation(SimParam p) : P_(std::move(p)), impl_(new Impl) {}
Simulation::~Simulation() { delete impl_; }

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
    throw std::logic_error("build_subsets_cpu: the CPU backend is not part of libvpmc");
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
    return d;
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
    if (kernel_deformation) attach_deformation(d, I, deformation_bound);
    size_t wb = vpmc_mc_binned_workspace_bytes(&d);
    vcheck(vpmc_mc_estimate_binned(&d, I.ws.get(wb), wb, stream), "monte_carlo_bs");
    // vel = v+ - v- (dt = 0 leaves the positions unchanged)
    vcheck(vpmc_advect(N, (float *)particles_.d_p(), (const float *)raw_ptr(I.vel_sign),
                       (float *)particles_.d_v(), 0.f, P_.Lx, P_.Ly, periodic ? 1 : 0, stream),
           "combine");
    vpmc_detail::hip_check(hipStreamSynchronize(stream), "estimate_velocity");
}

void Simulation::step_cpu(bool)
{
    throw std::logic_error("step_cpu: the CPU backend is not part of libvpmc");
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
