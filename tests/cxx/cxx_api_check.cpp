// cxx_api_check.cpp -- exercises the C++ API (the reference's interface) on
// the device and writes its results for tests/test_gpu_cxx_api.py.
//   usage: vpmc_cxx_check <input.bin> <outdir>
// input.bin: int32 Nx, int32 hwr, int32 N, float32 pos[2N], float32 b[N]
//   usage: vpmc_cxx_check <input.bin> <outdir> mcbs | hash
// input.bin: int32 Nx, hwr, n, N1, N2, n_own, seed; float32 dx; float32 px[n],
//   py[n], w[n]; int32 cnt_pos, cnt_neg; float32 cdf_pos[cnt_pos],
//   cdf_neg[cnt_neg], cum[2]; int32 nq; float32 q[2 nq]
// (the CPU backend's entry points: monte_carlo_bs_cpu, the samplers,
// step_cpu, evalVorticityAbs_cpu)
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <random>
#include <string>
#include <vector>

#include "cg_rbf.hpp"
#include "cg_rbf_init.cuh"
#include "monte_carlo_bs.cuh"
#include "monte_carlo_bs.hpp"
#include "spatial_hasher.hpp"
#include "viz_utils.hpp"
#include "vortex_particle_mc.hpp"

// the device helpers of monte_carlo_bs.cuh on n points: per point the view
// query, F v, the minimum image of p.x - 0.9, renorm_det1(F), K_eps(p - c)
__global__ void k_cuh_helpers(VorticityView v, const float2 *p, int n, float h_w, float *out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 F = make_float4(1.2f, 0.3f, -0.1f, 0.9f);
    const float2 fv = mul_mat2(F, p[i]);
    const float4 r = renorm_det1(F);
    const float2 k = bs_kernel_2d(make_float2(p[i].x - 0.5f, p[i].y - 0.5f), 1e-4f);
    float *o = out + (size_t)i * 10;
    o[0] = evalVorticityAbs_gpu(v, p[i], +1, false, false, h_w);
    o[1] = fv.x;
    o[2] = fv.y;
    o[3] = wrap_min_image(p[i].x - 0.9f, 1.f, true);
    o[4] = r.x;
    o[5] = r.y;
    o[6] = r.z;
    o[7] = r.w;
    o[8] = k.x;
    o[9] = k.y;
}

template <class T> static void dump(const std::string &path, const std::vector<T> &v)
{
    std::ofstream f(path, std::ios::binary);
    f.write(reinterpret_cast<const char *>(v.data()), v.size() * sizeof(T));
}

template <class T> static std::vector<T> take(std::ifstream &in, size_t n)
{
    std::vector<T> v(n);
    in.read(reinterpret_cast<char *>(v.data()), n * sizeof(T));
    return v;
}

static int run_mcbs(const char *input, const std::string &out)
{
    std::ifstream in(input, std::ios::binary);
    const std::vector<int> hdr = take<int>(in, 7);
    const int Nx = hdr[0], hwr = hdr[1], n = hdr[2], N1 = hdr[3], N2 = hdr[4], n_own = hdr[5];
    const uint32_t seed = (uint32_t)hdr[6];
    const float dx = take<float>(in, 1)[0];
    const std::vector<float> px = take<float>(in, n), py = take<float>(in, n),
                             w = take<float>(in, n);
    const std::vector<int> cnt = take<int>(in, 2);
    const std::vector<float> cp = take<float>(in, cnt[0]), cn = take<float>(in, cnt[1]),
                             cum = take<float>(in, 2);
    const int nq = take<int>(in, 1)[0];
    const std::vector<float2> q = take<float2>(in, nq);
    try {
        SimParam P(1e-3f, 1.f, 1.f, Nx, (float)hwr);
        P.N1 = N1;
        P.N2 = N2;
        std::vector<float2> pos(n);
        for (int i = 0; i < n; ++i) pos[i] = make_float2(px[i], py[i]);
        Simulation sim(P);
        sim.set_particles(pos, w);

        // monte_carlo_bs_cpu with the reference's FP32 CDFs and totals in the
        // reference's own MCBSCtx members (as src/cpu/simulation_step.cpp:13-27
        // fills them), owners 0 .. n_own-1, +1 then -1, one generator, aux
        // targets at dst -/+ dx
        MCBSCtx C = make_mcbs_ctx(sim, true);
        dvec<float> dcp(cp), dcn(cn);
        C.cdf_pos = &dcp;
        C.cdf_neg = &dcn;
        C.cdf_pos64 = nullptr;   // no FP64 CDFs: the FP32 members are used
        C.cdf_neg64 = nullptr;
        C.cnt_pos = -1;          // the CDF vectors' sizes, as in the reference
        C.cnt_neg = -1;
        C.cum_pos = cum[0];
        C.cum_neg = cum[1];
        std::mt19937 rng(seed);
        std::vector<float> est;
        for (int i = 0; i < n_own; ++i) {
            const float2 d = pos[i];
            const float2 dl{d.x - dx, d.y}, dr{d.x + dx, d.y}, dt{d.x, d.y + dx},
                db{d.x, d.y - dx};
            for (int sg = 0; sg < 2; ++sg) {
                const MCBSEstimate e = monte_carlo_bs_cpu(d, sg == 0 ? +1 : -1, 1, dl, dr, dt, db,
                                                          sim.particles_, C, N1, N2, rng);
                for (const float2 &v : {e.main, e.left, e.right, e.top, e.bottom}) {
                    est.push_back(v.x);
                    est.push_back(v.y);
                }
            }
        }
        dump(out + "/mcbs_est.bin", est);
        // the generator advanced by exactly the reference's draws
        std::mt19937 twin(seed);
        std::uniform_real_distribution<float> U(0.f, 1.f);
        const size_t used = (size_t)n_own * 2 * ((size_t)N1 * 3 + (size_t)N2 * 10);
        for (size_t k = 0; k < used; ++k) (void)U(twin);
        const std::vector<int> rng_ok = {U(rng) == U(twin) ? 1 : 0};
        dump(out + "/mcbs_rng.bin", rng_ok);

        // the samplers on mt19937(seed)'s first draws
        std::mt19937 g(seed);
        std::vector<float> smp;
        smp.push_back((float)sample_vorticity_idx_cdf(g, cp.data(), cnt[0]));
        const float2 disk = sample_disk_biot_savart(g, make_float2(0.3f, 0.6f), 0.7f);
        smp.push_back(disk.x);
        smp.push_back(disk.y);
        const float icdf[5] = {0.f, 0.3f, 0.5f, 0.7f, 1.f};
        smp.push_back(sample_kernel_radius(g, icdf, 5, 0.25f));
        dump(out + "/samplers.bin", smp);

        // evalVorticityAbs_cpu on the reference grid's views
        std::vector<float> qv;
        for (int sg = 0; sg < 2; ++sg)
            for (const float2 &p : q)
                qv.push_back(evalVorticityAbs_cpu(sg == 0 ? sim.makePosView() : sim.makeNegView(),
                                                  p, sg == 0 ? +1 : -1, false, false, P.h_w,
                                                  true));
        dump(out + "/evalq.bin", qv);

        // step_cpu from mt19937(seed), FP64 subsets of the simulation
        sim.rng = std::mt19937(seed);
        sim.step_cpu(true);
        dump(out + "/step_vel.bin", sim.particles_.vel.to_host());
        dump(out + "/step_pos.bin", sim.particles_.pos.to_host());
        dump(out + "/step_left.bin", sim.particles_.pos_left.to_host());
        std::mt19937 fresh(seed);
        const std::vector<int> member_ok = {sim.rng() == fresh() ? 1 : 0};
        dump(out + "/step_rng_member.bin", member_ok);

        // step_cpu over several replay chunks (N1 = N2 = 2000: 671 particles a
        // chunk) == monte_carlo_bs_cpu per particle on the same generator
        {
            SimParam P2 = P;
            P2.N1 = P2.N2 = 2000;
            Simulation s2(P2);
            s2.set_particles(pos, w);
            s2.rng = std::mt19937(seed);
            const MCBSCtx C2 = make_mcbs_ctx(s2, true);
            const std::vector<float2> x0 = s2.particles_.pos.to_host();
            std::mt19937 g2(seed);
            std::vector<float2> loop;
            const float2 z{0.f, 0.f};   // the aux points start at the origin, as in the reference
            for (int i = 0; i < 700; ++i) {
                const MCBSEstimate a = monte_carlo_bs_cpu(x0[i], +1, 1, z, z, z, z, s2.particles_,
                                                          C2, 2000, 2000, g2);
                const MCBSEstimate b = monte_carlo_bs_cpu(x0[i], -1, 1, z, z, z, z, s2.particles_,
                                                          C2, 2000, 2000, g2);
                loop.push_back(make_float2(a.main.x - b.main.x, a.main.y - b.main.y));
            }
            s2.step_cpu(true);
            std::vector<float2> v2 = s2.particles_.vel.to_host();
            v2.resize(700);
            dump(out + "/step_chunks_vel.bin", v2);
            dump(out + "/step_chunks_loop.bin", loop);
        }
    } catch (const std::exception &e) {
        std::cerr << "error: " << e.what() << std::endl;
        return 1;
    }
    std::cout << "cxx mcbs ok" << std::endl;
    return 0;
}

// hash mode: int32 ncase; per case int32 n, rsh; float32 px[n], py[n].
// Writes hash_<backend>_<case>.bin = cell_num, cell_acc, vp_cell, vp_map,
// cell_num after the sort (int32), from the host-span variants.
static int run_hash(const char *input, const std::string &out)
{
    std::ifstream in(input, std::ios::binary);
    const int ncase = take<int>(in, 1)[0];
    try {
        for (int c = 0; c < ncase; ++c) {
            const std::vector<int> h = take<int>(in, 2);
            const int n = h[0], rsh = h[1], cells = rsh * rsh;
            const std::vector<float> px = take<float>(in, n), py = take<float>(in, n);
            std::vector<int> sub(n);
            for (int i = 0; i < n; ++i) sub[i] = i;
            for (int omp = 0; omp < 2; ++omp) {
                std::vector<int> num(cells), acc(cells), cell(n), map(n);
                if (omp)
                    FillCells_omp(rsh, rsh, 1.f, 1.f, n, sub, px, py, num, acc, cell);
                else
                    FillCells_cpu(rsh, rsh, 1.f, 1.f, n, sub, px, py, num, acc, cell);
                std::vector<int> res(num);
                res.insert(res.end(), acc.begin(), acc.end());
                res.insert(res.end(), cell.begin(), cell.end());
                if (omp)
                    CountingSort_omp(n, num, acc, cell, map);
                else
                    CountingSort_cpu(n, num, acc, cell, map);
                res.insert(res.end(), map.begin(), map.end());
                res.insert(res.end(), num.begin(), num.end());
                dump(out + "/hash_" + (omp ? "omp_" : "cpu_") + std::to_string(c) + ".bin", res);
            }
        }
    } catch (const std::exception &e) {
        std::cerr << "error: " << e.what() << std::endl;
        return 1;
    }
    std::cout << "cxx hash ok" << std::endl;
    return 0;
}

int main(int argc, char **argv)
{
    if (argc >= 4 && std::string(argv[3]) == "mcbs") return run_mcbs(argv[1], argv[2]);
    if (argc >= 4 && std::string(argv[3]) == "hash") return run_hash(argv[1], argv[2]);
    if (argc < 3) {
        std::fprintf(stderr, "usage: %s input.bin outdir\n", argv[0]);
        return 2;
    }
    std::ifstream in(argv[1], std::ios::binary);
    int hdr[3];
    in.read(reinterpret_cast<char *>(hdr), sizeof(hdr));
    const int Nx = hdr[0], hwr = hdr[1], N = hdr[2];
    std::vector<float2> pos(N);
    std::vector<float> b(N);
    in.read(reinterpret_cast<char *>(pos.data()), N * sizeof(float2));
    in.read(reinterpret_cast<char *>(b.data()), N * sizeof(float));
    const std::string out = argv[2];
    try {
        SimParam P(1e-3f, 1.f, 1.f, Nx, (float)hwr);
        P.N1 = 64;
        P.N2 = 64;
        Simulation sim(P);
        sim.mode = EstimatorMode::Consistent;
        sim.seed = 7;
        sim.init_with(pos, b, 1e-6f, 1000);
        std::cout << "cg iterations " << sim.cg_iterations << " rel residual "
                  << sim.cg_rel_residual << std::endl;
        dump(out + "/omega.bin", sim.particles_.omega.to_host());

        // SpatialHasher over all particles on the reference grid
        int rsh = Nx / hwr / 2;
        SpatialHasher H(rsh, rsh, Backend::CUDA);
        H.resize(N);
        std::vector<int> iota(N);
        for (int i = 0; i < N; ++i) iota[i] = i;
        dvec<int> sub(iota);
        H.build(sim.particles_.pos, sub, 1.f, 1.f);
        dump(out + "/vp_cell.bin", H.data().vp_cell.to_host());
        dump(out + "/vp_sort.bin", H.data().vp_sort.to_host());
        dump(out + "/cell_acc.bin", H.data().cell_acc.to_host());
        dump(out + "/cell_num_after.bin", H.data().cell_num.to_host());

        // estimator at the initial positions
        std::vector<MCBSEstimate> est;
        monte_carlo_bs_gpu(sim, 64, 64, est);
        std::vector<float2> v(N);
        for (int i = 0; i < N; ++i) v[i] = est[i].main;
        dump(out + "/vel0.bin", v);

        // Ax on the strengths
        dvec<float> y(N);
        Ax_cuda(sim, sim.particles_.d_omega(), raw_ptr(y));
        (void)hipDeviceSynchronize();
        dump(out + "/Ax.bin", y.to_host());

        // naive direct sum (eps = 0.5 h_w)
        sim.compute_velocity_naive(true, 0.5f * sim.P_.h_w);
        dump(out + "/vel_naive.bin", sim.particles_.vel.to_host());

        // two plain steps
        sim.step_cuda(true);
        sim.step_cuda(true);
        (void)hipDeviceSynchronize();
        dump(out + "/pos2.bin", sim.particles_.pos.to_host());
        dump(out + "/vel2.bin", sim.particles_.vel.to_host());
        // the driver's output files (viz_utils.hpp) of this state
        dump_vorticity_png(sim, out + "/vort_per.png", 48, true);
        dump_vorticity_png(sim, out + "/vort.png", 48);
        dump_velocity_svg_centered(sim, out + "/svg/vel.svg", 0.2f, true);
        {   // monte_carlo_bs.cuh's device helpers, and the host query, at 16 points
            std::vector<float2> hp(16);
            for (int i = 0; i < 16; ++i)
                hp[i] = make_float2(0.03f + 0.061f * i, 0.97f - 0.057f * i);
            dvec<float2> dp(hp);
            dvec<float> res(160);
            const VorticityView pv = sim.makePosView();
            hipLaunchKernelGGL(k_cuh_helpers, dim3(1), dim3(64), 0, 0, pv, raw_ptr(dp), 16,
                               sim.P_.h_w, raw_ptr(res));
            (void)hipDeviceSynchronize();
            std::vector<float> h = res.to_host();
            for (int i = 0; i < 16; ++i)
                h.push_back(evalVorticityAbs_cpu(pv, hp[i], +1, false, false, sim.P_.h_w, true));
            dump(out + "/cuh_helpers.bin", h);
            dump(out + "/cuh_points.bin", hp);
        }

        // two steps of the CUDA step with aux targets and kernel deformation
        {
            Simulation sd(P);
            sd.mode = EstimatorMode::Consistent;
            sd.seed = 7;
            sd.kernel_deformation = true;
            sd.set_particles(pos, sim.particles_.omega.to_host());
            sd.step_cuda(true);
            sd.step_cuda(true);
            (void)hipDeviceSynchronize();
            dump(out + "/pos_def2.bin", sd.particles_.pos.to_host());
            dump(out + "/F_def2.bin", sd.particles_.F.to_host());
            sd.reinitialize(0, 0, 1e-6f, 1000);
            dump(out + "/pos_reinit.bin", sd.particles_.pos.to_host());
            dump(out + "/omega_reinit.bin", sd.particles_.omega.to_host());
        }

        // Simulation::eval_vorticity_cpu (w+ - w- of the reference-grid views,
        // hashed at the current positions as the reference's main loop does)
        {
            sim.do_spatial_hashing();
            std::vector<float> ev;
            for (int i = 0; i < 16; ++i)
                ev.push_back(sim.eval_vorticity_cpu(
                    make_float2(0.03f + 0.061f * i, 0.97f - 0.057f * i), true));
            dump(out + "/eval_vort.bin", ev);
        }

        // a bounded box: estimate_velocity(false), then step_cuda(false), in
        // both estimator modes (the direct estimator; no wrap)
        const std::vector<float> omega0 = sim.particles_.omega.to_host();
        for (int md = 0; md < 2; ++md) {
            Simulation sn(P);
            sn.mode = md ? EstimatorMode::Consistent : EstimatorMode::Reference;
            sn.seed = 7;
            sn.set_particles(pos, omega0);
            sn.estimate_velocity(false);
            dump(out + "/vel_np" + std::to_string(md) + ".bin", sn.particles_.vel.to_host());
            sn.step_cuda(false);
            (void)hipDeviceSynchronize();
            dump(out + "/pos_np" + std::to_string(md) + ".bin", sn.particles_.pos.to_host());
        }

        // copies own their buffers: a copy-constructed and a copy-assigned
        // simulation take the same step as the original, bit for bit, and
        // stepping the original leaves an earlier copy unchanged
        {
            Simulation c1 = sim;
            Simulation c2(SimParam(1e-3f, 1.f, 1.f, 16, 4.f));
            c2 = sim;
            const std::vector<float2> before = c1.particles_.pos.to_host();
            sim.step_cuda(true);
            const std::vector<float2> untouched = c1.particles_.pos.to_host();
            c1.step_cuda(true);
            c2.step_cuda(true);
            (void)hipDeviceSynchronize();
            const std::vector<float2> a = sim.particles_.pos.to_host(),
                                      b = c1.particles_.pos.to_host(),
                                      c = c2.particles_.pos.to_host();
            const size_t nb = a.size() * sizeof(float2);
            const bool same = b.size() == a.size() && c.size() == a.size() &&
                              std::memcmp(a.data(), b.data(), nb) == 0 &&
                              std::memcmp(a.data(), c.data(), nb) == 0;
            const bool kept = std::memcmp(before.data(), untouched.data(), nb) == 0 &&
                              std::memcmp(before.data(), a.data(), nb) != 0;
            dump(out + "/copy_ok.bin", std::vector<int>{same ? 1 : 0, kept ? 1 : 0});
        }

        sim.test_vorticity_grid(16);

        // Backend::CPU / OMP from host containers: the same hash
        for (Backend b : {Backend::CPU, Backend::OMP}) {
            SpatialHasher Hh(rsh, rsh, b);
            Hh.resize(N);
            Hh.build(sim.particles_.pos.to_host(), iota, 1.f, 1.f);
            if (Hh.data().vp_sort.to_host() != H.data().vp_sort.to_host() ||
                Hh.data().cell_acc.to_host() != H.data().cell_acc.to_host())
                throw std::runtime_error("host-container SpatialHasher differs");
        }
    } catch (const std::exception &e) {
        std::cerr << "error: " << e.what() << std::endl;
        return 1;
    }
    std::cout << "cxx api ok" << std::endl;
    return 0;
}
