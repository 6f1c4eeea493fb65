// cxx_api_check.cpp -- exercises the C++ API (the reference's interface) on
// the device and writes its results for tests/test_gpu_cxx_api.py.
//   usage: vpmc_cxx_check <input.bin> <outdir>
// input.bin: int32 Nx, int32 hwr, int32 N, float32 pos[2N], float32 b[N]
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include "cg_rbf.hpp"
#include "cg_rbf_init.cuh"
#include "monte_carlo_bs.hpp"
#include "spatial_hasher.hpp"
#include "vortex_particle_mc.hpp"

template <class T> static void dump(const std::string &path, const std::vector<T> &v)
{
    std::ofstream f(path, std::ios::binary);
    f.write(reinterpret_cast<const char *>(v.data()), v.size() * sizeof(T));
}

int main(int argc, char **argv)
{
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

        sim.test_vorticity_grid(16);

        bool threw = false;
        try {
            SpatialHasher bad(4, 4, Backend::CPU);
        } catch (const std::invalid_argument &) {
            threw = true;
        }
        if (!threw) throw std::runtime_error("Backend::CPU accepted");
        threw = false;
        try {
            sim.step_cpu(true);
        } catch (const std::logic_error &) {
            threw = true;
        }
        if (!threw) throw std::runtime_error("step_cpu accepted");
    } catch (const std::exception &e) {
        std::cerr << "error: " << e.what() << std::endl;
        return 1;
    }
    std::cout << "cxx api ok" << std::endl;
    return 0;
}
