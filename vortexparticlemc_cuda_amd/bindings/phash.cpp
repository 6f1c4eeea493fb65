// phash.cpp -- the reference's Python bindings module, phash
// (bindings/phash.cpp:6-14: a pybind11 module that exposes fill_dummy), here
// carrying the C++ API of the library (include/vortex_particle_mc.hpp:
// SimParam, Simulation; include/cg_rbf_init.cuh: conjugate_gradient_rbf) and
// the differentiable rollout as a torch.autograd.Function written in C++
// (torch::autograd::Function) over Simulation::step_cuda and
// Simulation::step_adjoint, the hand-written adjoint of the step.
//
// Host code only: every computation is a libvpmc kernel reached through the
// C++ API.  Arrays cross as numpy (host) or torch CUDA tensors (device).
#include <torch/extension.h>

#include <pybind11/numpy.h>
#include <hip/hip_runtime.h>

#include <stdexcept>
#include <string>
#include <vector>

#include "cg_rbf_init.cuh"
#include "monte_carlo_bs.hpp"
#include "vortex_particle_mc.hpp"
#include "vpmc.h"

namespace py = pybind11;
using torch::Tensor;

namespace {

using f32_array = py::array_t<float, py::array::c_style | py::array::forcecast>;

void hip_ok(hipError_t e, const char *what)
{
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

std::vector<float2> as_float2(const f32_array &a, const char *what)
{
    if (a.ndim() != 2 || a.shape(1) != 2)
        throw std::invalid_argument(std::string(what) + ": positions must be an (N, 2) array");
    const float *p = a.data();
    std::vector<float2> v((size_t)a.shape(0));
    for (size_t i = 0; i < v.size(); ++i) v[i] = make_float2(p[2 * i], p[2 * i + 1]);
    return v;
}

std::vector<float> as_floats(const f32_array &a)
{
    return std::vector<float>(a.data(), a.data() + a.size());
}

f32_array float2_array(const dvec<float2> &d, int n)
{
    std::vector<float2> h = d.to_host();
    f32_array out({(py::ssize_t)n, (py::ssize_t)2});
    float *o = out.mutable_data();
    for (int i = 0; i < n; ++i) {
        o[2 * i] = h[i].x;
        o[2 * i + 1] = h[i].y;
    }
    return out;
}

template <class T> py::array_t<T> flat_array(const dvec<T> &d, int n)
{
    std::vector<T> h = d.to_host();
    py::array_t<T> out((py::ssize_t)n);
    std::copy(h.begin(), h.begin() + n, out.mutable_data());
    return out;
}

Tensor device_f32(const Tensor &t, const char *what)
{
    if (!t.is_cuda()) throw std::invalid_argument(std::string(what) + ": a device tensor");
    return t.detach().to(torch::kFloat32).contiguous();
}

// positions (N, 2) and strengths (N) of a tensor pair into the simulation;
// the sampling tables and both fields follow the strengths
void load_state(Simulation &sim, const Tensor &pos, const Tensor &w)
{
    const int N = sim.particles_.N_cur;
    if (pos.dim() != 2 || pos.size(0) != N || pos.size(1) != 2 || w.numel() != N)
        throw std::invalid_argument("rollout: pos (N, 2) and w (N) for the simulation's N");
    hip_ok(hipDeviceSynchronize(), "rollout");
    hip_ok(hipMemcpy(sim.particles_.d_omega(), w.data_ptr<float>(), N * sizeof(float),
                     hipMemcpyDeviceToDevice),
           "rollout");
    hip_ok(hipMemcpy(sim.particles_.d_p(), pos.data_ptr<float>(), N * sizeof(float2),
                     hipMemcpyDeviceToDevice),
           "rollout");
    sim.build_subsets_cuda();
    sim.do_spatial_hashing();
}

Tensor positions_tensor(const Simulation &sim, const Tensor &like)
{
    Tensor out = torch::empty({sim.particles_.N_cur, 2}, like.options());
    hip_ok(hipMemcpy(out.data_ptr<float>(), sim.particles_.d_p(),
                     sim.particles_.N_cur * sizeof(float2), hipMemcpyDeviceToDevice),
           "rollout");
    return out;
}

void store_positions(Simulation &sim, const Tensor &p)
{
    hip_ok(hipMemcpy(sim.particles_.d_p(), p.data_ptr<float>(),
                     sim.particles_.N_cur * sizeof(float2), hipMemcpyDeviceToDevice),
           "rollout");
}

// particles.F as a (N, 4) tensor (kernel deformation: F is part of the state)
Tensor jac_tensor(Simulation &sim, const Tensor &like)
{
    Tensor out = torch::empty({sim.particles_.N_cur, 4}, like.options().dtype(torch::kFloat32));
    hip_ok(hipMemcpy(out.data_ptr<float>(), sim.particles_.d_F(),
                     sim.particles_.N_cur * sizeof(float4), hipMemcpyDeviceToDevice),
           "rollout");
    return out;
}

void store_jac(Simulation &sim, const Tensor &F)
{
    hip_ok(hipMemcpy(sim.particles_.d_F(), F.data_ptr<float>(),
                     sim.particles_.N_cur * sizeof(float4), hipMemcpyDeviceToDevice),
           "rollout");
}

// steps t0 .. t1-1 of the FP64 trajectory from X (device (N, 2) double, in
// place) and, with kernel deformation, F64 (device (N, 4) double, in place);
// states receives the state before each step (positions, then F)
void run_f64(Simulation &sim, Tensor &X, Tensor &F64, const Tensor &w64, int64_t t0,
             int64_t t1, std::vector<Tensor> *states)
{
    const bool def = sim.kernel_deformation;
    for (int64_t t = t0; t < t1; ++t) {
        if (states) {
            states->push_back(X.clone());
            if (def) states->push_back(F64.clone());
        }
        sim.step_f64((uint32_t)t, reinterpret_cast<double2 *>(X.data_ptr<double>()),
                     w64.data_ptr<double>(),
                     def ? reinterpret_cast<double4 *>(F64.data_ptr<double>()) : nullptr);
    }
}

// Positions after `steps` steps from pos0 with strengths w (Philox steps
// 0 .. steps-1), differentiable in pos0 and w under common random numbers.
// traj64: the FP64 forward of the frozen-sample step (Simulation::step_f64),
// positions returned in FP64, the step adjoints linearised at those states
// rounded to FP32; otherwise step_cuda(true).  The forward keeps the state
// every `checkpoint` steps; the backward recomputes each segment from its
// checkpoint and runs Simulation::step_adjoint backwards through it.  With
// kernel deformation the state is (positions, F): the rollout starts from
// the simulation's current F and its checkpoints keep F beside the
// positions (one or two tensors per checkpoint).
struct RolloutFunction : public torch::autograd::Function<RolloutFunction> {
    static Tensor forward(torch::autograd::AutogradContext *ctx, const Tensor &pos0,
                          const Tensor &w, Simulation *sim, int64_t steps, int64_t checkpoint,
                          bool traj64)
    {
        const Tensor p = device_f32(pos0, "rollout"), ww = device_f32(w, "rollout");
        load_state(*sim, p, ww);
        const bool def = sim->kernel_deformation;
        const int64_t K = checkpoint > 0 ? checkpoint : 1;
        std::vector<Tensor> ckpt;
        Tensor out;
        if (traj64) {
            Tensor X = pos0.detach().to(torch::kFloat64).contiguous().clone();
            const Tensor w64 = w.detach().to(torch::kFloat64).contiguous();
            Tensor F64 = def ? jac_tensor(*sim, p).to(torch::kFloat64).contiguous() : Tensor();
            for (int64_t t0 = 0; t0 < steps; t0 += K) {
                ckpt.push_back(X.clone());
                if (def) ckpt.push_back(F64.clone());
                run_f64(*sim, X, F64, w64, t0, std::min(steps, t0 + K), nullptr);
            }
            if (def) store_jac(*sim, F64.to(torch::kFloat32).contiguous());
            ckpt.push_back(w64);
            out = X;
        } else {
            for (int64_t t0 = 0; t0 < steps; t0 += K) {
                ckpt.push_back(positions_tensor(*sim, p));
                if (def) ckpt.push_back(jac_tensor(*sim, p));
                for (int64_t t = t0; t < std::min(steps, t0 + K); ++t) {
                    sim->step_index = (uint32_t)t;
                    sim->step_cuda(true);
                }
            }
        }
        sim->step_index = (uint32_t)steps;
        hip_ok(hipDeviceSynchronize(), "rollout");
        ctx->save_for_backward(ckpt);
        ctx->saved_data["sim"] = (int64_t)reinterpret_cast<intptr_t>(sim);
        ctx->saved_data["steps"] = steps;
        ctx->saved_data["K"] = K;
        ctx->saved_data["traj64"] = traj64;
        ctx->saved_data["def"] = def;
        ctx->saved_data["dtype"] = pos0.scalar_type();
        return traj64 ? out : positions_tensor(*sim, p).to(pos0.scalar_type());
    }

    static torch::autograd::variable_list backward(torch::autograd::AutogradContext *ctx,
                                                   torch::autograd::variable_list grad)
    {
        Simulation *sim = reinterpret_cast<Simulation *>((intptr_t)ctx->saved_data["sim"].toInt());
        const int64_t steps = ctx->saved_data["steps"].toInt(), K = ctx->saved_data["K"].toInt();
        const bool traj64 = ctx->saved_data["traj64"].toBool();
        const bool def = ctx->saved_data["def"].toBool();
        if (def != sim->kernel_deformation)
            throw std::invalid_argument("rollout backward: kernel_deformation changed");
        std::vector<Tensor> ckpt = ctx->get_saved_variables();
        Tensor w64;
        if (traj64) {   // the strengths ride behind the checkpoints
            w64 = ckpt.back();
            ckpt.pop_back();
        }
        const int per = def ? 2 : 1;   // tensors per checkpoint
        const int N = sim->particles_.N_cur;
        Tensor lam = grad[0].detach().to(torch::kFloat32).contiguous().clone();
        Tensor gw = torch::zeros({N}, lam.options());
        Tensor jbar = def ? torch::zeros({N, 4}, lam.options()) : Tensor();   // dL/dF_T = 0
        hip_ok(hipDeviceSynchronize(), "rollout backward");
        const int64_t nseg = (int64_t)ckpt.size() / per;
        for (int64_t seg = nseg - 1; seg >= 0; --seg) {
            const int64_t t0 = seg * K, t1 = std::min(steps, t0 + K);
            std::vector<Tensor> states;   // per step: positions (then F)
            if (traj64) {   // recompute in FP64, linearise at the rounded states
                Tensor X = ckpt[per * seg].clone();
                Tensor F64 = def ? ckpt[per * seg + 1].clone() : Tensor();
                run_f64(*sim, X, F64, w64, t0, t1, &states);
                for (Tensor &st : states) st = st.to(torch::kFloat32).contiguous();
            } else {
                store_positions(*sim, ckpt[per * seg]);
                if (def) store_jac(*sim, ckpt[per * seg + 1]);
                for (int64_t t = t0; t < t1; ++t) {   // recompute the segment
                    states.push_back(positions_tensor(*sim, lam));
                    if (def) states.push_back(jac_tensor(*sim, lam));
                    sim->step_index = (uint32_t)t;
                    sim->step_cuda(true);
                }
            }
            for (int64_t t = t1 - 1; t >= t0; --t) {
                store_positions(*sim, states[per * (t - t0)]);
                if (def) store_jac(*sim, states[per * (t - t0) + 1]);
                sim->step_adjoint((uint32_t)t, reinterpret_cast<float2 *>(lam.data_ptr<float>()),
                                  gw.data_ptr<float>(),
                                  def ? reinterpret_cast<float4 *>(jbar.data_ptr<float>())
                                      : nullptr);
            }
        }
        sim->step_index = (uint32_t)steps;
        hip_ok(hipDeviceSynchronize(), "rollout backward");
        const auto dt = ctx->saved_data["dtype"].toScalarType();
        return {lam.to(dt), gw, Tensor(), Tensor(), Tensor(), Tensor()};
    }
};

}  // namespace

PYBIND11_MODULE(phash, m)
{
    m.doc() = "Vortex-particle Monte Carlo on MI355X: the reference's phash module with its "
              "C++ API and the differentiable rollout";

    // bindings/phash.cpp:10-13
    m.def("fill_dummy", []() { return std::string("phash imported OK"); });

    py::class_<SimParam>(m, "SimParam")
        .def(py::init<float, float, float, int, float>(), py::arg("dt") = 0.01f,
             py::arg("Lx") = 1.f, py::arg("Ly") = 1.f, py::arg("resX") = 200,
             py::arg("hwr") = 16.f)
        .def_readwrite("dt", &SimParam::dt)
        .def_readwrite("Lx", &SimParam::Lx)
        .def_readwrite("Ly", &SimParam::Ly)
        .def_readwrite("ResolutionX", &SimParam::ResolutionX)
        .def_readwrite("hwr", &SimParam::hwr)
        .def_readwrite("N_max", &SimParam::N_max)
        .def_readwrite("N1", &SimParam::N1)
        .def_readwrite("N2", &SimParam::N2)
        .def_readwrite("ResolutionY", &SimParam::ResolutionY)
        .def_readwrite("h", &SimParam::h)
        .def_readwrite("h_w", &SimParam::h_w)
        .def_readwrite("length_iCDF", &SimParam::length_iCDF)
        .def_readwrite("dx", &SimParam::dx)
        .def_readwrite("periodic", &SimParam::periodic);

    py::enum_<EstimatorMode>(m, "EstimatorMode")
        .value("Reference", EstimatorMode::Reference)
        .value("Consistent", EstimatorMode::Consistent);

    py::class_<Simulation>(m, "Simulation")
        .def(py::init<SimParam>())
        .def("__copy__", [](const Simulation &s) { return Simulation(s); })
        .def("__deepcopy__", [](const Simulation &s, py::dict) { return Simulation(s); })
        .def_readwrite("P", &Simulation::P_)
        .def_readwrite("mode", &Simulation::mode)
        .def_readwrite("seed", &Simulation::seed)
        .def_readwrite("sim_id", &Simulation::sim_id)
        .def_readwrite("step_index", &Simulation::step_index)
        .def_readwrite("cells_per_hw", &Simulation::cells_per_hw)
        .def_readwrite("kernel_deformation", &Simulation::kernel_deformation)
        .def_readwrite("deformation_bound", &Simulation::deformation_bound)
        .def_readonly("cg_iterations", &Simulation::cg_iterations)
        .def_readonly("cg_rel_residual", &Simulation::cg_rel_residual)
        .def("init", &Simulation::init, py::arg("num_of_p"))
        .def("init_with",
             [](Simulation &s, const f32_array &pos, const f32_array &b, float tol, int max_iter) {
                 s.init_with(as_float2(pos, "init_with"), as_floats(b), tol, max_iter);
             },
             py::arg("pos"), py::arg("b"), py::arg("tol") = 1e-6f, py::arg("max_iter") = 1000)
        .def("set_particles",
             [](Simulation &s, const f32_array &pos, const f32_array &w) {
                 s.set_particles(as_float2(pos, "set_particles"), as_floats(w));
             },
             py::arg("pos"), py::arg("omega"))
        .def("do_spatial_hashing", &Simulation::do_spatial_hashing)
        .def("build_global_hash", &Simulation::build_global_hash)
        .def("build_subsets_cpu", &Simulation::build_subsets_cpu)
        .def("build_subsets_cuda", &Simulation::build_subsets_cuda)
        .def("prepare_rhs_b", &Simulation::prepare_rhs_b)
        .def("step_cpu", &Simulation::step_cpu, py::arg("periodic") = true)
        .def("step_cuda", &Simulation::step_cuda, py::arg("periodic") = true)
        .def("compute_velocity_naive", &Simulation::compute_velocity_naive,
             py::arg("periodic"), py::arg("eps_core"))
        .def("step_cuda_naive", &Simulation::step_cuda_naive, py::arg("periodic") = true)
        .def("estimate_velocity", &Simulation::estimate_velocity, py::arg("periodic") = true)
        .def("eval_vorticity_cpu",
             [](const Simulation &s, float x, float y, bool periodic) {
                 return s.eval_vorticity_cpu(make_float2(x, y), periodic);
             },
             py::arg("x"), py::arg("y"), py::arg("periodic") = true)
        .def("test_vorticity_grid", &Simulation::test_vorticity_grid, py::arg("res"))
        .def("reinitialize", &Simulation::reinitialize, py::arg("nx_l") = 0, py::arg("ny_l") = 0,
             py::arg("tol") = 1e-6f, py::arg("max_iter") = 1000)
        .def("subset_count", &Simulation::subset_count, py::arg("sign"))
        .def_property_readonly("N", [](const Simulation &s) { return s.particles_.N_cur; })
        .def("positions",
             [](const Simulation &s) { return float2_array(s.particles_.pos, s.particles_.N_cur); })
        .def("velocities",
             [](const Simulation &s) { return float2_array(s.particles_.vel, s.particles_.N_cur); })
        .def("omega",
             [](const Simulation &s) { return flat_array(s.particles_.omega, s.particles_.N_cur); });

    // include/cg_rbf_init.cuh:6-12 (the solve, then subsets and hashing)
    m.def("conjugate_gradient_rbf", &conjugate_gradient_rbf, py::arg("sim"),
          py::arg("tol") = 5e-3f, py::arg("max_iter") = 1000);

    m.def(
        "rollout",
        [](Simulation &sim, const Tensor &pos0, const Tensor &w, int64_t steps,
           int64_t checkpoint, const std::string &trajectory) {
            if (trajectory != "f64" && trajectory != "f32")
                throw std::invalid_argument("rollout: trajectory must be 'f64' or 'f32'");
            return RolloutFunction::apply(pos0, w, &sim, steps, checkpoint, trajectory == "f64");
        },
        py::arg("sim"), py::arg("pos0"), py::arg("w"), py::arg("steps"),
        py::arg("checkpoint") = 10, py::arg("trajectory") = "f64",
        "positions after `steps` steps from pos0 (device (N, 2)) with strengths w (device "
        "(N)); a torch.autograd.Function: backward runs the hand-written step adjoint "
        "(vpmc_step_adjoint) over checkpointed recomputation.  trajectory='f64' (default): "
        "the FP64 forward of the frozen-sample step, positions returned in FP64, gradients "
        "to ~1e-5 of the exact derivative; 'f32': step_cuda(true), ~3x cheaper, the FP32 "
        "rounding of the positions limits the gradient to ~1e-4.  The simulation must "
        "outlive the backward pass.  With kernel_deformation the state is (positions, F): "
        "the rollout starts from the simulation's current F (vpmc_step_adjoint_deformed).");
}
