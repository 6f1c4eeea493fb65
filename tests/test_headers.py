"""Every public header of the C++ drop-in compiles on its own (a translation
unit that includes only it), as the reference's callers include them."""
import glob
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HEADERS = sorted(os.path.basename(p) for p in glob.glob(os.path.join(ROOT, "include", "*"))
                 if p.endswith((".h", ".hpp", ".cuh")))


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not installed")
@pytest.mark.parametrize("header", HEADERS)
def test_header_is_self_contained(header, tmp_path):
    src = tmp_path / "tu.hip"
    src.write_text('#include "%s"\nint main() { return 0; }\n' % header)
    r = subprocess.run([HIPCC, "--offload-arch=gfx950", "-std=c++20", "-fsyntax-only",
                        "-I", os.path.join(ROOT, "include"), str(src)],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]


# A caller written against the reference's interface: MCBSCtx filled with
# the member types of include/monte_carlo_bs.hpp:15-35, as
# src/cpu/simulation_step.cpp:13-27 fills it; Simulation copied and queried
# with eval_vorticity_cpu (include/vortex_particle_mc.hpp:160,204-208);
# ParticleSet's pointer getters called from device code (:55-71).
REFERENCE_CALLER = r'''
#include <type_traits>
#include "monte_carlo_bs.hpp"

static_assert(std::is_same_v<decltype(MCBSCtx::cdf_pos), const dvec<float> *>);
static_assert(std::is_same_v<decltype(MCBSCtx::cdf_neg), const dvec<float> *>);
static_assert(std::is_same_v<decltype(MCBSCtx::cum_pos), float>);
static_assert(std::is_same_v<decltype(MCBSCtx::cum_neg), float>);
static_assert(std::is_copy_constructible_v<Simulation>);
static_assert(std::is_copy_assignable_v<Simulation>);

MCBSCtx fill_like_step_cpu(const dvec<int> &sub_pos_, const dvec<int> &sub_neg_,
                           const dvec<float> &cdf_pos_, const dvec<float> &cdf_neg_,
                           const dvec<float> &icdf_kernel, float cum_pos_, float cum_neg_,
                           const SimParam &P_, bool periodic, const Simulation &sim)
{
    MCBSCtx ctx;
    ctx.sub_pos = &sub_pos_;
    ctx.sub_neg = &sub_neg_;
    ctx.cdf_pos = &cdf_pos_;
    ctx.cdf_neg = &cdf_neg_;
    ctx.icdf_kernel = &icdf_kernel;
    ctx.kernel_L = 0;
    ctx.h_w = P_.h_w;
    ctx.periodic = periodic;
    ctx.Lx = P_.Lx;  ctx.Ly = P_.Ly;
    ctx.cum_pos = cum_pos_;
    ctx.cum_neg = cum_neg_;
    ctx.pos_view = sim.makePosView();
    ctx.neg_view = sim.makeNegView();
    return ctx;
}

float query(const Simulation &s)
{
    Simulation copy = s;
    return copy.eval_vorticity_cpu(make_float2(0.5f, 0.5f), true);
}

__device__ const float2 *device_positions(const ParticleSet &p) { return p.d_p(); }
__device__ float *device_strengths(ParticleSet &p) { return p.d_omega(); }

int main() { return 0; }
'''


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not installed")
def test_reference_style_caller_compiles(tmp_path):
    src = tmp_path / "caller.hip"
    src.write_text(REFERENCE_CALLER)
    r = subprocess.run([HIPCC, "--offload-arch=gfx950", "-std=c++20", "-fsyntax-only",
                        "-I", os.path.join(ROOT, "include"), str(src)],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-3000:]
