"""C ABI library: loads without a GPU, exports every declared symbol, host-side
queries behave.  CPU only (no kernel launches)."""
import ctypes as C
import os
import re

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    src = open(os.path.join(ROOT, "include", "vpmc.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(vpmc_[a-z0-9_]+)\s*\(", src)))


def test_library_loads_and_exports_everything():
    from vortexparticlemc_cuda_amd import _capi
    L = _capi.lib()
    names = _declared()
    assert len(names) >= 20
    for n in names:
        assert hasattr(L, n), n
    assert sorted(_capi.EXPORTS) == names
    assert L.vpmc_version() >= 100
    assert L.vpmc_status_string(0) == b"ok"


def test_library_has_gfx950_code():
    path = os.path.join(ROOT, "vortexparticlemc_cuda_amd", "lib", "libvpmc.so")
    blob = open(path, "rb").read()
    assert b"gfx950" in blob


def test_grid_and_workspace_queries():
    from vortexparticlemc_cuda_amd import _capi
    L = _capi.lib()
    nx, ny = C.c_int(), C.c_int()
    assert L.vpmc_field_grid(0.08, 1.0, 1.0, 8, C.byref(nx), C.byref(ny)) == 0
    assert (nx.value, ny.value) == (100, 100)
    assert L.vpmc_field_grid(0.0078125, 1.0, 1.0, 4, C.byref(nx), C.byref(ny)) == 0
    assert nx.value == 512
    # support wider than the box can hold without overlap -> grid error
    assert L.vpmc_field_grid(0.6, 1.0, 1.0, 2, C.byref(nx), C.byref(ny)) == 4
    assert L.vpmc_field_grid(-1.0, 1.0, 1.0, 2, C.byref(nx), C.byref(ny)) == 1
    assert L.vpmc_hash_workspace_bytes(1000, 36) >= 4000
    assert L.vpmc_field_workspace_bytes(40000, 100, 100) >= 2 * 40000 * 4 + 10000 * 4
    assert L.vpmc_subsets_workspace_bytes(40000) >= 40000 * 24
    assert L.vpmc_cg_workspace_bytes(40000, 100, 100) > 40000 * 16


def test_argument_errors_need_no_gpu():
    """invalid arguments are rejected before any device work"""
    from vortexparticlemc_cuda_amd import _capi
    L = _capi.lib()
    assert L.vpmc_fill_cells(0, 4, 1.0, 1.0, 10, None, None, None, None, None, None, 0,
                             None) == 1
    assert L.vpmc_build_guide(None, None, 3, None, None) == 1
    assert L.vpmc_mc_estimate(None, None) == 1
    d = _capi.McDesc()
    d.mode = 7
    assert L.vpmc_mc_estimate(C.byref(d), None) == 1


def test_sim_param_semantics():
    """include/vortex_particle_mc.hpp:96-134: the constructor's defaults win
    over the member initialisers"""
    from vortexparticlemc_cuda_amd.sim import SimParam
    P = SimParam()
    assert P.dt == np.float32(0.01) and P.hwr == np.float32(16) and P.ResolutionX == 200
    assert P.h == np.float32(0.005) and P.ResolutionY == 200
    assert P.h_w == np.float32(0.005) * np.float32(16) and P.dx == np.float32(0.5) * P.h_w
    assert P.rsh == (6, 6) and P.N1 == 1000 and P.N_max == 6000
    P = SimParam(dt=1e-3, resX=1024, hwr=8)
    assert P.rsh == (64, 64) and abs(float(P.h_w) - 8 / 1024) < 1e-9


def test_oracle_not_imported_by_product():
    """the product package never imports the oracle"""
    pkg = os.path.join(ROOT, "vortexparticlemc_cuda_amd")
    for dirpath, _, files in os.walk(pkg):
        for f in files:
            if f.endswith((".py", ".hip", ".hpp", ".cpp", ".h")):
                text = open(os.path.join(dirpath, f)).read()
                assert "import oracle" not in text and "from oracle" not in text, f
                assert "vpmc_oracle" not in text and "liboracle" not in text, f


def test_chunked_workspace_bounds_the_large_reference_step():
    """the aux + F step of the large run (1,048,576 owners, N1 = N2 = 256:
    3.2e9 sample slots) gets a finite workspace, bounded by the owner chunk
    (chunk_samples; default 2^30 slots), and a smaller chunk a smaller one.
    Host-side query only: the device pointers are never dereferenced."""
    from vortexparticlemc_cuda_amd import _capi
    L = _capi.lib()
    hw = 8.0 / 1024
    nx, ny = C.c_int(), C.c_int()
    assert L.vpmc_field_grid(hw, 1.0, 1.0, 8, C.byref(nx), C.byref(ny)) == 0
    fake = C.c_void_p(256)
    d = _capi.McDesc()
    d.mode, d.N1, d.N2, d.h_w, d.Lx, d.Ly, d.periodic = 1, 256, 256, hw, 1.0, 1.0, 1
    d.n_owner = 1 << 20
    d.owner_xy = d.vel_sign = d.vel_aux = fake
    d.aux, d.aux_dx = 1, hw / 2
    for s in range(2):
        S = d.set[s]
        S.n, S.cdf, S.total_dev, S.guide, S.guide_log2, S.set_xy = 1 << 19, fake, fake, fake, 20, fake
        S.field = _capi.Field(nx.value, ny.value, 1.0, 1.0, 1 << 19, fake, fake, fake, None, 0.0)
    full = L.vpmc_mc_binned_workspace_bytes(C.byref(d))
    assert 0 < full < 64 << 30
    d.chunk_samples = 1 << 24
    small = L.vpmc_mc_binned_workspace_bytes(C.byref(d))
    assert 0 < small < full // 16
    d.chunk_samples = 2 * (256 + 5 * 256) - 1   # one owner's samples exceed a chunk
    assert L.vpmc_mc_binned_workspace_bytes(C.byref(d)) == 0


def test_phash_module_loads():
    """the bindings module (bindings/phash.cpp:6-14) imports without a GPU,
    keeps fill_dummy, and its SimParam has the reference's constructor
    precedence (include/vortex_particle_mc.hpp:96-134)"""
    import vortexparticlemc_cuda_amd as v
    phash = v.phash()
    import phash as again
    assert again is phash
    assert phash.fill_dummy() == "phash imported OK"
    P = phash.SimParam()
    assert abs(P.dt - 0.01) < 1e-9 and P.ResolutionX == 200 and P.hwr == 16.0
    assert P.N1 == 1000 and P.N_max == 6000 and P.periodic
    P = phash.SimParam(dt=1e-3, resX=100, hwr=16)
    assert abs(P.h - 0.01) < 1e-7 and abs(P.h_w - 0.16) < 1e-6 and abs(P.dx - 0.08) < 1e-7
    for name in ("step_cuda", "step_cpu", "eval_vorticity_cpu", "do_spatial_hashing",
                 "estimate_velocity", "reinitialize"):
        assert hasattr(phash.Simulation, name)
    assert callable(phash.rollout) and callable(phash.conjugate_gradient_rbf)
