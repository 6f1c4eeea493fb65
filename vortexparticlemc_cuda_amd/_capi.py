"""ctypes bindings for libvpmc.so (include/vpmc.h).

The library is loaded from vortexparticlemc_cuda_amd/lib/libvpmc.so (built
in-tree by _build.py).  There is no fallback: if the library is missing or a
call fails, the error is raised.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libvpmc.so")

VPMC_OK = 0
MODE_REFERENCE = 0
MODE_CONSISTENT = 1
ALGO_BINNED = 0
ALGO_DIRECT = 1

vp = C.c_void_p
ci = C.c_int
cf = C.c_float
cd = C.c_double
csz = C.c_size_t


class VpmcError(RuntimeError):
    def __init__(self, status, where):
        self.status = status
        msg = "%s failed: %s (status %d)" % (where, _lib.vpmc_status_string(status).decode(),
                                              status)
        if status == 2:
            msg += ", hipError %d" % _lib.vpmc_last_hip_error()
        super().__init__(msg)


class View(C.Structure):
    _fields_ = [("pos_xy", vp), ("w", vp), ("N", ci), ("vp_sort", vp), ("cell_acc", vp),
                ("rshx", ci), ("rshy", ci), ("Lx", cf), ("Ly", cf), ("cnt_view", ci),
                ("sub", vp)]


class Field(C.Structure):
    _fields_ = [("nx", ci), ("ny", ci), ("Lx", cf), ("Ly", cf), ("n", ci), ("pts", vp),
                ("cell_start", vp), ("order", vp), ("gq", vp), ("gq_bound", cf)]


class SignSet(C.Structure):
    _fields_ = [("n", ci), ("count_dev", vp), ("cdf", vp), ("total_dev", vp), ("guide", vp),
                ("guide_log2", ci), ("set_xy", vp), ("field", Field), ("set_jac", vp)]


class McDesc(C.Structure):
    _fields_ = [("mode", ci), ("N1", ci), ("N2", ci), ("h_w", cf), ("Lx", cf), ("Ly", cf),
                ("periodic", ci), ("seed", C.c_uint64), ("step", C.c_uint32),
                ("sim", C.c_uint32), ("uniforms", vp), ("aux_draws", ci), ("n_owner", ci),
                ("owner_xy", vp), ("set", SignSet * 2), ("vel_sign", vp),
                ("eval_count", vp), ("algo", ci), ("eval_events", vp * 2), ("spl", ci),
                ("eval_threads", ci), ("aux", ci), ("aux_dx", cf), ("vel_aux", vp),
                ("jac", vp), ("owner_base", ci), ("chunk_samples", C.c_uint64)]


class EnsembleDesc(C.Structure):
    _fields_ = [("n_sim", ci), ("h_offset", vp), ("h_h_w", vp), ("h_L", vp), ("N1", ci),
                ("N2", ci), ("mode", ci), ("cells_per_hw", ci), ("spl", ci),
                ("eval_threads", ci), ("seed", C.c_uint64), ("dt", cf), ("pos_xy", vp), ("omega", vp),
                ("vel_xy", vp), ("eval_events", vp * 2), ("jac", vp)]


_lib = None


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError("libvpmc.so is not built (%s); run "
                          "`python -c 'import __graft_entry__ as g; g.build()'`" % LIB_PATH)
    L = C.CDLL(LIB_PATH)
    L.vpmc_status_string.restype = C.c_char_p
    L.vpmc_status_string.argtypes = [ci]
    L.vpmc_last_hip_error.restype = ci
    L.vpmc_version.restype = ci
    L.vpmc_hash_workspace_bytes.restype = csz
    L.vpmc_hash_workspace_bytes.argtypes = [ci, ci]
    L.vpmc_fill_cells.argtypes = [ci, ci, cf, cf, ci, vp, vp, vp, vp, vp, vp, csz, vp]
    L.vpmc_counting_sort.argtypes = [ci, ci, vp, vp, vp, vp, vp, csz, vp]
    L.vpmc_subsets_workspace_bytes.restype = csz
    L.vpmc_subsets_workspace_bytes.argtypes = [ci]
    L.vpmc_build_subsets.argtypes = [ci, vp, vp, vp, vp, vp, vp, vp, vp, csz, vp]
    L.vpmc_build_guide.argtypes = [vp, vp, ci, vp, vp]
    L.vpmc_sample_idx.argtypes = [ci, vp, vp, vp, vp, ci, vp, vp]
    L.vpmc_query_view.argtypes = [C.POINTER(View), ci, vp, ci, ci, ci, cf, ci, vp, vp]
    L.vpmc_field_grid.argtypes = [cf, cf, cf, ci, C.POINTER(ci), C.POINTER(ci)]
    L.vpmc_field_workspace_bytes.restype = csz
    L.vpmc_field_workspace_bytes.argtypes = [ci, ci, ci]
    L.vpmc_field_build.argtypes = [ci, ci, cf, cf, ci, vp, vp, vp, ci, vp, vp, vp, vp, vp, csz,
                                   vp]
    L.vpmc_probe_fma.argtypes = [ci, ci, vp, vp]
    L.vpmc_cg_lattice_workspace_bytes.restype = csz
    L.vpmc_cg_lattice_workspace_bytes.argtypes = [ci, ci, cf, cf, cf, ci]
    L.vpmc_cg_rbf_lattice.argtypes = [ci, ci, vp, vp, vp, vp, cf, cf, cf, ci, C.c_double, ci,
                                      C.POINTER(ci), C.POINTER(C.c_double), vp, csz, vp]
    L.vpmc_reinit_workspace_bytes.restype = csz
    L.vpmc_reinit_workspace_bytes.argtypes = [ci, ci, ci, cf, cf, cf, ci, ci]
    L.vpmc_reinit.argtypes = [ci, vp, vp, vp, cf, cf, cf, ci, ci, ci, C.c_double, ci, vp, vp,
                              C.POINTER(ci), C.POINTER(C.c_double), vp, csz, vp]
    L.vpmc_field_build_jac.argtypes = [ci, ci, cf, cf, ci, vp, vp, vp, ci, vp, vp, vp, vp, vp,
                                       vp, vp, vp, csz, vp]
    L.vpmc_field_query.argtypes = [C.POINTER(Field), cf, ci, vp, vp, vp]
    L.vpmc_mc_estimate.argtypes = [C.POINTER(McDesc), vp]
    L.vpmc_mc_binned_workspace_bytes.restype = csz
    L.vpmc_mc_binned_workspace_bytes.argtypes = [C.POINTER(McDesc)]
    L.vpmc_mc_estimate_binned.argtypes = [C.POINTER(McDesc), vp, csz, vp]
    L.vpmc_advect.argtypes = [ci, vp, vp, vp, cf, cf, cf, ci, vp]
    L.vpmc_update_jacobian.argtypes = [ci, vp, vp, cf, cf, vp, vp]
    L.vpmc_step_workspace_bytes.restype = csz
    L.vpmc_step_workspace_bytes.argtypes = [C.POINTER(McDesc)]
    L.vpmc_step.argtypes = [C.POINTER(McDesc), ci, vp, vp, vp, vp, vp, cf, vp, csz, vp]
    L.vpmc_step_noalloc.argtypes = [C.POINTER(McDesc), ci, vp, vp, vp, vp, vp, cf, vp]
    L.vpmc_rbf_workspace_bytes.restype = csz
    L.vpmc_rbf_workspace_bytes.argtypes = [ci, ci, ci]
    L.vpmc_rbf_matvec.argtypes = [C.POINTER(Field), cf, vp, vp, vp]
    L.vpmc_rbf_matvec_f64.argtypes = [C.POINTER(Field), cf, vp, vp, vp]
    L.vpmc_rbf_matvec_particles.argtypes = [ci, vp, vp, vp, cf, cf, cf, ci, vp, csz, vp]
    L.vpmc_cg_workspace_bytes.restype = csz
    L.vpmc_cg_workspace_bytes.argtypes = [ci, ci, ci]
    L.vpmc_cg_rbf.argtypes = [ci, vp, vp, vp, cf, cf, cf, ci, cd, ci, ci, C.POINTER(ci),
                              C.POINTER(cd), vp, csz, vp]
    L.vpmc_cg_rbf_f64.argtypes = [ci, vp, vp, vp, vp, cf, cf, cf, ci, cd, ci, C.POINTER(ci),
                                  C.POINTER(cd), vp, csz, vp]
    L.vpmc_direct_sum_eps.argtypes = [ci, vp, ci, ci, vp, vp, cd, ci, cd, vp, vp]
    L.vpmc_direct_sum_blob.argtypes = [ci, vp, ci, vp, vp, cd, cd, vp, vp]
    ep = C.POINTER(EnsembleDesc)
    L.vpmc_ensemble_workspace_bytes.restype = csz
    L.vpmc_ensemble_workspace_bytes.argtypes = [ep]
    L.vpmc_ensemble_init.argtypes = [ep, vp, csz, vp]
    L.vpmc_ensemble_step.argtypes = [ep, vp, csz, vp]
    L.vpmc_ensemble_set_step.argtypes = [ep, vp, csz, C.c_uint32, vp]
    L.vpmc_ensemble_graph_create.argtypes = [ep, vp, csz, ci, vp, C.POINTER(vp)]
    L.vpmc_graph_launch.argtypes = [vp, vp]
    L.vpmc_graph_destroy.argtypes = [vp]
    mp = C.POINTER(McDesc)
    L.vpmc_mc_adjoint_workspace_bytes.restype = csz
    L.vpmc_mc_adjoint_workspace_bytes.argtypes = [mp]
    L.vpmc_mc_adjoint.argtypes = [mp, vp, vp, vp, vp, vp, vp, csz, vp]
    L.vpmc_step_adjoint_workspace_bytes.restype = csz
    L.vpmc_step_adjoint_workspace_bytes.argtypes = [mp]
    L.vpmc_step_adjoint.argtypes = [mp, ci, vp, vp, vp, vp, cf, vp, vp, vp, csz, vp]
    L.vpmc_step_f64_workspace_bytes.restype = csz
    L.vpmc_step_f64_workspace_bytes.argtypes = [mp]
    L.vpmc_step_f64.argtypes = [mp, ci, vp, vp, vp, vp, vp, cd, vp, csz, vp]
    L.vpmc_step_adjoint_deformed_workspace_bytes.restype = csz
    L.vpmc_step_adjoint_deformed_workspace_bytes.argtypes = [mp]
    L.vpmc_step_adjoint_deformed.argtypes = [mp, ci, vp, vp, vp, vp, cf, vp, vp, vp, vp, csz,
                                             vp]
    L.vpmc_step_f64_deformed.argtypes = [mp, ci, vp, vp, vp, vp, vp, cd, vp, vp, csz, vp]
    L.vpmc_grid_vorticity_workspace_bytes.restype = csz
    L.vpmc_grid_vorticity_workspace_bytes.argtypes = [ci, ci, ci]
    L.vpmc_grid_vorticity.argtypes = [ci, vp, vp, cf, cf, cf, ci, ci, ci, vp, vp, csz, vp]
    L.vpmc_grid_vorticity_adjoint.argtypes = [ci, vp, vp, cf, cf, cf, ci, ci, vp, vp, vp, vp]
    L.vpmc_grid_vorticity_f64.argtypes = [ci, vp, vp, cd, cd, cd, ci, ci, vp, vp]
    _lib = L
    return L


def check(status, where):
    if status != VPMC_OK:
        raise VpmcError(status, where)
    return status


# every symbol include/vpmc.h declares (the CPU suite checks the exports)
EXPORTS = [
    "vpmc_status_string", "vpmc_last_hip_error", "vpmc_version", "vpmc_probe_fma",
    "vpmc_hash_workspace_bytes", "vpmc_fill_cells", "vpmc_counting_sort",
    "vpmc_subsets_workspace_bytes", "vpmc_build_subsets", "vpmc_build_guide", "vpmc_sample_idx",
    "vpmc_query_view", "vpmc_field_grid", "vpmc_field_workspace_bytes", "vpmc_field_build",
    "vpmc_field_build_jac", "vpmc_field_query", "vpmc_mc_estimate", "vpmc_mc_binned_workspace_bytes",
    "vpmc_mc_estimate_binned", "vpmc_advect", "vpmc_update_jacobian", "vpmc_step_workspace_bytes",
    "vpmc_step", "vpmc_step_noalloc", "vpmc_rbf_workspace_bytes", "vpmc_rbf_matvec", "vpmc_rbf_matvec_f64", "vpmc_rbf_matvec_particles",
    "vpmc_cg_rbf", "vpmc_cg_rbf_f64", "vpmc_cg_workspace_bytes", "vpmc_cg_lattice_workspace_bytes",
    "vpmc_cg_rbf_lattice", "vpmc_reinit_workspace_bytes",
    "vpmc_reinit", "vpmc_direct_sum_eps", "vpmc_direct_sum_blob",
    "vpmc_ensemble_workspace_bytes", "vpmc_ensemble_init", "vpmc_ensemble_step",
    "vpmc_ensemble_set_step", "vpmc_ensemble_graph_create", "vpmc_graph_launch",
    "vpmc_graph_destroy", "vpmc_mc_adjoint_workspace_bytes", "vpmc_mc_adjoint",
    "vpmc_step_adjoint_workspace_bytes", "vpmc_step_adjoint", "vpmc_step_f64_workspace_bytes",
    "vpmc_step_f64", "vpmc_step_adjoint_deformed_workspace_bytes", "vpmc_step_adjoint_deformed",
    "vpmc_step_f64_deformed", "vpmc_grid_vorticity_workspace_bytes", "vpmc_grid_vorticity",
    "vpmc_grid_vorticity_adjoint", "vpmc_grid_vorticity_f64",
]
