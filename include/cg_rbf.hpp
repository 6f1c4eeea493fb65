// cg_rbf.hpp -- RBF strength fit, host entry points.
// Replaces include/cg_rbf.hpp:4-10.  The reference's CPU solver
// (src/cpu/cg_rbf.cpp) does not build; here solve_cg_rbf_cpu keeps its name
// and arguments and solves A w = omega_field on the device with the
// semantics of the reference's working CG (src/cuda/cg_rbf_init.cu:53-116:
// diagonal included), stopping at ||r|| / ||b|| < rel_tol.  The solve is
// FP64 MINRES, which also holds where the Gram matrix is indefinite and the
// reference's CG would meet p.Ap <= 0; alpha_tol is accepted and unused.
#pragma once

#include "vortex_particle_mc.hpp"

void build_global_hash_all(Simulation &sim);

void solve_cg_rbf_cpu(Simulation &sim, float h_w, float rel_tol = 3e-3f,
                      float alpha_tol = 1e-8f, int max_iter = 1000);
