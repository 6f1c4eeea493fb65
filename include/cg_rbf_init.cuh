// cg_rbf_init.cuh -- matrix-free RBF Gram operator and CG (device).
// Replaces include/cg_rbf_init.cuh:6-12 (cudaStream_t becomes hipStream_t).
#pragma once

#include "vortex_particle_mc.hpp"

// y_i = sum_j x_j W(|x_i - x_j|) over all particles (x, y device, particle order)
void Ax_cuda(const Simulation &sim, const float *x, float *y, hipStream_t stream = 0);

// fits particles_.omega to particles_.omega_field (FP64 MINRES: the Gram
// matrix is indefinite, where the reference's FP32 CG can break down), then
// rebuilds subsets and hashes
void conjugate_gradient_rbf(Simulation &sim, float tol = 5e-3f, int max_iter = 1000);
