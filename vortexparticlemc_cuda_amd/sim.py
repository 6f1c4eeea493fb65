"""Python mirror of the reference's Simulation / SimParam for the GPU path.

SimParam follows include/vortex_particle_mc.hpp:96-134 (member defaults, then
the constructor's defaults, which win: dt 0.01, hwr 16, ResolutionX 200).
Simulation holds device buffers (torch tensors) and drives libvpmc.so:
init (positions, CG-RBF strengths), build_subsets, do_spatial_hashing, step.
"""
import ctypes as C

import numpy as np
import torch

from . import ops
from ._capi import (ALGO_BINNED, ALGO_DIRECT, MODE_CONSISTENT, MODE_REFERENCE, McDesc, SignSet,
                    check, lib)

f32 = np.float32


class SimParam:
    """include/vortex_particle_mc.hpp:96-134 (FP32 derived values)"""

    def __init__(self, dt=0.01, Lx=1.0, Ly=1.0, resX=200, hwr=16.0):
        self.dt = f32(dt)
        self.Lx = f32(Lx)
        self.Ly = f32(Ly)
        self.ResolutionX = int(resX)
        self.hwr = f32(hwr)
        self.N_max = 6000
        self.N1 = 1000
        self.N2 = 1000
        self.length_iCDF = 10
        self.periodic = True
        assert self.ResolutionX > 0, "ResolutionX must be positive"
        self.h = f32(self.Lx / f32(self.ResolutionX))
        self.ResolutionY = int(np.round(self.Ly / self.h)) or 1
        self.h_w = f32(self.h * self.hwr)
        self.dx = f32(f32(0.5) * self.h_w)

    @property
    def rsh(self):
        """reference hash grid (src/vortex_particle_mc.cu:125-126)"""
        return (self.ResolutionX // int(self.hwr) // 2, self.ResolutionY // int(self.hwr) // 2)


def lattice_positions(nx, ny=None, L=1.0):
    """cell-centred lattice, particle index i*Ny + j"""
    ny = nx if ny is None else ny
    i, j = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")
    x = ((i + 0.5) * (L / nx)).ravel()
    y = ((j + 0.5) * (L / ny)).ravel()
    return np.stack([x, y], 1).astype(np.float32)


class Simulation:
    """One periodic simulation on the device (the reference's Simulation,
    include/vortex_particle_mc.hpp:136-192, for its GPU path)."""

    def __init__(self, P, device="cuda", cells_per_hw=8, mode=MODE_CONSISTENT, seed=0, sim=0,
                 algo=ALGO_BINNED, spl=0, eval_threads=0, deform=False, gq_bound=0.0):
        self.P = P
        # kernel deformation (step_cuda's particles.F): F per particle, the
        # fields carry its quadratic forms, each step refreshes it from the
        # five estimates (src/cuda/simulation_step.cu:51-72)
        self.deform = bool(deform)
        self.gq_bound = float(gq_bound)
        self.jac = None
        self.algo = int(algo)
        self.spl = int(spl)
        self.eval_threads = int(eval_threads)
        self.ws = None
        self.device = torch.device(device)
        self.cells_per_hw = int(cells_per_hw)
        self.mode = int(mode)
        self.seed = int(seed)
        self.sim = int(sim)
        self.step_index = 0
        # sample slots per owner chunk of the binned estimator (0: the
        # library's default, 2^30); results do not depend on it
        self.chunk_samples = 0
        self.pos = None
        self.omega = None

    # -- particles ---------------------------------------------------------
    def set_particles(self, pos, omega):
        self.pos = torch.as_tensor(np.asarray(pos, np.float32), device=self.device).contiguous()
        self.omega = torch.as_tensor(np.asarray(omega, np.float32),
                                     device=self.device).contiguous()
        self.N = self.pos.shape[0]
        self.build_subsets()
        self._alloc()

    def init_strengths_cg(self, pos, b, tol=1e-6, max_iter=1000, refine=True, lattice=None):
        """Simulation::init + conjugate_gradient_rbf (src/vortex_particle_mc.cu:77-120,
        src/cuda/cg_rbf_init.cu:53-116): strengths w solving A w = b.  With refine
        the solve is FP64 MINRES; strengths are FP32, so with an FP32 b, tol below
        ~1e-7 only amplifies the rounding noise of b in the Gram matrix's
        near-null modes.  lattice=(nx, ny): the particles are that cell-centred
        lattice; b is taken in FP64 and MINRES uses the lattice's circulant
        preconditioner."""
        pos_t = torch.as_tensor(np.asarray(pos, np.float32), device=self.device).contiguous()
        dt = np.float64 if lattice is not None else np.float32
        b_t = torch.as_tensor(np.asarray(b, dt), device=self.device).contiguous()
        w, it, rr, st = ops.cg_rbf(pos_t, b_t, float(self.P.h_w), float(self.P.Lx),
                                   float(self.P.Ly), tol=tol, max_iter=max_iter, refine=refine,
                                   cells_per_hw=self.cells_per_hw, lattice=lattice)
        self.cg_iters, self.cg_rel_res, self.cg_status = it, rr, st
        self.set_particles(pos, w.cpu().numpy())
        return it, rr, st

    # -- subsets / CDF (build_subsets_cuda) --------------------------------
    def build_subsets(self):
        s = ops.build_subsets(self.omega)
        self.sub = s
        counts = s["counts"].cpu().numpy()
        self.cnt = (int(counts[0]), int(counts[1]))
        self.glog2 = ops.guide_log2_for(self.N)
        self.guide = [ops.build_guide(s["cdf_pos"], s["counts"][0:1], self.glog2),
                      ops.build_guide(s["cdf_neg"], s["counts"][1:2], self.glog2)]

    def _alloc(self):
        P = self.P
        self.nx, self.ny = ops.field_grid(float(P.h_w), float(P.Lx), float(P.Ly),
                                          self.cells_per_hw)
        self.fields = [ops.FieldBuffers(self.nx, self.ny, P.Lx, P.Ly, self.cnt[k], self.device,
                                        deformed=self.deform, gq_bound=self.gq_bound)
                       for k in range(2)]
        if self.deform and (self.jac is None or self.jac.shape[0] != max(self.N, 1)):
            self.jac = torch.zeros((max(self.N, 1), 4), dtype=torch.float32, device=self.device)
            self.jac[:, 0] = 1.0   # F = I
            self.jac[:, 3] = 1.0
        self.vel_sign = torch.zeros((2 * max(self.N, 1), 2), dtype=torch.float32,
                                    device=self.device)
        self.vel = torch.zeros((max(self.N, 1), 2), dtype=torch.float32, device=self.device)
        self.eval_count = torch.zeros(2, dtype=torch.int64, device=self.device)
        self.vel_aux = None

    def desc(self, N1, N2, uniforms=None, aux_draws=0, count=False, aux=False):
        P = self.P
        sets = []
        names = ("pos", "neg")
        for k in range(2):
            fb = self.fields[k]
            sets.append(SignSet(self.cnt[k], ops._p(self.sub["counts"][k:k + 1]),
                                ops._p(self.sub["cdf_" + names[k]]),
                                ops._p(self.sub["totals"][k:k + 1]), ops._p(self.guide[k]),
                                self.glog2, ops._p(fb.set_xy), fb.struct(),
                                ops._p(fb.set_jac)))
        self._U = None if uniforms is None else torch.as_tensor(
            uniforms, dtype=torch.float32, device=self.device).contiguous()
        d = McDesc()
        d.mode = self.mode
        d.N1, d.N2 = int(N1), int(N2)
        d.h_w, d.Lx, d.Ly = float(P.h_w), float(P.Lx), float(P.Ly)
        d.periodic = int(P.periodic)
        d.seed = self.seed
        d.step = self.step_index
        d.sim = self.sim
        d.uniforms = ops._p(self._U) if self._U is not None else None
        d.aux_draws = int(aux_draws)
        d.n_owner = self.N
        d.owner_xy = ops._p(self.pos)
        d.set[0], d.set[1] = sets[0], sets[1]
        d.vel_sign = ops._p(self.vel_sign)
        d.eval_count = ops._p(self.eval_count) if count else None
        d.algo = self.algo
        d.spl = self.spl
        d.eval_threads = self.eval_threads
        if aux:   # the four targets at dst -/+ dx (SimParam::dx = h_w / 2)
            if self.vel_aux is None:
                self.vel_aux = torch.zeros((8 * max(self.N, 1), 2), dtype=torch.float32,
                                           device=self.device)
            d.aux = 1
            d.aux_dx = float(P.dx)
            d.vel_aux = ops._p(self.vel_aux)
        d.jac = ops._p(self.jac) if self.deform else None
        d.chunk_samples = int(self.chunk_samples)
        return d

    def _ws_for(self, d):
        need = lib().vpmc_step_workspace_bytes(C.byref(d))
        if self.ws is None or self.ws.numel() < need:
            self.ws = ops.workspace(need, self.device)
        return self.ws

    def launch_estimate(self, d):
        """the estimator kernels for descriptor d (fields already built)"""
        if self.algo == ALGO_BINNED:
            ws = self._ws_for(d)
            check(lib().vpmc_mc_estimate_binned(C.byref(d), ops._p(ws), ws.numel(),
                                                ops._stream()), "vpmc_mc_estimate_binned")
        else:
            check(lib().vpmc_mc_estimate(C.byref(d), ops._stream()), "vpmc_mc_estimate")

    def rebuild_fields(self):
        """do_spatial_hashing on the fine grid for both sign sets"""
        ws = ops.workspace(lib().vpmc_field_workspace_bytes(max(self.cnt), self.nx, self.ny),
                           self.device)
        for k, key in enumerate(("sub_pos", "sub_neg")):
            ops.field_build(self.fields[k], self.pos, self.omega, sub=self.sub[key], abs_w=True,
                            ws=ws, jac=self.jac if self.deform else None)

    def estimate(self, N1, N2, uniforms=None, aux_draws=0, count=False, aux=False):
        """velocity estimate at the current positions (no advection); with aux
        also at the four targets dst -/+ dx (self.vel_aux, [sign][L,R,T,B][owner])"""
        self.rebuild_fields()
        d = self.desc(N1, N2, uniforms, aux_draws, count, aux)
        self.launch_estimate(d)
        vs = self.vel_sign[:2 * self.N].view(2, self.N, 2)
        return vs[0] - vs[1]

    def aux_velocities(self):
        """u = v+ - v- at the targets L, R, T, B of the last aux estimate: (4, N, 2)"""
        va = self.vel_aux[:8 * self.N].view(2, 4, self.N, 2)
        return va[0] - va[1]

    def update_jacobian(self, dt=None, jac=None):
        """update_jacobian_from5 (src/cuda/monte_carlo_bs.cu:274-288) from the
        last aux estimate: F = renorm(I + dt/(2 dx) [uR - uL | uT - uB]), (N, 4)"""
        dt = float(self.P.dt if dt is None else dt)
        if jac is None:
            jac = torch.empty((max(self.N, 1), 4), dtype=torch.float32, device=self.device)
        check(lib().vpmc_update_jacobian(self.N, ops._p(self.vel_sign), ops._p(self.vel_aux), dt,
                                         float(self.P.dx), ops._p(jac), ops._stream()),
              "vpmc_update_jacobian")
        return jac[:self.N]

    def reinit(self, nx=None, ny=None, tol=1e-6, max_iter=1000):
        """reinitialisation (README.md:319-321, 416-420): particles to the
        nx x ny cell-centred lattice (default: ResolutionX x ResolutionY),
        strengths re-fitted to the current (deformed) field by the RBF solve,
        F reset to I; subsets and fields rebuilt.  Returns (iterations,
        relative FP64 residual, status)."""
        P = self.P
        nx = P.ResolutionX if nx is None else int(nx)
        ny = P.ResolutionY if ny is None else int(ny)
        pos, w, it, rr, st = ops.reinit(self.pos, self.omega, float(P.h_w), nx, ny,
                                        jac=self.jac[:self.N] if self.deform else None,
                                        Lx=float(P.Lx), Ly=float(P.Ly), tol=tol,
                                        max_iter=max_iter, cells_per_hw=self.cells_per_hw)
        self.jac = None   # _alloc makes F = I for the new particle count
        self.set_particles(pos.cpu().numpy(), w.cpu().numpy())
        return it, rr, st

    def set_jacobian(self, jac):
        """the deformation F per particle, (N, 4) row-major (F.x F.y; F.z F.w)"""
        if not self.deform:
            raise ValueError("set_jacobian: Simulation(deform=True) carries F")
        j = torch.as_tensor(np.asarray(jac, np.float32), device=self.device).reshape(-1, 4)
        if j.shape[0] != self.N:
            raise ValueError("set_jacobian: one F per particle")
        self.jac[:self.N] = j

    def step(self, N1=None, N2=None, dt=None, uniforms=None, aux_draws=0):
        """one step: fields, estimator, advection (step_cuda); a deformed
        simulation also estimates the four aux targets and replaces F by
        update_jacobian_from5 of them before the move"""
        N1 = self.P.N1 if N1 is None else N1
        N2 = self.P.N2 if N2 is None else N2
        dt = float(self.P.dt if dt is None else dt)
        d = self.desc(N1, N2, uniforms, aux_draws, aux=self.deform)
        self._ws_for(d)
        check(lib().vpmc_step(C.byref(d), self.N, ops._p(self.pos), ops._p(self.omega),
                              ops._p(self.sub["sub_pos"]), ops._p(self.sub["sub_neg"]),
                              ops._p(self.vel), dt, ops._p(self.ws), self.ws.numel(),
                              ops._stream()), "vpmc_step")
        self.step_index += 1
        return self.vel[:self.N]
