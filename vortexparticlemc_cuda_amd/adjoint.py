"""Differentiable Monte Carlo vortex-particle steps (torch.autograd over the
library's hand-written adjoint kernels).

    mc_velocity(sim, pos, w)                   u = v+ - v- at the particles
    rollout(sim, pos0, w, steps, checkpoint, trajectory)
                                               positions after `steps` steps
    grid_vorticity(pos, w, h_w, gnx, gny)      omega_h on a gnx x gny grid

Derivatives are taken under common random numbers with the samples frozen
(vpmc_mc_adjoint in include/vpmc.h states the estimator that is
differentiated): the Philox stream of step t is keyed by (seed, t), so a
recomputed step replays its samples exactly.  The sampling tables (sign
subsets, |w| CDF, guide) are built from the strengths passed to the forward,
which is where the derivative is taken.

`rollout` keeps the positions every `checkpoint` steps; the backward pass
recomputes each segment forwards from its checkpoint and then runs the step
adjoints backwards through it, so memory grows with steps / checkpoint +
checkpoint states, not with steps.

rollout_f64 / grid_vorticity_f64 evaluate the same frozen-sample functions in
FP64 (vpmc_step_f64, vpmc_grid_vorticity_f64), for finite-difference checks.

A Simulation(deform=True) differentiates the deformed step (aux targets and
F, vpmc_step_adjoint_deformed): F is part of the state, the rollout starts
from sim's current F, and the checkpoints keep it with the positions.
"""
import ctypes as C

import torch

from . import ops
from ._capi import check, lib


def _set_nominal(sim, w):
    """make w the strengths the sampling tables and fields use"""
    w = w.detach().to(torch.float32).contiguous()
    if sim.omega is None or sim.omega.shape != w.shape:
        raise ValueError("strengths must match the simulation's particle count")
    sim.omega.copy_(w)
    counts = sim.cnt
    sim.build_subsets()
    if sim.cnt != counts:   # sign subsets changed size: field buffers follow
        sim._alloc()


def _adj_ws(sim, d, deformed=False):
    fn = (lib().vpmc_step_adjoint_deformed_workspace_bytes if deformed
          else lib().vpmc_step_adjoint_workspace_bytes)
    need = fn(C.byref(d))
    if need == 0:
        raise ValueError("the adjoint needs the consistent, periodic, Philox-driven estimator")
    if getattr(sim, "adj_ws", None) is None or sim.adj_ws.numel() < need:
        sim.adj_ws = ops.workspace(need, sim.device)
    return sim.adj_ws


class _McVelocity(torch.autograd.Function):
    @staticmethod
    def forward(ctx, pos, w, sim):
        sim.pos.copy_(pos.detach())
        _set_nominal(sim, w)
        u = sim.estimate(sim.P.N1, sim.P.N2).clone()
        ctx.sim = sim
        ctx.step = sim.step_index
        ctx.save_for_backward(pos.detach().clone())
        return u

    @staticmethod
    def backward(ctx, gu):
        sim = ctx.sim
        (pos,) = ctx.saved_tensors
        sim.pos.copy_(pos)
        sim.rebuild_fields()
        d = sim.desc(sim.P.N1, sim.P.N2)
        d.step = ctx.step
        gpos = torch.zeros_like(pos)
        gw = torch.zeros(pos.shape[0], dtype=torch.float32, device=pos.device)
        ws = _adj_ws(sim, d)
        check(lib().vpmc_mc_adjoint(C.byref(d), ops._p(sim.sub["sub_pos"]),
                                    ops._p(sim.sub["sub_neg"]),
                                    ops._p(gu.to(torch.float32).contiguous()), ops._p(gpos),
                                    ops._p(gw), ops._p(ws), ws.numel(), ops._stream()),
              "vpmc_mc_adjoint")
        return gpos, gw, None


def mc_velocity(sim, pos, w):
    """the estimator's velocity u = v+ - v- at the particles (owners = the
    particles), differentiable in pos and w; uses sim's step index"""
    return _McVelocity.apply(pos, w, sim)


def _state(sim):
    """the FP32 state a step starts from: positions, and F when deformed"""
    return (sim.pos.clone(), sim.jac.clone() if sim.deform else None)


def _set_state(sim, st):
    sim.pos.copy_(st[0])
    if sim.deform:
        sim.jac.copy_(st[1])


def _run(sim, t0, t1, keep=None):
    """advance sim's state from step t0 to t1 (forward steps); keep: list that
    receives the state before each step"""
    for t in range(t0, t1):
        if keep is not None:
            keep.append(_state(sim))
        sim.step_index = t
        sim.step()
    sim.step_index = t1


def _step_f64(sim, X, w64, t, ws, F64=None):
    """one frozen-sample step of the FP64 positions X in place (vpmc_step_f64);
    a deformed simulation also advances its FP64 F (F64, (N, 4)) in place
    (vpmc_step_f64_deformed)"""
    P = sim.P
    d = sim.desc(P.N1, P.N2, aux=sim.deform)
    d.step = t
    need = lib().vpmc_step_f64_workspace_bytes(C.byref(d))
    if ws is None or ws.numel() < need:
        ws = ops.workspace(need, sim.device)
    args = (C.byref(d), sim.N, ops._p(X), ops._p(w64), ops._p(sim.omega),
            ops._p(sim.sub["sub_pos"]), ops._p(sim.sub["sub_neg"]), float(P.dt))
    if sim.deform:
        check(lib().vpmc_step_f64_deformed(*args, ops._p(F64), ops._p(ws), ws.numel(),
                                           ops._stream()), "vpmc_step_f64_deformed")
    else:
        check(lib().vpmc_step_f64(*args, ops._p(ws), ws.numel(), ops._stream()),
              "vpmc_step_f64")
    return ws


def _run_f64(sim, X, w64, t0, t1, keep=None, F64=None):
    ws = None
    for t in range(t0, t1):
        if keep is not None:
            keep.append((X.clone(), F64.clone() if F64 is not None else None))
        ws = _step_f64(sim, X, w64, t, ws, F64)
    sim.step_index = t1


def _step_adjoint(sim, st, t, lam, gw, jbar=None):
    """one step backwards at state st = (positions, F) of step t: lam (dL/dx')
    becomes dL/dx in place, gw accumulates dL/dw; a deformed simulation also
    turns jbar (dL/dF', (N, 4)) into dL/dF in place"""
    P = sim.P
    _set_state(sim, st)
    d = sim.desc(P.N1, P.N2, aux=sim.deform)
    d.step = t
    ws = _adj_ws(sim, d, sim.deform)
    args = (C.byref(d), sim.N, ops._p(sim.pos), ops._p(sim.omega), ops._p(sim.sub["sub_pos"]),
            ops._p(sim.sub["sub_neg"]), float(P.dt), ops._p(lam))
    if sim.deform:
        check(lib().vpmc_step_adjoint_deformed(*args, ops._p(jbar), ops._p(gw), ops._p(ws),
                                               ws.numel(), ops._stream()),
              "vpmc_step_adjoint_deformed")
    else:
        check(lib().vpmc_step_adjoint(*args, ops._p(gw), ops._p(ws), ws.numel(),
                                      ops._stream()), "vpmc_step_adjoint")


def step_vjp(sim, pos, F, t, lam, jbar=None):
    """the vector-Jacobian product of step t at positions pos (float32 (N, 2))
    and, for a deformed simulation, F (float32 (N, 4)), with sim's strengths
    as the nominal ones: returns (dL/dx, dL/dw, dL/dF) for the cotangents
    lam = dL/dx' and jbar = dL/dF' (dL/dF None when undeformed)"""
    lam = lam.to(torch.float32).contiguous().clone()
    gw = torch.zeros(sim.N, dtype=torch.float32, device=lam.device)
    if sim.deform:
        jbar = (torch.zeros((sim.N, 4), dtype=torch.float32, device=lam.device) if jbar is None
                else jbar.to(torch.float32).contiguous().clone())
        st = (pos.to(torch.float32), F.to(torch.float32))
    else:
        jbar, st = None, (pos.to(torch.float32), None)
    _step_adjoint(sim, st, int(t), lam, gw, jbar)
    return lam, gw, jbar


def step_f64(sim, X, w64, t, F64=None):
    """FP64 frozen-sample step t from positions X (float64 (N, 2)) with
    strengths w64: returns the new positions and, for a deformed simulation,
    the new F from F64 (float64 (N, 4)); the sampling tables and MIS density
    stay those of sim's nominal strengths"""
    X = X.to(torch.float64).contiguous().clone()
    w64 = w64.to(torch.float64).contiguous()
    F = F64.to(torch.float64).contiguous().clone() if sim.deform else None
    _step_f64(sim, X, w64, int(t), None, F)
    return X, F


class _Rollout(torch.autograd.Function):
    @staticmethod
    def forward(ctx, pos0, w, sim, steps, checkpoint, traj64):
        _set_nominal(sim, w)
        K = max(1, int(checkpoint))
        ckpt = {}
        F0 = sim.jac[:sim.N].clone() if sim.deform else None   # F the rollout starts from
        if traj64:
            X = pos0.detach().to(torch.float64).contiguous().clone()
            w64 = w.detach().to(torch.float64).contiguous()
            F64 = F0.to(torch.float64).contiguous() if sim.deform else None
            for t0 in range(0, steps, K):
                ckpt[t0] = (X.clone(), F64.clone() if F64 is not None else None)
                _run_f64(sim, X, w64, t0, min(steps, t0 + K), F64=F64)
            ctx.w64 = w64
            out = X
            if sim.deform:
                sim.jac[:sim.N].copy_(F64)
        else:
            sim.pos.copy_(pos0.detach())
            for t0 in range(0, steps, K):
                ckpt[t0] = _state(sim)
                _run(sim, t0, min(steps, t0 + K))
            out = sim.pos.clone()
        ctx.sim, ctx.steps, ctx.K, ctx.ckpt, ctx.traj64 = sim, steps, K, ckpt, traj64
        ctx.in_dtype = pos0.dtype
        return out

    @staticmethod
    def backward(ctx, gX):
        sim, steps, K = ctx.sim, ctx.steps, ctx.K
        lam = gX.to(torch.float32).contiguous().clone()
        gw = torch.zeros(sim.N, dtype=torch.float32, device=lam.device)
        jbar = (torch.zeros((sim.N, 4), dtype=torch.float32, device=lam.device)
                if sim.deform else None)   # the final F does not enter the loss
        for t0 in reversed(range(0, steps, K)):
            t1 = min(steps, t0 + K)
            states = []
            if ctx.traj64:   # recompute in FP64, linearise at the rounded states
                X, F64 = (v.clone() if v is not None else None for v in ctx.ckpt[t0])
                _run_f64(sim, X, ctx.w64, t0, t1, keep=states, F64=F64)
                states = [(x.to(torch.float32), f.to(torch.float32) if f is not None else None)
                          for x, f in states]
            else:
                _set_state(sim, ctx.ckpt[t0])
                _run(sim, t0, t1, keep=states)
            for t in reversed(range(t0, t1)):
                _step_adjoint(sim, states[t - t0], t, lam, gw, jbar)
        sim.step_index = steps
        return lam.to(ctx.in_dtype), gw, None, None, None, None


def rollout(sim, pos0, w, steps, checkpoint=10, trajectory="f64"):
    """positions after `steps` plain steps (Philox steps 0 .. steps-1) from
    pos0 with strengths w; differentiable in pos0 and w.

    trajectory="f64" (default): the positions advance by the FP64 forward of
    the same frozen-sample step (vpmc_step_f64) and are returned in FP64; the
    step adjoints linearise at those states rounded to FP32, and the gradient
    matches the exact derivative to about 1e-5 of its largest entry.
    trajectory="f32": the FP32 step (vpmc_step), about 3x cheaper; the
    positions carry FP32 rounding accumulated over the steps, which bounds
    the match at about 1e-4 on a loss that compares nearly equal fields."""
    if trajectory not in ("f32", "f64"):
        raise ValueError("trajectory must be 'f32' or 'f64'")
    return _Rollout.apply(pos0, w, sim, int(steps), int(checkpoint), trajectory == "f64")


class _GridVorticity(torch.autograd.Function):
    @staticmethod
    def forward(ctx, pos, w, h_w, gnx, gny, L, cells_per_hw, f64):
        in_dtype = pos.dtype
        if f64:   # FP64 sums at the positions as given (FP64 ones stay FP64)
            out = grid_vorticity_f64(pos.detach(), w.detach(), h_w, gnx, gny, L)
        pos = pos.detach().to(torch.float32).contiguous()
        w = w.detach().to(torch.float32).contiguous()
        n = pos.shape[0]
        if not f64:
            nx, ny = ops.field_grid(h_w, L, L, cells_per_hw)
            ws = ops.workspace(lib().vpmc_grid_vorticity_workspace_bytes(n, nx, ny), pos.device)
            out = torch.empty(gnx * gny, dtype=torch.float32, device=pos.device)
            check(lib().vpmc_grid_vorticity(n, ops._p(pos), ops._p(w), h_w, L, L, cells_per_hw,
                                            gnx, gny, ops._p(out), ops._p(ws), ws.numel(),
                                            ops._stream()), "vpmc_grid_vorticity")
        ctx.save_for_backward(pos, w)
        ctx.args = (h_w, gnx, gny, L)
        ctx.in_dtype = in_dtype
        return out

    @staticmethod
    def backward(ctx, g):
        pos, w = ctx.saved_tensors
        h_w, gnx, gny, L = ctx.args
        gpos = torch.zeros_like(pos)
        gw = torch.zeros_like(w)
        check(lib().vpmc_grid_vorticity_adjoint(pos.shape[0], ops._p(pos), ops._p(w), h_w, L, L,
                                                gnx, gny,
                                                ops._p(g.to(torch.float32).contiguous()),
                                                ops._p(gpos), ops._p(gw), ops._stream()),
              "vpmc_grid_vorticity_adjoint")
        return gpos.to(ctx.in_dtype), gw, None, None, None, None, None, None


def grid_vorticity(pos, w, h_w, gnx, gny, L=1.0, cells_per_hw=8, f64=False):
    """omega_h(z) = sum_j w_j W(|z - x_j|) at the centres of a gnx x gny grid
    (index a*gny + b), differentiable in pos and w.  f64: the values are
    summed in FP64 (all particles per grid point), for losses that subtract
    nearly equal fields; the derivative is the FP32 adjoint either way."""
    return _GridVorticity.apply(pos, w, float(h_w), int(gnx), int(gny), float(L),
                                int(cells_per_hw), bool(f64))


# ---- FP64 forwards of the same functions (finite-difference checks) ------------

def rollout_f64(sim, pos0, w, steps, F0=None):
    """FP64 positions after `steps` frozen-sample steps from pos0 (float64,
    (N, 2)) with strengths w (float64); the sampling tables and the MIS
    density stay those of sim's nominal strengths.  A deformed simulation
    starts from F0 ((N, 4); default: sim's current F) and advances F in FP64."""
    X = pos0.to(torch.float64).contiguous().clone()
    w = w.to(torch.float64).contiguous()
    F64 = None
    if sim.deform:
        F64 = (sim.jac[:sim.N] if F0 is None else F0).to(torch.float64).contiguous().clone()
    ws = None
    for t in range(steps):
        ws = _step_f64(sim, X, w, t, ws, F64)
    return X


def grid_vorticity_f64(pos, w, h_w, gnx, gny, L=1.0):
    pos = pos.to(torch.float64).contiguous()
    w = w.to(torch.float64).contiguous()
    out = torch.empty(gnx * gny, dtype=torch.float64, device=pos.device)
    check(lib().vpmc_grid_vorticity_f64(pos.shape[0], ops._p(pos), ops._p(w), float(h_w),
                                        float(L), float(L), int(gnx), int(gny), ops._p(out),
                                        ops._stream()), "vpmc_grid_vorticity_f64")
    return out
