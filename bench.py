"""Benchmark of the plain GPU path (one MI355X).

Workload (default): Taylor-Green vorticity 4 pi sin(2 pi x) sin(2 pi y) on the
periodic unit box, Nx = Ny = 200 lattice (40,000 particles), hwr = 16
(h_w = 0.08), N1 = N2 = 128, dt = 1e-3, consistent estimator, Philox seed 0,
strengths from CG-RBF (FP64 MINRES, tol 1e-6).  A step is one
pass of the hot path: rebuild both sign fields from the current positions,
estimate both signs, advect.  Inputs are resident in HBM before timing.

Prints one JSON line: particle-steps/s over K timed steps, the estimator's
roofline (algorithmic FLOPs per launch / its event-timed duration), the RMS
velocity error against the FP64 direct sum, and the reference's CPU path
(oracle/_ref, compiled from its sources) timed on a bounded sample of the same
workload on this host.
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FLOP_PER_EVAL = 15      # d 2, r^2 3, sqrt 1, q 1, spline 5, weight+accumulate 3
FLOP_PER_SAMPLE = 30    # MIS contribution ~20, sampling ~10
PEAK_FP32 = 157.3e12    # MI355X FP32 (vector) peak, FLOP/s
PEAK_HBM = 8.0e12


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def setup(Nx, hwr, N1, N2, device, seed=0, cells_per_hw=8, spl=0, eval_threads=0):
    import torch

    import vortexparticlemc_cuda_amd as v
    from vortexparticlemc_cuda_amd import ops
    from vortexparticlemc_cuda_amd.sim import SimParam, Simulation, lattice_positions

    P = SimParam(dt=1e-3, resX=Nx, hwr=hwr)
    P.N1, P.N2 = N1, N2
    pos = lattice_positions(Nx)
    b = (4 * np.pi * np.sin(2 * np.pi * pos[:, 0].astype(np.float64)) *
         np.sin(2 * np.pi * pos[:, 1].astype(np.float64))).astype(np.float32)
    sim = Simulation(P, device=device, cells_per_hw=cells_per_hw, mode=v.MODE_CONSISTENT,
                     seed=seed, spl=spl, eval_threads=eval_threads)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    it, rr, st = sim.init_strengths_cg(pos, b, tol=1e-6, refine=True)
    torch.cuda.synchronize()
    t_cg = time.perf_counter() - t0
    return sim, dict(cg_iters=it, cg_rel_res=rr, cg_status=st, cg_seconds=t_cg)


def _event_pair(s):
    """two timing events, recorded once so that their hipEvent_t exists"""
    import torch
    e = (torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
    e[0].record(s)
    e[1].record(s)
    return e


PLAIN_TUNE = tuple((c, s, t) for c in (8, 12, 16) for s in (1, 2, 4) for t in (64, 128, 256))


def tune_plain(sim, steps=3, deform=False):
    """(cells_per_hw, spl, eval_threads) with the shortest plain step on this
    workload; every candidate computes identical velocities (the knobs change
    only the launch shape and the grid), timed with HIP events.  deform: the
    same for the step with aux targets and kernel deformation."""
    import torch

    from vortexparticlemc_cuda_amd.sim import Simulation
    pos = sim.pos.cpu().numpy()
    om = sim.omega.cpu().numpy()
    times = {}
    for cph, spl, th in PLAIN_TUNE:
        try:
            t = Simulation(sim.P, device=sim.device, cells_per_hw=cph, mode=sim.mode,
                           seed=sim.seed, spl=spl, eval_threads=th, deform=deform)
            t.set_particles(pos, om)
            stepper = (lambda k: [t.step() for _ in range(k)]) if deform else \
                (lambda k: run_steps(t, k))
            stepper(1)
            a = torch.cuda.Event(enable_timing=True)
            b = torch.cuda.Event(enable_timing=True)
            a.record()
            stepper(steps)
            b.record()
            torch.cuda.synchronize()
        except Exception:   # stencil does not fit this workgroup / grid
            continue
        times["%d,%d,%d" % (cph, spl, th)] = a.elapsed_time(b) / steps
        del t
    best = min(times, key=times.get)
    return tuple(int(v) for v in best.split(",")), times


def run_steps(sim, K, events=None):
    """K plain steps.  If events is a dict of lists, it receives per step the
    event pairs around the whole estimator ("est", torch's stream) and around
    its evaluation kernel ("eval", recorded by the library on the launch
    stream)."""
    import torch

    from vortexparticlemc_cuda_amd import ops

    s = torch.cuda.current_stream()
    P = sim.P
    pairs = []
    if events is not None:
        pairs = [(_event_pair(s), _event_pair(s)) for _ in range(K)]
    for i in range(K):
        d = sim.desc(P.N1, P.N2)
        ws = sim._ws_for(d)   # one buffer serves the field builds and the estimator
        for k, key in enumerate(("sub_pos", "sub_neg")):
            ops.field_build(sim.fields[k], sim.pos, sim.omega, sub=sim.sub[key], abs_w=True,
                            ws=ws)
        if events is not None:
            est, ev = pairs[i]
            d.eval_events[0] = ev[0].cuda_event
            d.eval_events[1] = ev[1].cuda_event
            est[0].record(s)
        sim.launch_estimate(d)
        if events is not None:
            est[1].record(s)
            events["est"].append(est)
            events["eval"].append(ev)
        ops.advect(sim.pos, sim.vel_sign[:2 * sim.N], float(P.dt), float(P.Lx), float(P.Ly))
        sim.step_index += 1


def copy_probe(nbytes, reps=3):
    """a plain device copy of nbytes (reads nbytes, writes nbytes), for
    calibrating HBM byte counters"""
    import torch
    n = nbytes // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda:0")
    b = torch.empty_like(a)
    for _ in range(reps):
        b.copy_(a)
    torch.cuda.synchronize()
    log("copy probe: %d x %d bytes" % (reps, n * 4))


# algorithmic HBM bytes per particle of the HBM-bound stages (hash/field build,
# |omega| CDF, advection): positions, strengths and indices read, records and
# keys written once
HBM_BYTES_FIELD = 28.0
HBM_BYTES_CDF = 16.0
HBM_BYTES_ADVECT = 24.0


def hbm_stages(sim, reps=5):
    """achieved GB/s of the HBM-bound stages on sim's particles: the two
    sign fields' builds (both together, 28 B per particle), the subsets with
    FP64 CDFs (16 B per particle) and advection (24 B per particle); HIP
    events around reps launches of each on the current stream"""
    import torch

    from vortexparticlemc_cuda_amd import ops
    from vortexparticlemc_cuda_amd._capi import lib
    N = sim.N
    ws = ops.workspace(lib().vpmc_field_workspace_bytes(max(sim.cnt), sim.nx, sim.ny), sim.device)
    vel_sign = torch.zeros((2 * N, 2), dtype=torch.float32, device=sim.device)
    pos = sim.pos.clone()

    def fields():
        for k, key in enumerate(("sub_pos", "sub_neg")):
            ops.field_build(sim.fields[k], sim.pos, sim.omega, sub=sim.sub[key], abs_w=True,
                            ws=ws, jac=sim.jac if sim.deform else None)

    stages = (("field_build", fields, HBM_BYTES_FIELD),
              ("subsets_cdf", lambda: ops.build_subsets(sim.omega), HBM_BYTES_CDF),
              ("advect", lambda: ops.advect(pos, vel_sign, 0.0), HBM_BYTES_ADVECT))
    out = {}
    for name, fn, bpp in stages:
        fn()
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record()
        for _ in range(reps):
            fn()
        b.record()
        torch.cuda.synchronize()
        t = a.elapsed_time(b) / 1e3 / reps
        gbs = bpp * N / t
        out[name] = {"ms": t * 1e3, "bytes_per_particle": bpp, "achieved_GBps": gbs / 1e9,
                     "frac_of_hbm": gbs / PEAK_HBM}
    return out


def fma_probe(blocks=2048, iters=20000, reps=3):
    """measured unpacked FP32 FMA rate (FLOP/s) of this device: the best of
    reps launches of vpmc_probe_fma, timed with HIP events on the stream"""
    import torch

    from vortexparticlemc_cuda_amd import ops
    from vortexparticlemc_cuda_amd._capi import check, lib
    out = torch.zeros(blocks, dtype=torch.float32, device="cuda:0")
    best = None
    for _ in range(reps + 1):
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record()
        check(lib().vpmc_probe_fma(blocks, iters, ops._p(out), ops._stream()), "vpmc_probe_fma")
        b.record()
        torch.cuda.synchronize()
        t = a.elapsed_time(b) / 1e3
        best = t if best is None else min(best, t)   # the first launch warms up
    return blocks * 256.0 * iters * 32.0 / best


def count_evals(sim):
    import torch

    from vortexparticlemc_cuda_amd import ops
    from vortexparticlemc_cuda_amd._capi import check, lib

    sim.eval_count.zero_()
    sim.rebuild_fields()
    d = sim.desc(sim.P.N1, sim.P.N2, count=True)
    sim.launch_estimate(d)
    torch.cuda.synchronize()
    c = sim.eval_count.cpu().numpy()
    return int(c[0]), int(c[1])


def velocity_error(sim):
    """RMS |u_MC - u_oracle| / RMS |u_oracle| at the current positions; oracle =
    FP64 direct sum of the blob field with the periodic kernel (device)"""
    import torch

    from vortexparticlemc_cuda_amd import ops

    u = sim.estimate(sim.P.N1, sim.P.N2).double()
    pos64 = sim.pos.double()
    ud = ops.direct_sum_blob(pos64, pos64, sim.omega.double(), float(sim.P.h_w))
    x = pos64[:, 0]
    y = pos64[:, 1]
    ut = torch.stack([torch.sin(2 * np.pi * x) * torch.cos(2 * np.pi * y),
                      -torch.cos(2 * np.pi * x) * torch.sin(2 * np.pi * y)], 1)
    rms = lambda a: float(torch.sqrt((a * a).sum(1).mean()))  # noqa: E731
    return dict(rms_err_vs_direct=rms(u - ud) / rms(ud), direct_vs_tg=rms(ud - ut) / rms(ut),
                rms_err_vs_tg=rms(u - ut) / rms(ut))


def deform_bench(args, device, base):
    """the reference's CUDA step semantics (step_cuda): five estimates per
    owner and sign (owner + the aux targets at dst -/+ dx, branch-1 samples
    shared), F = update_jacobian_from5 of them, deformed kernels W(|F^-1 d|)
    in every query and F u in the branch-1 jitter; same workload and launch
    shape as the plain step"""
    import ctypes as C

    import torch

    from vortexparticlemc_cuda_amd import ops
    from vortexparticlemc_cuda_amd._capi import check, lib
    from vortexparticlemc_cuda_amd.sim import Simulation

    P = base.P
    cph, spl, th = args.cells_per_hw, args.spl, args.eval_threads
    tuned = None
    if not args.no_tune:
        (cph, spl, th), tune_ms = tune_plain(base, deform=True)
        tuned = {"cells_per_hw": cph, "spl": spl, "eval_threads": th, "candidates_ms": tune_ms}
        log("tuned deformed step: cells_per_hw=%d spl=%d eval_threads=%d (%.3f ms)"
            % (cph, spl, th, tune_ms["%d,%d,%d" % (cph, spl, th)]))
    sim = Simulation(P, device=device, cells_per_hw=cph, mode=base.mode, seed=base.seed,
                     spl=spl, eval_threads=th, deform=True)
    sim.set_particles(base.pos.cpu().numpy(), base.omega.cpu().numpy())
    for _ in range(args.warmup):
        sim.step(P.N1, P.N2)
    s = torch.cuda.current_stream()
    K = args.deform_steps
    pairs = [_event_pair(s) for _ in range(K)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(K):
        d = sim.desc(P.N1, P.N2, aux=True)
        d.eval_events[0] = pairs[i][0].cuda_event
        d.eval_events[1] = pairs[i][1].cuda_event
        ws = sim._ws_for(d)
        check(lib().vpmc_step(C.byref(d), sim.N, ops._p(sim.pos), ops._p(sim.omega),
                              ops._p(sim.sub["sub_pos"]), ops._p(sim.sub["sub_neg"]),
                              ops._p(sim.vel), float(P.dt), ops._p(ws), ws.numel(),
                              ops._stream()), "vpmc_step")
        sim.step_index += 1
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    eval_s = float(np.mean([a.elapsed_time(b) for a, b in pairs])) / 1e3
    sim.eval_count.zero_()
    sim.rebuild_fields()
    d = sim.desc(P.N1, P.N2, count=True, aux=True)
    sim.launch_estimate(d)
    torch.cuda.synchronize()
    n_in, n_vis = (int(v) for v in sim.eval_count.cpu().numpy())
    F = sim.jac[:sim.N].double()
    det = F[:, 0] * F[:, 3] - F[:, 1] * F[:, 2]
    sv = torch.linalg.svdvals(F.view(-1, 2, 2))
    # reinitialisation of the deformed state onto the lattice (FP64 field at
    # the lattice, MINRES refit to 1e-6, F = I), timed as a whole
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    pos_r, w_r, it_r, rr_r, st_r = ops.reinit(sim.pos, sim.omega, float(P.h_w), P.ResolutionX,
                                              P.ResolutionY, jac=sim.jac[:sim.N],
                                              Lx=float(P.Lx), Ly=float(P.Ly), tol=1e-6,
                                              cells_per_hw=sim.cells_per_hw)
    torch.cuda.synchronize()
    t_r = time.perf_counter() - t0
    out = {"workload": "the plain workload with the CUDA step's aux targets and kernel "
                       "deformation (F reset from the five estimates every step)",
           "particle_steps_per_s": sim.N * K / el, "ms_per_step": el / K * 1e3, "steps": K,
           "samples_per_step": sim.N * 2 * (P.N1 + 5 * P.N2),
           "eval_kernel": "k_bin_eval (deformed records)", "eval_avg_ms": eval_s * 1e3,
           "kernel_evals_per_launch": n_in, "candidates_per_launch": n_vis,
           "achieved_flops": FLOP_PER_EVAL * n_in / eval_s,
           "frac_of_fp32_peak": FLOP_PER_EVAL * n_in / eval_s / PEAK_FP32,
           "tuned": tuned,
           "reinit": {"ms": t_r * 1e3, "iterations": it_r, "rel_residual": rr_r,
                      "status": st_r, "lattice": [P.ResolutionX, P.ResolutionY]},
           "max_abs_det_minus_1": float((det - 1).abs().max()),
           "max_sigma_max_F": float(sv[:, 0].max())}
    log("deform", out)
    return out


def ensemble_bench(args, device):
    """The batched ensemble (64 periodic simulations, Nx drawn from
    {64, 96, 128, 192, 256}, 8 Gaussian vortices each, hwr = 8, N1 = N2 = 64,
    dt = 5e-4): particle-steps/s of the batched path (one launch per stage,
    eager and replayed from a HIP graph) and of the plain per-simulation path
    on the same ensemble.  Both compute the same estimator (bit-identical per
    simulation), so their RMS velocity errors are equal by construction."""
    import torch

    from vortexparticlemc_cuda_amd import ops
    from vortexparticlemc_cuda_amd._capi import ALGO_DIRECT
    from vortexparticlemc_cuda_amd.ensemble import Ensemble, autotune
    from vortexparticlemc_cuda_amd.sim import SimParam, Simulation
    from vortexparticlemc_cuda_amd.workloads import gaussian_ensemble

    data = gaussian_ensemble(args.ens_sims)
    hwr, N1, N2, dt = 8, args.ens_N, args.ens_N, 5e-4
    Ps, pos, w = [], [], []
    t0 = time.perf_counter()
    for nx, p, om, _ in data:
        P = SimParam(dt=dt, resX=nx, hwr=hwr)
        P.N1, P.N2 = N1, N2
        h = 1.0 / nx
        b = (om * 1.0).astype(np.float32)
        # the reference's tolerance (src/cuda/cg_rbf_init.cu:53): tighter fits
        # of these fields amplify the Gram matrix's near-null modes into
        # alternating strengths, which inflate the estimator's variance
        wt, it, rr, st = ops.cg_rbf(torch.as_tensor(p, device=device), torch.as_tensor(b, device=device),
                                    float(P.h_w), tol=5e-3, max_iter=300, refine=True)
        Ps.append(P)
        pos.append(p)
        w.append(wt.cpu().numpy())
    t_cg = time.perf_counter() - t0
    hw = [float(P.h_w) for P in Ps]
    N = sum(p.shape[0] for p in pos)
    log("ensemble: %d sims, %d particles, CG init %.1f s" % (len(pos), N, t_cg))
    cph, spl, th, tune_ms = autotune(pos, w, hw, N1=N1, N2=N2, device=device)
    log("autotune: cells_per_hw=%d spl=%d eval_threads=%d" % (cph, spl, th))
    ens = Ensemble(pos, w, hw, N1=N1, N2=N2, dt=dt, seed=0, cells_per_hw=cph, spl=spl,
                   eval_threads=th, device=device)
    ens.step(1)
    v0 = ens.vel.clone()
    p0 = [torch.as_tensor(p, device=device) for p in pos]
    # RMS error vs the FP64 blob direct sum on the four smallest simulations
    small = sorted(range(len(pos)), key=lambda s: pos[s].shape[0])[:4]
    num = den = 0.0
    for s in small:
        pd = p0[s].double()
        ud = ops.direct_sum_blob(pd, pd, torch.as_tensor(w[s], device=device).double(), hw[s])
        um = v0[int(ens.offsets[s]):int(ens.offsets[s + 1])].double()
        num += float(((um - ud) ** 2).sum())
        den += float((ud ** 2).sum())
    err = (num / den) ** 0.5
    K = args.ens_steps
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ens.step(K)
    torch.cuda.synchronize()
    t_eager = time.perf_counter() - t0
    # the evaluation kernel alone: HIP events the library records around it
    s0 = torch.cuda.current_stream()
    ev_pairs = [_event_pair(s0) for _ in range(K)]
    for pr in ev_pairs:
        ens.step(1, events=pr)
    torch.cuda.synchronize()
    ens_eval_s = float(np.mean([a.elapsed_time(b) for a, b in ev_pairs])) / 1e3
    side = torch.cuda.Stream()
    with torch.cuda.stream(side):
        ens.capture(K)
        ens.replay(K)   # first replay uploads the graph
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ens.replay(K)
        torch.cuda.synchronize()
        t_graph = time.perf_counter() - t0
    ens.close()
    batched = N * K / min(t_eager, t_graph)
    del ens
    torch.cuda.empty_cache()
    # the same ensemble with kernel deformation (the CUDA step: four aux
    # targets per owner, F in the fields and the jitter, F from the five
    # estimates), same launch shape
    ensd = Ensemble(pos, w, hw, N1=N1, N2=N2, dt=dt, seed=0, cells_per_hw=cph, spl=spl,
                    eval_threads=th, device=device, deform=True)
    ensd.step(1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ensd.step(K)
    torch.cuda.synchronize()
    t_def = time.perf_counter() - t0
    ensd.close()
    del ensd
    torch.cuda.empty_cache()
    log("batched deformed: %.1f ms/step" % (t_def / K * 1e3))
    # plain path: one Simulation per member, stepped one after the other
    sims = []
    for s in range(len(pos)):
        sim = Simulation(Ps[s], device=device, cells_per_hw=cph, seed=0, sim=s, spl=spl,
                         eval_threads=th)
        sim.set_particles(pos[s], w[s])
        sims.append(sim)
    # evaluations within h_w and candidates visited by the ensemble's step:
    # the counting pass of each member at its first step (same Philox streams
    # and launch shape as the ensemble)
    n_in = n_vis = 0
    for sim in sims:
        sim.eval_count.zero_()
        sim.estimate(N1, N2, count=True)
        c = sim.eval_count.cpu().numpy()
        n_in += int(c[0])
        n_vis += int(c[1])
    for sim in sims:
        sim.step()
    Kp = max(1, args.ens_plain_steps)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(Kp):
        for sim in sims:
            sim.step()
    torch.cuda.synchronize()
    t_plain = time.perf_counter() - t0
    plain = N * Kp / t_plain
    # the same per-simulation loop with the reference's launch structure: one
    # workgroup per (owner, sign) sampling and querying on its own
    # (monte_carlo_bs_kernel_safe, src/cuda/monte_carlo_bs.cu:144-271)
    for sim in sims:
        sim.algo = ALGO_DIRECT
    for sim in sims:
        sim.step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for sim in sims:
        sim.step()
    torch.cuda.synchronize()
    t_direct = time.perf_counter() - t0
    plain_direct = N / t_direct
    del sims
    torch.cuda.empty_cache()
    log("batched %.3g particle-steps/s (eager %.1f ms/step, graph %.1f ms/step), plain %.3g"
        % (batched, t_eager / K * 1e3, t_graph / K * 1e3, plain))
    return {"workload": "64 periodic simulations, Nx from {64,96,128,192,256} (seed 20261203), "
                        "8 Gaussian vortices each, hwr=8, N1=N2=%d, dt=5e-4, consistent "
                        "estimator, CG-RBF strengths" % N1,
            "simulations": len(pos), "particles": N,
            "batched_particle_steps_per_s": batched,
            "batched_ms_per_step_eager": t_eager / K * 1e3,
            "batched_ms_per_step_graph": t_graph / K * 1e3,
            "plain_particle_steps_per_s": plain, "plain_ms_per_step": t_plain / Kp * 1e3,
            "batched_over_plain": batched / plain, "steps_timed": K, "plain_steps_timed": Kp,
            "plain_direct_particle_steps_per_s": plain_direct,
            "plain_direct_ms_per_step": t_direct * 1e3,
            "batched_over_plain_direct": batched / plain_direct,
            "rms_err_vs_direct": err, "cells_per_hw": cph, "spl": spl, "eval_threads": th,
            "autotune_ms": tune_ms,
            "deformed": {"ms_per_step": t_def / K * 1e3, "particle_steps_per_s": N * K / t_def,
                         "samples_per_pair": N1 + 5 * N2},
            "roofline": {"kernel": "k_ens_eval", "bound": "valu",
                         "achieved": FLOP_PER_EVAL * n_in / ens_eval_s, "peak": PEAK_FP32,
                         "unit": "FLOP/s", "frac": FLOP_PER_EVAL * n_in / ens_eval_s / PEAK_FP32,
                         "avg_launch_ms": ens_eval_s * 1e3, "kernel_evals_per_launch": n_in,
                         "candidates_per_launch": n_vis,
                         "frac_of_unpacked": FLOP_PER_EVAL * n_in / ens_eval_s / (PEAK_FP32 / 2)},
            "cg_init_s": t_cg}


def large_bench(args, device):
    """The large run: double shear layer (rho = 1/30, delta = 0.05) on the
    periodic unit box, Nx = Ny = 1024 (1,048,576 particles), hwr = 8,
    N1 = N2 = 256, dt = 2.5e-4; strengths from the RBF solve to 1e-10 (FP64
    MINRES on the FP64 vorticity, preconditioned by the lattice's circulant
    Gram matrix); plain steps with a tuned launch shape; RMS error
    against the FP64 blob direct sum on 4,096 uniform probe points (seed
    20261203), the estimator evaluated at those probes on the same fields."""
    import torch

    from vortexparticlemc_cuda_amd import ops
    from vortexparticlemc_cuda_amd.sim import SimParam, Simulation, lattice_positions
    from vortexparticlemc_cuda_amd.workloads import double_shear

    Nx, hwr, N = args.large_Nx, 8, args.large_N
    P = SimParam(dt=2.5e-4, resX=Nx, hwr=hwr)
    P.N1 = P.N2 = N
    pos = lattice_positions(Nx)
    b = double_shear(pos[:, 0], pos[:, 1])   # FP64 right-hand side
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    w, it, rr, st = ops.cg_rbf(torch.as_tensor(pos, device=device),
                               torch.as_tensor(b, device=device), float(P.h_w), tol=1e-10,
                               max_iter=3000, refine=True, cells_per_hw=8, lattice=(Nx, Nx))
    torch.cuda.synchronize()
    t_cg = time.perf_counter() - t0
    log("large: %d particles, RBF solve %d iterations, rel. residual %.3g (status %d), %.2f s"
        % (pos.shape[0], it, rr, st, t_cg))
    sim = Simulation(P, device=device, cells_per_hw=args.large_cells_per_hw, seed=0,
                     spl=args.large_spl, eval_threads=args.large_threads)
    sim.set_particles(pos, w.cpu().numpy())
    del w
    # error at probe points: the same fields, owners = probes
    rng = np.random.default_rng(20261203)
    probes = torch.as_tensor(rng.random((args.large_probes, 2)).astype(np.float32), device=device)
    sim.rebuild_fields()
    d = sim.desc(N, N)
    d.n_owner = probes.shape[0]
    d.owner_xy = ops._p(probes)
    vs = torch.zeros((2 * probes.shape[0], 2), dtype=torch.float32, device=device)
    d.vel_sign = ops._p(vs)
    sim.launch_estimate(d)
    um = (vs[:probes.shape[0]] - vs[probes.shape[0]:]).double()
    ud = ops.direct_sum_blob(probes.double(), sim.pos.double(), sim.omega.double(),
                             float(P.h_w))
    err = float(torch.sqrt(((um - ud) ** 2).sum() / (ud ** 2).sum()))
    log("large: RMS error vs FP64 blob direct sum at %d probes: %.4g" % (probes.shape[0], err))
    del vs, um, ud
    hbm = hbm_stages(sim)
    log("large: HBM-bound stages", hbm)
    # launch shape: a few candidates, one timed step each, on the one
    # simulation (its buffers are re-laid for each grid)
    tune = {}
    if args.large_tune:
        for cph, spl, th in ((6, 1, 256), (8, 1, 256), (8, 2, 256), (12, 1, 256)):
            sim.cells_per_hw, sim.spl, sim.eval_threads = cph, spl, th
            sim._alloc()
            try:
                run_steps(sim, 1)
                a = torch.cuda.Event(enable_timing=True)
                b = torch.cuda.Event(enable_timing=True)
                a.record()
                run_steps(sim, 1)
                b.record()
                torch.cuda.synchronize()
            except Exception:   # stencil does not fit this workgroup / grid
                continue
            tune["%d,%d,%d" % (cph, spl, th)] = a.elapsed_time(b)
        best = min(tune, key=tune.get)
        sim.cells_per_hw, sim.spl, sim.eval_threads = (int(v) for v in best.split(","))
        sim._alloc()
        log("large: tuned cells_per_hw=%d spl=%d eval_threads=%d (%s)"
            % (sim.cells_per_hw, sim.spl, sim.eval_threads, tune))
    run_steps(sim, 2)
    torch.cuda.synchronize()
    events = {"est": [], "eval": []}
    K = args.large_steps
    t0 = time.perf_counter()
    run_steps(sim, K, events)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    est_s = float(np.mean([a.elapsed_time(b) for a, b in events["est"]])) / 1e3
    eval_s = float(np.mean([a.elapsed_time(b) for a, b in events["eval"]])) / 1e3
    n_in, n_vis = count_evals(sim)
    flop = FLOP_PER_EVAL * n_in
    n = sim.N
    log("large: %.1f ms/step, estimator %.1f ms, evaluation %.1f ms, %.3g evals/launch"
        % (el / K * 1e3, est_s * 1e3, eval_s * 1e3, n_in))
    out = {"workload": "double shear layer (rho=1/30, delta=0.05) on periodic [0,1]^2, "
                       "Nx=Ny=%d (%d particles), hwr=8, N1=N2=%d, dt=2.5e-4, consistent "
                       "estimator, Philox seed 0, RBF strengths to 1e-10" % (Nx, n, N),
           "particles": n, "steps_timed": K, "ms_per_step": el / K * 1e3,
           "particle_steps_per_s": n * K / el, "estimator_ms": est_s * 1e3,
           "rms_err_vs_direct": err, "probes": int(probes.shape[0]),
           "rbf_iters": it, "rbf_rel_res": rr, "rbf_status": st, "rbf_seconds": t_cg,
           "hbm_stages": hbm,
           "cells_per_hw": sim.cells_per_hw, "spl": sim.spl, "eval_threads": sim.eval_threads,
           "tune_ms": tune,
           "roofline": {"kernel": "k_bin_eval", "bound": "valu", "achieved": flop / eval_s,
                        "peak": PEAK_FP32, "unit": "FLOP/s", "frac": flop / eval_s / PEAK_FP32,
                        "avg_launch_ms": eval_s * 1e3, "kernel_evals_per_launch": n_in,
                        "candidates_per_launch": n_vis,
                        "samples_per_launch": n * 2 * 2 * N,
                        "frac_of_unpacked": flop / eval_s / (PEAK_FP32 / 2)}}
    del sim
    torch.cuda.empty_cache()
    return out


def adjoint_bench(args, device):
    """The adjoint workload: Taylor-Green on the periodic unit box, Nx = Ny =
    128 (16,384 particles), hwr = 16, N1 = N2 = 64, dt = 1e-3, T = 50 steps,
    checkpoint every 10.  Strengths: the RBF solution w_TG of the Taylor-Green
    vorticity plus 0.1 N(0,1) per particle in vorticity units (scaled by
    rms(w_TG) / rms(omega_TG), the lattice's vorticity-to-strength factor).
    Loss: mean squared difference on a 64 x 64 grid between the final
    vorticity (B-spline sum) and that of the unperturbed w_TG evolved with
    the same Philox stream.  Times the forward rollout and the checkpointed
    backward (recompute + step adjoints), and checks dL/dx0 and dL/dw against
    central differences of the FP64 forward of the same frozen-sample rollout
    on `adj_fd` coordinates (half positions, half strengths; eps = 1e-4 for
    positions, 1e-4 max|w| for strengths)."""
    import torch

    from vortexparticlemc_cuda_amd import adjoint as A
    from vortexparticlemc_cuda_amd import ops
    from vortexparticlemc_cuda_amd.sim import SimParam, Simulation, lattice_positions
    from vortexparticlemc_cuda_amd.workloads import taylor_green

    Nx, hwr, N, T, K, G = args.adj_Nx, 16, 64, args.adj_steps, 10, 64
    P = SimParam(dt=1e-3, resX=Nx, hwr=hwr)
    P.N1 = P.N2 = N
    pos = lattice_positions(Nx)
    om_tg = taylor_green(pos[:, 0], pos[:, 1])
    pos_t = torch.as_tensor(pos, device=device)
    w_tg, it, rr, st = ops.cg_rbf(pos_t, torch.as_tensor(om_tg.astype(np.float32), device=device),
                                  float(P.h_w), tol=1e-6, refine=True)
    w_tg = w_tg.clone()
    scale = float(torch.sqrt((w_tg.double() ** 2).mean())) / float(np.sqrt((om_tg ** 2).mean()))
    rng = np.random.default_rng(20261203)
    w0 = (w_tg.double() + torch.as_tensor(0.1 * scale * rng.standard_normal(pos.shape[0]),
                                          device=device)).float()
    h = float(P.h_w)
    sim = Simulation(P, device=device, cells_per_hw=8, seed=0)
    sim.set_particles(pos, w_tg.cpu().numpy())
    with torch.no_grad():
        X_ref = A.rollout(sim, pos_t, w_tg, T, checkpoint=K)
        ref = A.grid_vorticity(X_ref, w_tg, h, G, G, f64=True)
    p = pos_t.clone().requires_grad_(True)
    wt = w0.clone().requires_grad_(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    XT = A.rollout(sim, p, wt, T, checkpoint=K)
    om = A.grid_vorticity(XT, wt, h, G, G, f64=True)
    loss = ((om - ref) ** 2).mean()
    torch.cuda.synchronize()
    t_fwd = time.perf_counter() - t0
    t0 = time.perf_counter()
    loss.backward()
    torch.cuda.synchronize()
    t_bwd = time.perf_counter() - t0
    gp = p.grad.double().cpu().numpy()
    gw = wt.grad.double().cpu().numpy()
    n = pos.shape[0]
    log("adjoint: loss %.6g, forward %.3f s, backward %.3f s (%d steps, %d particles)"
        % (float(loss.detach()), t_fwd, t_bwd, T, n))
    # the same gradient along the FP64 trajectory of the frozen-sample step
    p2 = pos_t.clone().requires_grad_(True)
    wt2 = w0.clone().requires_grad_(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    XT2 = A.rollout(sim, p2, wt2, T, checkpoint=K, trajectory="f64")
    loss2 = ((A.grid_vorticity(XT2, wt2, h, G, G, f64=True) - ref) ** 2).mean()
    torch.cuda.synchronize()
    t_fwd64 = time.perf_counter() - t0
    t0 = time.perf_counter()
    loss2.backward()
    torch.cuda.synchronize()
    t_bwd64 = time.perf_counter() - t0
    grads = {"": (p2.grad.double().cpu().numpy(), wt2.grad.double().cpu().numpy()),
             "_f32traj": (gp, gw)}
    log("adjoint (FP64 trajectory): loss %.9g, forward %.3f s, backward %.3f s"
        % (float(loss2.detach()), t_fwd64, t_bwd64))
    out = {"workload": "Taylor-Green + 0.1 N(0,1) strength noise, Nx=Ny=%d (%d particles), "
                       "hwr=16, N1=N2=64, dt=1e-3, T=%d, checkpoint every %d, loss = MSE of "
                       "the final B-spline vorticity on a 64x64 grid vs the unperturbed run"
                       % (Nx, n, T, K),
           "loss": float(loss.detach()), "forward_s": t_fwd, "backward_s": t_bwd,
           "forward_particle_steps_per_s": n * T / t_fwd,
           "adjoint_particle_steps_per_s": n * T / t_bwd,
           "grad_pos_max": float(np.abs(gp).max()), "grad_w_max": float(np.abs(gw).max()),
           "f64_trajectory": {"loss": float(loss2.detach()), "forward_s": t_fwd64,
                              "backward_s": t_bwd64}}
    if args.adj_fd > 0:
        p64 = pos_t.double()
        w64 = w0.double()

        def L(pp, ww):
            X = A.rollout_f64(sim, pp, ww, T)
            o = A.grid_vorticity_f64(X, ww, h, G, G)
            return float(((o - ref) ** 2).mean())

        t0 = time.perf_counter()
        L0 = L(p64, w64)
        wn = w0.abs().cpu().numpy()
        cand = np.nonzero(wn > 0.2 * wn.max())[0]
        pick = rng.choice(cand, size=args.adj_fd, replace=False)
        axis = rng.integers(2, size=args.adj_fd)
        out.update({"fd_coordinates": int(args.adj_fd), "fd_loss_fp64": L0})
        # eps = 1e-4 as specified, and 1e-6: the frozen-sample velocity has the
        # 1/r Biot-Savart kernel at branch-1 samples close to their owner, so
        # the wider difference carries truncation error of its own
        # headline: the gradient along the FP64 trajectory ("" keys); the
        # FP32-trajectory gradient ("_f32traj") beside it
        for tag, eps_x, eps_w in (("", 1e-4, 1e-4 * float(w0.abs().max())),
                                  ("_eps1e-6", 1e-6, 1e-6 * float(w0.abs().max()))):
            fds = []
            for c, i in enumerate(pick):
                if c % 2 == 0:
                    k = int(axis[c])
                    a, b = p64.clone(), p64.clone()
                    a[i, k] += eps_x
                    b[i, k] -= eps_x
                    fds.append((L(a, w64) - L(b, w64)) / (2 * eps_x))
                else:
                    a, b = w64.clone(), w64.clone()
                    a[i] += eps_w
                    b[i] -= eps_w
                    fds.append((L(p64, a) - L(p64, b)) / (2 * eps_w))
            for gtag, (gpx, gwx) in grads.items():
                ex, ew, rel = [], [], []
                for c, i in enumerate(pick):
                    fd = fds[c]
                    if c % 2 == 0:
                        g = gpx[i, int(axis[c])]
                        ex.append(abs(fd - g) / np.abs(gpx).max())
                    else:
                        g = gwx[i]
                        ew.append(abs(fd - g) / np.abs(gwx).max())
                    rel.append(abs(fd - g) / max(abs(fd), 1e-300))
                log("adjoint%s vs FP64 central differences (eps %.0e): max err/max|g| "
                    "positions %.3g, strengths %.3g; median per-coordinate rel err %.3g"
                    % (gtag, eps_x, max(ex), max(ew), np.median(rel)))
                k2 = tag + gtag
                out.update({"fd_max_err_pos" + k2: max(ex), "fd_max_err_w" + k2: max(ew),
                            "fd_median_rel_err" + k2: float(np.median(rel)),
                            "fd_max_rel_err" + k2: float(np.max(rel))})
        out["fd_seconds"] = time.perf_counter() - t0
    del sim
    torch.cuda.empty_cache()
    return out


def read_traffic(kernel, args):
    """HBM bytes per launch of `kernel` at this workload, from the committed
    counter summary (FETCH_SIZE + WRITE_SIZE, copy-calibrated); None if absent"""
    path = os.path.join(ROOT, "profiles", "traffic_%s_tg%d.txt" % (kernel, args.Nx))
    if not os.path.exists(path):
        return None
    want = "N1=%d N2=%d hwr=%d cells_per_hw=%d spl=%d eval_threads=%d" % (
        args.N1, args.N2, args.hwr, args.cells_per_hw, args.spl, args.eval_threads)
    ok = False
    for line in open(path):
        if line.startswith("workload:"):
            ok = want in line
        if ok and line.startswith("hbm_bytes_per_launch:"):
            return float(line.split(":", 1)[1])
    return None


def cpu_baseline(Nx, hwr, N1, N2, target_s=10.0):
    """the reference's monte_carlo_bs_cpu (its compiled src/cpu sources,
    oracle/_ref) over a bounded subset of owners of the same workload, aux = 0
    (the work the GPU path does): single-threaded like its step_cpu, and over
    the host's cores (OpenMP over owners, one generator per thread)"""
    import os

    import oracle as O
    R = O.ref_lib()
    if R is None:
        return None
    x, y = O.lattice(Nx)
    x = x.astype(np.float32)
    y = y.astype(np.float32)
    w = (O.tg_vorticity(x, y) / O.tg_lambda11(Nx, hwr)).astype(np.float32)
    rsh = O.rsh_of(Nx, hwr)
    hw = np.float32(hwr / Nx)
    n = x.size

    def timed(m, threads):
        vel = np.zeros(2 * m, np.float32)
        t0 = time.perf_counter()
        if threads == 1:
            R.ref_mc_velocities(n, x, y, w, rsh, rsh, 1.0, 1.0, hw, 1, N1, N2, 0, 5489, 0, m,
                                vel)
        else:
            R.ref_mc_velocities_mt(n, x, y, w, rsh, rsh, 1.0, 1.0, hw, 1, N1, N2, 5489, 0, m,
                                   threads, vel)
        return time.perf_counter() - t0

    def run(threads):
        probe = 8 * threads
        t1 = timed(probe, threads)
        m = int(max(probe, min(n, target_s / max(t1 / probe, 1e-9))))
        t = timed(m, threads)
        return m, t

    threads = max(1, min(16, os.cpu_count() or 1))
    m1, t1 = run(1)
    mt, tt = run(threads)
    cpu = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            cpu = next(l.split(":", 1)[1].strip() for l in f if l.startswith("model name"))
    except (OSError, StopIteration):
        pass
    return {"value": mt / tt, "unit": "particle-steps/s", "cores": threads, "kind": "reference",
            "sample": "reference monte_carlo_bs_cpu (src/cpu, compiled from source) for %d of "
                      "%d owners, both signs, N1=N2=%d, aux=0, OpenMP over owners on %d threads "
                      "(%s), %.1f s, scaled per particle" % (mt, n, N1, threads, cpu, tt),
            "single_thread": {"value": m1 / t1, "cores": 1,
                              "sample": "%d owners, %.1f s" % (m1, t1)}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--Nx", type=int, default=200)
    ap.add_argument("--hwr", type=int, default=16)
    ap.add_argument("--N1", type=int, default=128)
    ap.add_argument("--N2", type=int, default=128)
    ap.add_argument("--cells-per-hw", type=int, default=8,
                    help="fine-grid cells per h_w (with --no-tune; else tuned)")
    ap.add_argument("--spl", type=int, default=0, help="samples per lane (with --no-tune)")
    ap.add_argument("--eval-threads", type=int, default=0,
                    help="evaluation workgroup size (with --no-tune)")
    ap.add_argument("--algo", choices=["binned", "direct"], default="binned")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-error", action="store_true")
    ap.add_argument("--no-tune", action="store_true",
                    help="skip the launch-shape tuning of the plain step")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="also run a plain copy of MIB MiB (HBM counter calibration)")
    ap.add_argument("--deform-steps", type=int, default=10,
                    help="steps of the deformed (aux + F) step on the plain workload (0: skip)")
    ap.add_argument("--ens-sims", type=int, default=64,
                    help="simulations in the batched-ensemble measurement (0: skip it)")
    ap.add_argument("--ens-N", type=int, default=64, help="N1 = N2 of the ensemble")
    ap.add_argument("--ens-steps", type=int, default=10)
    ap.add_argument("--ens-plain-steps", type=int, default=2)
    ap.add_argument("--large-Nx", type=int, default=1024,
                    help="lattice side of the large double-shear run (0: skip it)")
    ap.add_argument("--large-N", type=int, default=256, help="N1 = N2 of the large run")
    ap.add_argument("--large-steps", type=int, default=10)
    ap.add_argument("--large-probes", type=int, default=4096)
    ap.add_argument("--large-cells-per-hw", type=int, default=12)
    ap.add_argument("--large-spl", type=int, default=1)
    ap.add_argument("--large-threads", type=int, default=256)
    ap.add_argument("--large-tune", type=int, default=1,
                    help="try a few launch shapes on the large run (0: use the flags)")
    ap.add_argument("--adj-Nx", type=int, default=128,
                    help="lattice side of the adjoint workload (0: skip it)")
    ap.add_argument("--adj-steps", type=int, default=50)
    ap.add_argument("--adj-fd", type=int, default=32,
                    help="coordinates checked by FP64 central differences")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU; this build has no multi-GPU path "
              "(--gpus %d requested)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import torch
    device = torch.device("cuda:0")
    log("setup Nx=%d hwr=%d N1=%d N2=%d" % (args.Nx, args.hwr, args.N1, args.N2))
    sim, init = setup(args.Nx, args.hwr, args.N1, args.N2, device,
                      cells_per_hw=args.cells_per_hw, spl=args.spl,
                      eval_threads=args.eval_threads)
    sim.algo = 0 if args.algo == "binned" else 1
    log("cg init", init)
    tuned = None
    if args.algo == "binned" and not args.no_tune:
        (cph, spl, th), tune_ms = tune_plain(sim)
        tuned = {"cells_per_hw": cph, "spl": spl, "eval_threads": th, "candidates_ms": tune_ms}
        log("tuned plain step: cells_per_hw=%d spl=%d eval_threads=%d (%.3f ms)"
            % (cph, spl, th, tune_ms["%d,%d,%d" % (cph, spl, th)]))
        from vortexparticlemc_cuda_amd.sim import Simulation
        s2 = Simulation(sim.P, device=device, cells_per_hw=cph, mode=sim.mode, seed=sim.seed,
                        spl=spl, eval_threads=th)
        s2.set_particles(sim.pos.cpu().numpy(), sim.omega.cpu().numpy())
        sim = s2
        args.cells_per_hw, args.spl, args.eval_threads = cph, spl, th
    err = None
    if not args.no_error:
        err = velocity_error(sim)
        log("error", err)
    run_steps(sim, args.warmup)
    torch.cuda.synchronize()
    log("warmup done")
    events = {"est": [], "eval": []}
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    run_steps(sim, args.steps, events)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    est_avg_s = float(np.mean([a.elapsed_time(b) for a, b in events["est"]])) / 1e3
    eval_avg_s = float(np.mean([a.elapsed_time(b) for a, b in events["eval"]])) / 1e3
    log("timed %d steps in %.3f s, estimator avg %.3f ms, evaluation kernel avg %.3f ms"
        % (args.steps, elapsed, est_avg_s * 1e3, eval_avg_s * 1e3))
    n_in, n_vis = count_evals(sim)
    N = sim.N
    S = N * 2 * (args.N1 + args.N2)
    eval_kernel = "k_bin_eval" if args.algo == "binned" else "k_mc_estimate"
    # the evaluation kernel's algorithmic work: 15 FLOP per in-support pair
    # (for the direct kernel, which also samples, 30 FLOP per sample on top)
    flop = FLOP_PER_EVAL * n_in + (FLOP_PER_SAMPLE * S if args.algo == "direct" else 0)
    achieved = flop / eval_avg_s
    traffic = read_traffic(eval_kernel, args)
    hbm = hbm_stages(sim)
    log("HBM-bound stages (plain workload)", hbm)
    fma_rate = fma_probe()
    log("measured unpacked FP32 FMA rate %.3e FLOP/s" % fma_rate)
    if args.copy_probe:
        copy_probe(args.copy_probe << 20)
    deform = deform_bench(args, device, sim) if args.deform_steps > 0 else None
    batched = ensemble_bench(args, device) if args.ens_sims > 0 else None
    large = large_bench(args, device) if args.large_Nx > 0 else None
    adj = adjoint_bench(args, device) if args.adj_Nx > 0 else None
    cpu = None if args.no_cpu else cpu_baseline(args.Nx, args.hwr, args.N1, args.N2)
    value = N * args.steps / elapsed
    line = {
        "metric": "particle-steps/s at fixed RMS vel. error vs FP64 direct sum; batched vs plain",
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (FP64 CDF/Omega)",
        "data": "synthetic",
        "config": {"workload": "plain GPU path: Taylor-Green on periodic [0,1]^2, "
                               "Nx=Ny=%d lattice (%d particles), hwr=%d, N1=N2=%d, dt=1e-3, "
                               "consistent estimator, Philox seed 0, CG-RBF strengths"
                               % (args.Nx, N, args.hwr, args.N1),
                   "particles": N, "N1": args.N1, "N2": args.N2, "hwr": args.hwr,
                   "cells_per_hw": args.cells_per_hw, "estimator_algo": args.algo,
                   "tuned": tuned,
                   "parallelism": "single GPU"},
        "roofline": {"kernel": eval_kernel, "bound": "valu", "achieved": achieved,
                     "peak": PEAK_FP32, "unit": "FLOP/s", "frac": achieved / PEAK_FP32,
                     "traffic": traffic,
                     "flop_per_launch": flop, "kernel_evals_per_launch": n_in,
                     "candidates_per_launch": n_vis, "samples_per_launch": S,
                     "avg_launch_ms": eval_avg_s * 1e3,
                     "unpacked_fp32_issue_peak": PEAK_FP32 / 2,
                     "frac_of_unpacked": achieved / (PEAK_FP32 / 2),
                     "measured_unpacked_fma_flops": fma_rate,
                     "frac_of_measured_unpacked": achieved / fma_rate},
        "kernel_evals_per_s": n_in / eval_avg_s,
        "hbm_stages": hbm,
        "estimator_ms": est_avg_s * 1e3,
        "estimator_fraction_of_step": est_avg_s / (elapsed / args.steps),
        "error": err,
        "cg_init": init,
        "deform": deform,
        "batched": batched,
        "large": large,
        "adjoint": adj,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
