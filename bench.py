"""Benchmark of the vortex-particle Monte Carlo step on one MI355X.

Headline (--steps K, --warmup W): the batched ensemble.  64 independent
periodic simulations on [0,1]^2 with Nx drawn from {64, 96, 128, 192, 256}
(seed 20261203), 8 Gaussian vortices each, hwr = 8, N1 = N2 = 64, dt = 5e-4,
consistent estimator (FP32 bulk, FP64 |omega| CDF and Omega), RBF strengths.
A step is one pass of the hot path for every simulation: rebuild both sign
fields, generate and bin the samples, evaluate, reduce, advect; all
simulations share one launch per stage and K steps replay from one HIP graph.
`value` is particle-steps/s over exactly K timed steps with the inputs
resident in HBM.  Beside it: the RMS velocity error against the FP64 blob
direct sum, the tuned plain per-simulation path and the reference's launch
structure on the same ensemble (batched_over_plain,
batched_over_reference_launch), and the evaluation kernel's roofline.

Also measured: the plain GPU path at the reference's default size
(Taylor-Green, Nx = 200, hwr = 16, N1 = N2 = 128) with an error-versus-samples
sweep and one step at the smallest N1 = N2 that reaches 1e-2 RMS error; the
CUDA step's semantics (aux targets + F); the large double-shear run
(1,048,576 particles, N1 = N2 = 256) with the reference step at that size in
owner chunks; the adjoint workload with its finite-difference check; and the
reference's own CPU path (oracle/_ref, compiled from its sources) on a bounded
sample of the headline workload and on the 10,000-particle Taylor-Green case
in full.
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

    rshx, rshy = sim.P.rsh

    def ref_hash():   # FillCells + CountingSort over every particle on the reference grid
        num, acc, cell = ops.fill_cells(rshx, rshy, float(sim.P.Lx), float(sim.P.Ly), sim.pos)
        ops.counting_sort(num, acc, cell)

    stages = (("field_build", fields, HBM_BYTES_FIELD),
              ("reference_grid_hash", ref_hash, HBM_BYTES_FIELD),
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


def ensemble_data(args, device):
    """the 64 Gaussian-vortex simulations with their RBF strengths"""
    import torch

    from vortexparticlemc_cuda_amd import ops
    from vortexparticlemc_cuda_amd.sim import SimParam
    from vortexparticlemc_cuda_amd.workloads import gaussian_ensemble

    data = gaussian_ensemble(args.ens_sims)
    Ps, pos, w = [], [], []
    t0 = time.perf_counter()
    for nx, p, om, _ in data:
        P = SimParam(dt=5e-4, resX=nx, hwr=8)
        P.N1, P.N2 = args.ens_N, args.ens_N
        # the reference's tolerance (src/cuda/cg_rbf_init.cu:53): tighter fits
        # of these fields amplify the Gram matrix's near-null modes into
        # alternating strengths, which inflate the estimator's variance
        wt, it, rr, st = ops.cg_rbf(torch.as_tensor(p, device=device),
                                    torch.as_tensor(om.astype(np.float32), device=device),
                                    float(P.h_w), tol=5e-3, max_iter=300, refine=True)
        Ps.append(P)
        pos.append(p)
        w.append(wt.cpu().numpy())
    return Ps, pos, w, time.perf_counter() - t0


def ensemble_error(ens, pos, w, hw, probes, device):
    """RMS |u_MC - u_direct| / RMS |u_direct| of the ensemble's last velocities
    (estimated at the positions `pos`) over up to `probes` random owners of
    every simulation, against the FP64 blob direct sum over that simulation"""
    import torch

    from vortexparticlemc_cuda_amd import ops
    rng = np.random.default_rng(20261203)
    v = ens.vel
    num = den = 0.0
    for s in range(len(pos)):
        n = pos[s].shape[0]
        idx = np.sort(rng.choice(n, size=min(n, probes), replace=False))
        pd = torch.as_tensor(pos[s], device=device).double()
        it = torch.as_tensor(idx, device=device)
        ud = ops.direct_sum_blob(pd[it].contiguous(), pd,
                                 torch.as_tensor(w[s], device=device).double(), hw[s])
        um = v[int(ens.offsets[s]):int(ens.offsets[s + 1])][it].double()
        num += float(((um - ud) ** 2).sum())
        den += float((ud ** 2).sum())
    return (num / den) ** 0.5


def plain_ensemble_loop(Ps, pos, w, shape, device, seed=0, N1=None):
    """one plain Simulation per member (the per-simulation GPU path); N1:
    N1 = N2 other than the members' own"""
    import copy

    from vortexparticlemc_cuda_amd.sim import Simulation
    cph, spl, th = shape
    sims = []
    for s in range(len(pos)):
        P = Ps[s]
        if N1 is not None:
            P = copy.copy(P)
            P.N1 = P.N2 = int(N1)
        sim = Simulation(P, device=device, cells_per_hw=cph, seed=seed, sim=s, spl=spl,
                         eval_threads=th)
        sim.set_particles(pos[s], w[s])
        sims.append(sim)
    return sims


def time_plain_loop(sims, steps):
    import torch
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        for sim in sims:
            sim.step()
    torch.cuda.synchronize()
    return time.perf_counter() - t0


def ensemble_headline(args, device):
    """The headline workload: 64 independent periodic simulations, Nx drawn
    from {64, 96, 128, 192, 256} (seed 20261203), 8 Gaussian vortices each,
    hwr = 8, N1 = N2 = 64, dt = 5e-4, consistent estimator, RBF strengths.
    Times exactly `steps` batched steps replayed from one HIP graph after
    `warmup` eager steps (and one untimed replay that uploads the graph).
    Beside it: the RMS error of the batched estimate against the FP64 blob
    direct sum (probe owners of every member), the tuned plain per-simulation
    path over >= 10 steps and the reference's launch structure (one workgroup
    per owner and sign) on the same ensemble, each with its own label, the
    k_ens_eval roofline (HIP events the library records around the kernel),
    and the same ensemble with kernel deformation."""
    import torch

    from vortexparticlemc_cuda_amd._capi import ALGO_DIRECT
    from vortexparticlemc_cuda_amd.ensemble import Ensemble, autotune

    N1 = args.ens_N
    dt = 5e-4
    Ps, pos, w, t_cg = ensemble_data(args, device)
    hw = [float(P.h_w) for P in Ps]
    N = sum(p.shape[0] for p in pos)
    sizes = sorted(set(int(P.ResolutionX) for P in Ps))
    log("ensemble: %d sims (Nx in %s), %d particles, RBF init %.1f s"
        % (len(pos), sizes, N, t_cg))
    if args.ens_shape:   # a fixed launch shape (profiling), no tuning
        cph, spl, th = (int(v) for v in args.ens_shape.split(","))
        tune_ms = {}
        log("launch shape (given): cells_per_hw=%d spl=%d eval_threads=%d" % (cph, spl, th))
    else:
        cph, spl, th, tune_ms = autotune(pos, w, hw, N1=N1, N2=N1, device=device)
        log("autotune: cells_per_hw=%d spl=%d eval_threads=%d" % (cph, spl, th))
    ens = Ensemble(pos, w, hw, N1=N1, N2=N1, dt=dt, seed=0, cells_per_hw=cph, spl=spl,
                   eval_threads=th, device=device)
    ens.step(1)   # step 0: velocities at the initial positions
    err = ensemble_error(ens, pos, w, hw, args.ens_probes, device)
    log("ensemble: RMS error vs FP64 blob direct sum %.4g (%d probe owners per member)"
        % (err, args.ens_probes))
    K, W = args.steps, args.warmup
    ens.step(W)
    torch.cuda.synchronize()
    log("ensemble: %d warmup steps done" % W)
    side = torch.cuda.Stream()
    with torch.cuda.stream(side):
        ens.capture(K)
        log("ensemble: graph of %d steps captured" % K)
        ens.replay(K)   # the first replay uploads the graph (untimed)
        torch.cuda.synchronize()
        log("ensemble: first replay done")
        t0 = time.perf_counter()
        ens.replay(K)
        torch.cuda.synchronize()
        t_graph = time.perf_counter() - t0
    torch.cuda.synchronize()
    log("ensemble: timed replay %.2f ms/step" % (t_graph / K * 1e3))
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
    eval_s = float(np.mean([a.elapsed_time(b) for a, b in ev_pairs])) / 1e3
    ens.close()
    batched = N * K / t_graph
    del ens
    torch.cuda.empty_cache()
    log("batched: %.3g particle-steps/s (graph %.2f ms/step, eager %.2f ms/step)"
        % (batched, t_graph / K * 1e3, t_eager / K * 1e3))
    # the same ensemble with kernel deformation (the CUDA step: four aux
    # targets per owner, F in the fields and the jitter), same launch shape
    ensd = Ensemble(pos, w, hw, N1=N1, N2=N1, dt=dt, seed=0, cells_per_hw=cph, spl=spl,
                    eval_threads=th, device=device, deform=True)
    ensd.step(1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ensd.step(args.ens_deform_steps)
    torch.cuda.synchronize()
    t_def = (time.perf_counter() - t0) / args.ens_deform_steps
    ensd.close()
    del ensd
    torch.cuda.empty_cache()
    log("batched deformed: %.1f ms/step" % (t_def * 1e3))
    # evaluations within h_w and candidates visited per step: a counting pass
    # of each member at step 0 with the ensemble's launch shape (the same
    # Philox streams and samples as the batched step 0)
    sims = plain_ensemble_loop(Ps, pos, w, (cph, spl, th), device)
    n_in = n_vis = 0
    for sim in sims:
        sim.eval_count.zero_()
        sim.estimate(N1, N1, count=True)
        c = sim.eval_count.cpu().numpy()
        n_in += int(c[0])
        n_vis += int(c[1])
    # the tuned plain path: the per-simulation loop with the launch shape
    # that steps it fastest (one timed step per candidate)
    plain_tune = {}
    for shape in ((cph, spl, th),) + tuple(s for s in PLAIN_ENS_TUNE if s != (cph, spl, th)):
        if shape != (cph, spl, th):
            del sims
            torch.cuda.empty_cache()
            try:
                sims = plain_ensemble_loop(Ps, pos, w, shape, device)
            except Exception:
                sims = None
                continue
        for sim in sims:
            sim.step()
        plain_tune["%d,%d,%d" % shape] = time_plain_loop(sims, 1) * 1e3
    best = min(plain_tune, key=plain_tune.get)
    pshape = tuple(int(v) for v in best.split(","))
    del sims
    torch.cuda.empty_cache()
    sims = plain_ensemble_loop(Ps, pos, w, pshape, device)
    for sim in sims:
        sim.step()
    Kp = max(1, args.ens_plain_steps)
    t_plain = time_plain_loop(sims, Kp)
    plain = N * Kp / t_plain
    log("plain per-simulation path (shape %s): %.3g particle-steps/s (%.1f ms/step)"
        % (best, plain, t_plain / Kp * 1e3))
    # the reference's launch structure per simulation: one workgroup per
    # (owner, sign) sampling and querying on its own
    # (monte_carlo_bs_kernel_safe, src/cuda/monte_carlo_bs.cu:144-271)
    for sim in sims:
        sim.algo = ALGO_DIRECT
        sim.step()
    Kd = max(1, args.ens_direct_steps)
    t_direct = time_plain_loop(sims, Kd)
    per_launch = N * Kd / t_direct
    del sims
    torch.cuda.empty_cache()
    trend = [{"N1_eq_N2": N1, "rms_err_vs_direct": err, "batched_ms_per_step": t_graph / K * 1e3,
              "plain_ms_per_step": t_plain / Kp * 1e3, "batched_over_plain": batched / plain}]
    for n1 in args.ens_N_trend:   # lower errors: more samples, same launch shapes
        trend.append(ensemble_trend_point(Ps, pos, w, hw, n1, dt, (cph, spl, th), pshape,
                                          args.ens_probes, device))
    flop = FLOP_PER_EVAL * n_in
    return {"workload": "64 periodic simulations on [0,1]^2, Nx from {64,96,128,192,256} "
                        "(seed 20261203), 8 Gaussian vortices each, hwr=8, N1=N2=%d, dt=5e-4, "
                        "consistent estimator, FP32 bulk with FP64 CDF/Omega, RBF strengths "
                        "(tol 5e-3)" % N1,
            "simulations": len(pos), "particles": N, "lattice_sides": sizes,
            "samples_per_step": N * 2 * 2 * N1,
            "batched_particle_steps_per_s": batched,
            "batched_ms_per_step_graph": t_graph / K * 1e3,
            "batched_ms_per_step_eager": t_eager / K * 1e3,
            "steps_timed": K, "warmup": W,
            "rms_err_vs_direct": err, "error_probe_owners_per_member": args.ens_probes,
            "plain_particle_steps_per_s": plain, "plain_ms_per_step": t_plain / Kp * 1e3,
            "plain_steps_timed": Kp, "plain_launch_shape": best, "plain_tune_ms": plain_tune,
            "batched_over_plain": batched / plain,
            "reference_launch_particle_steps_per_s": per_launch,
            "reference_launch_ms_per_step": t_direct / Kd * 1e3,
            "batched_over_reference_launch": batched / per_launch,
            "reference_launch_note": "per-simulation loop with one workgroup per (owner, sign), "
                                     "the launch structure of monte_carlo_bs_kernel_safe",
            "target_batched_over_plain": 5.0,
            # batched over plain as the error falls (N1 = N2 grows; the two
            # paths compute the same estimates, so their errors are equal)
            "error_trend": trend,
            "cells_per_hw": cph, "spl": spl, "eval_threads": th, "autotune_ms": tune_ms,
            # kernel evaluations within h_w per second: over the evaluation
            # kernel's own time, and over the whole batched step
            "kernel_evals_per_s": n_in / eval_s,
            "kernel_evals_per_s_step": n_in / (t_graph / K),
            "deformed": {"ms_per_step": t_def * 1e3, "particle_steps_per_s": N / t_def,
                         "samples_per_pair": N1 + 5 * N1},
            "roofline": {"kernel": "k_ens_eval", "bound": "valu",
                         "achieved": flop / eval_s, "peak": PEAK_FP32, "unit": "FLOP/s",
                         "frac": flop / eval_s / PEAK_FP32, "avg_launch_ms": eval_s * 1e3,
                         "flop_per_launch": flop, "kernel_evals_per_launch": n_in,
                         "candidates_per_launch": n_vis,
                         # the whole step: 15 FLOP per evaluation within h_w and 30 per
                         # sample (MIS contribution and sampling) over the step's time
                         "step_frac_of_fp32_peak": (flop + FLOP_PER_SAMPLE * N * 2 * 2 * N1) /
                                                   (t_graph / K) / PEAK_FP32},
            "rbf_init_s": t_cg, "_data": (Ps, pos, w)}


def ensemble_trend_point(Ps, pos, w, hw, n1, dt, shape, pshape, probes, device, K=3):
    """the ensemble at N1 = N2 = n1: its RMS error (step 0), the batched step
    (eager, launch shape `shape`) and the plain per-simulation loop (launch
    shape `pshape`), each timed over K steps"""
    import torch

    from vortexparticlemc_cuda_amd.ensemble import Ensemble
    cph, spl, th = shape
    ens = Ensemble(pos, w, hw, N1=n1, N2=n1, dt=dt, seed=0, cells_per_hw=cph, spl=spl,
                   eval_threads=th, device=device)
    ens.step(1)
    err = ensemble_error(ens, pos, w, hw, probes, device)
    ens.step(1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ens.step(K)
    torch.cuda.synchronize()
    tb = (time.perf_counter() - t0) / K
    ens.close()
    del ens
    torch.cuda.empty_cache()
    sims = plain_ensemble_loop(Ps, pos, w, pshape, device, N1=n1)
    for sim in sims:
        sim.step()
    tp = time_plain_loop(sims, K) / K
    del sims
    torch.cuda.empty_cache()
    log("ensemble at N1=N2=%d: RMS error %.4g, batched %.1f ms/step, plain %.1f ms/step (%.2fx)"
        % (n1, err, tb * 1e3, tp * 1e3, tp / tb))
    return {"N1_eq_N2": n1, "rms_err_vs_direct": err, "batched_ms_per_step": tb * 1e3,
            "plain_ms_per_step": tp * 1e3, "batched_over_plain": tp / tb,
            "steps_timed": K, "timing": "eager steps (no graph) for both paths"}


PLAIN_ENS_TUNE = ((8, 1, 256), (6, 1, 256), (5, 1, 256), (4, 1, 256), (8, 2, 256), (12, 1, 256))


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
                        "samples_per_launch": n * 2 * 2 * N},
           # the whole estimator's algorithmic rate: every kernel of the
           # estimate (generation, sort, evaluation, reduction) in the time
           "estimator_frac_of_fp32_peak": flop / est_s / PEAK_FP32}
    shape = (sim.cells_per_hw, sim.spl, sim.eval_threads)
    pos_now, w_now = sim.pos.cpu().numpy(), sim.omega.cpu().numpy()
    del sim
    torch.cuda.empty_cache()
    if args.large_ref_steps > 0:
        # the reference's step semantics at this size (step_cuda,
        # src/cuda/simulation_step.cu:51-73: aux targets and F every step):
        # N1 + 5 N2 samples per owner and sign, 3.2e9 sample slots a step, run
        # in owner chunks of 2^30 slots
        simd = Simulation(P, device=device, cells_per_hw=shape[0], seed=0, spl=shape[1],
                          eval_threads=shape[2], deform=True)
        simd.set_particles(pos_now, w_now)
        simd.step(N, N)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.large_ref_steps):
            simd.step(N, N)
        torch.cuda.synchronize()
        t_ref = (time.perf_counter() - t0) / args.large_ref_steps
        F = simd.jac[:simd.N].double()
        det = F[:, 0] * F[:, 3] - F[:, 1] * F[:, 2]
        out["reference_step"] = {
            "semantics": "aux targets at dst -/+ dx and kernel deformation F every step",
            "samples_per_step": n * 2 * (N + 5 * N), "chunk_samples": 1 << 30,
            "ms_per_step": t_ref * 1e3, "particle_steps_per_s": n / t_ref,
            "steps_timed": args.large_ref_steps,
            "max_abs_det_minus_1": float((det - 1).abs().max())}
        log("large: reference step (aux + F) %.1f ms/step" % (t_ref * 1e3))
        del simd
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
        X_ref = A.rollout(sim, pos_t, w_tg, T, checkpoint=K, trajectory="f32")
        ref = A.grid_vorticity(X_ref, w_tg, h, G, G, f64=True)
    p = pos_t.clone().requires_grad_(True)
    wt = w0.clone().requires_grad_(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    XT = A.rollout(sim, p, wt, T, checkpoint=K, trajectory="f32")
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
    if args.adj_deform_fd > 0:
        out["deformed"] = adjoint_deformed_bench(args, device, pos, pos_t, w_tg, w0, rng)
    return out


def adjoint_deformed_bench(args, device, pos, pos_t, w_tg, w0, rng):
    """The adjoint workload with the reference's deformed CUDA step (aux
    targets and F every step, F = I at the start; the reference run deformed
    too): forward and backward along the FP64 trajectory, and dL/dx0, dL/dw
    against central differences (eps = 1e-6) of the FP64 deformed rollout on
    `adj_deform_fd` coordinates."""
    import torch

    from vortexparticlemc_cuda_amd import adjoint as A
    from vortexparticlemc_cuda_amd.sim import SimParam, Simulation

    Nx, hwr, N, T, K, G = args.adj_Nx, 16, 64, args.adj_steps, 10, 64
    P = SimParam(dt=1e-3, resX=Nx, hwr=hwr)
    P.N1 = P.N2 = N
    h = float(P.h_w)
    sim = Simulation(P, device=device, cells_per_hw=8, seed=0, deform=True)
    sim.set_particles(pos, w_tg.cpu().numpy())
    n = pos.shape[0]
    I = torch.tensor([1.0, 0.0, 0.0, 1.0], device=device).repeat(n, 1)
    with torch.no_grad():
        X_ref = A.rollout_f64(sim, pos_t.double(), w_tg.double(), T, F0=I)
        ref = A.grid_vorticity_f64(X_ref, w_tg.double(), h, G, G)
    sim.set_particles(pos, w_tg.cpu().numpy())   # F = I again
    p = pos_t.clone().requires_grad_(True)
    wt = w0.clone().requires_grad_(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    XT = A.rollout(sim, p, wt, T, checkpoint=K, trajectory="f64")
    loss = ((A.grid_vorticity(XT, wt, h, G, G, f64=True) - ref) ** 2).mean()
    torch.cuda.synchronize()
    t_fwd = time.perf_counter() - t0
    max_det = float((sim.jac[:n, 0] * sim.jac[:n, 3] - sim.jac[:n, 1] * sim.jac[:n, 2] - 1)
                    .abs().max())
    max_FmI = float((sim.jac[:n] - I).abs().max())
    t0 = time.perf_counter()
    loss.backward()
    torch.cuda.synchronize()
    t_bwd = time.perf_counter() - t0
    gp = p.grad.double().cpu().numpy()
    gw = wt.grad.double().cpu().numpy()
    log("adjoint, deformed step: loss %.9g, forward %.3f s, backward %.3f s, max |F_T - I| %.3g"
        % (float(loss.detach()), t_fwd, t_bwd, max_FmI))
    p64, w64 = pos_t.double(), w0.double()

    def L(pp, ww):
        X = A.rollout_f64(sim, pp, ww, T, F0=I)
        o = A.grid_vorticity_f64(X, ww, h, G, G)
        return float(((o - ref) ** 2).mean())

    t0 = time.perf_counter()
    wn = w0.abs().cpu().numpy()
    cand = np.nonzero(wn > 0.2 * wn.max())[0]
    pick = rng.choice(cand, size=args.adj_deform_fd, replace=False)
    axis = rng.integers(2, size=args.adj_deform_fd)
    eps_x, eps_w = 1e-6, 1e-6 * float(w0.abs().max())
    ex, ew = [], []
    for c, i in enumerate(pick):
        if c % 2 == 0:
            k = int(axis[c])
            a, b = p64.clone(), p64.clone()
            a[i, k] += eps_x
            b[i, k] -= eps_x
            fd = (L(a, w64) - L(b, w64)) / (2 * eps_x)
            ex.append(abs(fd - gp[i, k]) / np.abs(gp).max())
        else:
            a, b = w64.clone(), w64.clone()
            a[i] += eps_w
            b[i] -= eps_w
            fd = (L(p64, a) - L(p64, b)) / (2 * eps_w)
            ew.append(abs(fd - gw[i]) / np.abs(gw).max())
    t_fd = time.perf_counter() - t0
    log("adjoint, deformed step, vs FP64 central differences (eps 1e-6): max err/max|g| "
        "positions %.3g, strengths %.3g (%d coordinates, %.1f s)"
        % (max(ex), max(ew), len(pick), t_fd))
    del sim
    torch.cuda.empty_cache()
    return {"semantics": "aux targets at dst -/+ dx and kernel deformation F every step "
                         "(F = I at the start), FP64 trajectory",
            "loss": float(loss.detach()), "forward_s": t_fwd, "backward_s": t_bwd,
            "max_abs_F_minus_I": max_FmI, "max_abs_det_minus_1": max_det,
            "fd_coordinates": int(len(pick)), "fd_max_err_pos_eps1e-6": max(ex),
            "fd_max_err_w_eps1e-6": max(ew), "fd_seconds": t_fd}


def read_traffic(kernel, workload, want):
    """HBM bytes per launch of `kernel` on `workload`, from the committed
    counter summary profiles/traffic_<kernel>_<workload>.txt (FETCH_SIZE +
    WRITE_SIZE, copy-calibrated) when its workload line carries `want`;
    None if absent"""
    path = os.path.join(ROOT, "profiles", "traffic_%s_%s.txt" % (kernel, workload))
    if not os.path.exists(path):
        return None
    ok = False
    for line in open(path):
        if line.startswith("workload:"):
            ok = want in line
        if ok and line.startswith("hbm_bytes_per_launch:"):
            return float(line.split(":", 1)[1])
    return None


def fixed_error(sim, args, target=1e-2):
    """RMS error against the FP64 blob direct sum versus N1 = N2 on the plain
    workload's initial state, the smallest N1 = N2 that the 1/sqrt(N1 + N2)
    law fitted to those points predicts for `target` (plus 15 %), the error
    measured there and one timed step there (owners in chunks)"""
    import torch

    from vortexparticlemc_cuda_amd import ops
    pos0 = sim.pos.clone()
    pos64 = pos0.double()
    ud = ops.direct_sum_blob(pos64, pos64, sim.omega.double(), float(sim.P.h_w))
    rms = lambda a: float(torch.sqrt((a * a).sum(1).mean()))  # noqa: E731
    sweep = {}
    for n in args.error_sweep:
        sim.pos.copy_(pos0)
        u = sim.estimate(n, n).double()
        sweep[n] = rms(u - ud) / rms(ud)
    ns = np.array(sorted(sweep), np.float64)
    es = np.array([sweep[int(n)] for n in ns])
    c = float(np.exp(np.mean(np.log(es) + 0.5 * np.log(2 * ns))))   # err = c / sqrt(N1+N2)
    n_star = int(np.ceil((c / target) ** 2 / 2 * 1.15))
    S = sim.N * 2 * 2 * n_star
    out = {"target": target, "oracle": "FP64 blob direct sum at every particle",
           "sweep_N1_eq_N2": {str(int(n)): sweep[int(n)] for n in ns},
           "fit_err_times_sqrt_N1_plus_N2": c, "N1_eq_N2": n_star, "samples_per_step": S}
    if S > args.fixed_error_max_samples:
        out["skipped"] = "%.3g samples a step exceed --fixed-error-max-samples" % S
        return out
    sim.pos.copy_(pos0)
    u = sim.estimate(n_star, n_star).double()
    out["rms_err_vs_direct"] = rms(u - ud) / rms(ud)
    sim.pos.copy_(pos0)
    sim.step(n_star, n_star)   # warm-up at this size (workspace, chunks)
    sim.pos.copy_(pos0)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    sim.step(n_star, n_star)
    torch.cuda.synchronize()
    t = time.perf_counter() - t0
    sim.pos.copy_(pos0)
    sim.ws = None
    torch.cuda.empty_cache()
    out.update({"ms_per_step": t * 1e3, "particle_steps_per_s": sim.N / t})
    log("fixed error: N1=N2=%d, RMS error %.4g, %.1f ms/step" % (n_star, out["rms_err_vs_direct"],
                                                                t * 1e3))
    return out


def plain_section(args, device):
    """The plain GPU path at the reference's default size: Taylor-Green,
    Nx = Ny = 200 (40,000 particles), hwr = 16, N1 = N2 = 128, dt = 1e-3,
    consistent estimator, Philox seed 0, RBF strengths; tuned launch shape;
    `plain_steps` timed steps; the k_bin_eval roofline; the HBM-bound stages;
    the fixed-error measurement"""
    import torch

    from vortexparticlemc_cuda_amd.sim import Simulation
    log("plain: setup Nx=%d hwr=%d N1=%d N2=%d" % (args.Nx, args.hwr, args.N1, args.N2))
    sim, init = setup(args.Nx, args.hwr, args.N1, args.N2, device,
                      cells_per_hw=args.cells_per_hw, spl=args.spl,
                      eval_threads=args.eval_threads)
    sim.algo = 0 if args.algo == "binned" else 1
    tuned = None
    if args.algo == "binned" and not args.no_tune:
        (cph, spl, th), tune_ms = tune_plain(sim)
        tuned = {"cells_per_hw": cph, "spl": spl, "eval_threads": th, "candidates_ms": tune_ms}
        log("tuned plain step: cells_per_hw=%d spl=%d eval_threads=%d (%.3f ms)"
            % (cph, spl, th, tune_ms["%d,%d,%d" % (cph, spl, th)]))
        s2 = Simulation(sim.P, device=device, cells_per_hw=cph, mode=sim.mode, seed=sim.seed,
                        spl=spl, eval_threads=th)
        s2.set_particles(sim.pos.cpu().numpy(), sim.omega.cpu().numpy())
        sim = s2
        args.cells_per_hw, args.spl, args.eval_threads = cph, spl, th
    err = None if args.no_error else velocity_error(sim)
    fixed = None
    if not args.no_error and args.error_sweep:
        fixed = fixed_error(sim, args)
    run_steps(sim, args.plain_warmup)
    events = {"est": [], "eval": []}
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    run_steps(sim, args.plain_steps, events)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    est_s = float(np.mean([a.elapsed_time(b) for a, b in events["est"]])) / 1e3
    eval_s = float(np.mean([a.elapsed_time(b) for a, b in events["eval"]])) / 1e3
    n_in, n_vis = count_evals(sim)
    N = sim.N
    S = N * 2 * (args.N1 + args.N2)
    kernel = "k_bin_eval" if args.algo == "binned" else "k_mc_estimate"
    flop = FLOP_PER_EVAL * n_in + (FLOP_PER_SAMPLE * S if args.algo == "direct" else 0)
    want = "N1=%d N2=%d hwr=%d cells_per_hw=%d spl=%d eval_threads=%d" % (
        args.N1, args.N2, args.hwr, args.cells_per_hw, args.spl, args.eval_threads)
    hbm = hbm_stages(sim)
    log("plain: %.3f ms/step, estimator %.3f ms, evaluation %.3f ms"
        % (elapsed / args.plain_steps * 1e3, est_s * 1e3, eval_s * 1e3))
    out = {"workload": "Taylor-Green on periodic [0,1]^2, Nx=Ny=%d lattice (%d particles), "
                       "hwr=%d, N1=N2=%d, dt=1e-3, consistent estimator, Philox seed 0, RBF "
                       "strengths" % (args.Nx, N, args.hwr, args.N1),
           "particles": N, "steps_timed": args.plain_steps,
           "particle_steps_per_s": N * args.plain_steps / elapsed,
           "ms_per_step": elapsed / args.plain_steps * 1e3, "estimator_ms": est_s * 1e3,
           "estimator_fraction_of_step": est_s / (elapsed / args.plain_steps),
           "tuned": tuned, "error": err, "fixed_error": fixed, "rbf_init": init,
           "kernel_evals_per_s": n_in / eval_s, "hbm_stages": hbm,
           "roofline": {"kernel": kernel, "bound": "valu", "achieved": flop / eval_s,
                        "peak": PEAK_FP32, "unit": "FLOP/s", "frac": flop / eval_s / PEAK_FP32,
                        "traffic": read_traffic(kernel, "tg%d" % args.Nx, want),
                        "flop_per_launch": flop, "kernel_evals_per_launch": n_in,
                        "candidates_per_launch": n_vis, "samples_per_launch": S,
                        "avg_launch_ms": eval_s * 1e3}}
    return sim, out


def _threads():
    return max(1, min(16, os.cpu_count() or 1))


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            return next(l.split(":", 1)[1].strip() for l in f if l.startswith("model name"))
    except (OSError, StopIteration):
        return "unknown"


def cpu_baseline_ensemble(Ps, pos, w, N1, target_s=10.0):
    """the reference's monte_carlo_bs_cpu (its src/cpu sources compiled into
    oracle/_ref) over a bounded sample of the ensemble's owners: four members
    of different sizes, both signs, aux = 0 (the work the GPU step does),
    OpenMP over owners; particle-steps/s of the sample, which scales to the
    whole ensemble because every query here visits the same 3 x 3 cells of
    1152 same-sign candidates (hwr = 8)"""
    import oracle as O
    R = O.ref_lib()
    if R is None:
        return None
    threads = _threads()
    picked, seen = [], set()
    for s, P in enumerate(Ps):
        if P.ResolutionX not in seen:
            seen.add(P.ResolutionX)
            picked.append(s)
    picked = picked[:4]

    def run(s, m):
        x = np.ascontiguousarray(pos[s][:, 0], np.float32)
        y = np.ascontiguousarray(pos[s][:, 1], np.float32)
        ww = np.ascontiguousarray(w[s], np.float32)
        rsh = O.rsh_of(Ps[s].ResolutionX, 8)
        vel = np.zeros(2 * m, np.float32)
        t0 = time.perf_counter()
        R.ref_mc_velocities_mt(x.size, x, y, ww, rsh, rsh, 1.0, 1.0, np.float32(Ps[s].h_w), 1,
                               N1, N1, 5489, 0, m, threads, vel)
        return time.perf_counter() - t0

    probe = 8 * threads
    rate = probe / run(picked[0], probe)
    owners = tsum = 0.0
    for s in picked:
        m = int(max(probe, min(pos[s].shape[0], rate * target_s / len(picked))))
        tsum += run(s, m)
        owners += m
    return {"value": owners / tsum, "unit": "particle-steps/s", "cores": threads,
            "kind": "reference",
            "sample": "reference monte_carlo_bs_cpu (src/cpu, compiled from source) for %d owners "
                      "of members %s of the 64-simulation ensemble (Nx %s), both signs, "
                      "N1=N2=%d, aux=0, OpenMP over owners on %d threads (%s), %.1f s"
                      % (owners, picked, [Ps[s].ResolutionX for s in picked], N1, threads,
                         _cpu_model(), tsum)}


def cpu_config1_full(steps=10):
    """the plain CPU-runnable workload in full: Taylor-Green, Nx = Ny = 100
    (10,000 particles), hwr = 16, N1 = N2 = 64, dt = 1e-3, 10 steps, each the
    reference's monte_carlo_bs_cpu for every particle and both signs (aux = 0)
    on all host threads, then x += u dt and the periodic wrap
    (src/cpu/simulation_step.cpp:46-71); strengths from the lattice's closed
    form of the RBF solve"""
    import oracle as O
    R = O.ref_lib()
    if R is None:
        return None
    threads = _threads()
    Nx, hwr, N1, dt = 100, 16, 64, np.float32(1e-3)
    x, y = O.lattice(Nx)
    x = x.astype(np.float32)
    y = y.astype(np.float32)
    w = (O.tg_vorticity(x, y) / O.tg_lambda11(Nx, hwr)).astype(np.float32)
    rsh = O.rsh_of(Nx, hwr)
    n = x.size
    vel = np.zeros(2 * n, np.float32)
    t0 = time.perf_counter()
    for k in range(steps):
        R.ref_mc_velocities_mt(n, x, y, w, rsh, rsh, 1.0, 1.0, np.float32(hwr / Nx), 1, N1, N1,
                               5489 + k, 0, n, threads, vel)
        x = x + vel[0::2] * dt
        y = y + vel[1::2] * dt
        x = (x - np.floor(x)).astype(np.float32)
        y = (y - np.floor(y)).astype(np.float32)
    t = time.perf_counter() - t0
    return {"particle_steps_per_s": n * steps / t, "seconds": t, "steps": steps,
            "particles": n, "cores": threads, "kind": "reference",
            "workload": "Taylor-Green, Nx=Ny=100, hwr=16, N1=N2=64, dt=1e-3, 10 steps, aux=0 "
                        "(the reference's step_cpu runs aux=1, about 3x this work)"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20, help="timed batched steps (headline)")
    ap.add_argument("--warmup", type=int, default=5, help="untimed batched steps first")
    ap.add_argument("--ens-sims", type=int, default=64)
    ap.add_argument("--ens-N", type=int, default=64, help="N1 = N2 of the ensemble")
    ap.add_argument("--ens-probes", type=int, default=512,
                    help="probe owners per member for the ensemble's error")
    ap.add_argument("--ens-shape", default="",
                    help="cells_per_hw,spl,eval_threads of the ensemble instead of the tuner's")
    ap.add_argument("--ens-plain-steps", type=int, default=10)
    ap.add_argument("--ens-N-trend", type=int, nargs="*", default=[256],
                    help="N1 = N2 values of the batched-over-plain error trend (none: skip)")
    ap.add_argument("--ens-direct-steps", type=int, default=2)
    ap.add_argument("--ens-deform-steps", type=int, default=5)
    ap.add_argument("--Nx", type=int, default=200)
    ap.add_argument("--hwr", type=int, default=16)
    ap.add_argument("--N1", type=int, default=128)
    ap.add_argument("--N2", type=int, default=128)
    ap.add_argument("--plain-steps", type=int, default=20)
    ap.add_argument("--plain-warmup", type=int, default=5)
    ap.add_argument("--cells-per-hw", type=int, default=8,
                    help="fine-grid cells per h_w of the plain path (with --no-tune)")
    ap.add_argument("--spl", type=int, default=0, help="samples per lane (with --no-tune)")
    ap.add_argument("--eval-threads", type=int, default=0,
                    help="evaluation workgroup size (with --no-tune)")
    ap.add_argument("--algo", choices=["binned", "direct"], default="binned")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-error", action="store_true")
    ap.add_argument("--no-tune", action="store_true",
                    help="skip the launch-shape tuning of the plain step")
    ap.add_argument("--error-sweep", type=int, nargs="*", default=[32, 128, 512, 2048, 8192],
                    help="N1 = N2 values of the plain workload's error sweep (none: skip it "
                         "and the fixed-error step)")
    ap.add_argument("--fixed-error-max-samples", type=float, default=4e10)
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="also run a plain copy of MIB MiB (HBM counter calibration)")
    ap.add_argument("--deform-steps", type=int, default=10,
                    help="steps of the deformed (aux + F) step on the plain workload (0: skip)")
    ap.add_argument("--large-Nx", type=int, default=1024,
                    help="lattice side of the large double-shear run (0: skip it)")
    ap.add_argument("--large-N", type=int, default=256, help="N1 = N2 of the large run")
    ap.add_argument("--large-steps", type=int, default=10)
    ap.add_argument("--large-ref-steps", type=int, default=2,
                    help="timed steps of the reference step semantics (aux + F) at the large size")
    ap.add_argument("--large-probes", type=int, default=4096)
    ap.add_argument("--large-cells-per-hw", type=int, default=12)
    ap.add_argument("--large-spl", type=int, default=1)
    ap.add_argument("--large-threads", type=int, default=256)
    ap.add_argument("--large-tune", type=int, default=1,
                    help="try a few launch shapes on the large run (0: use the flags)")
    ap.add_argument("--adj-Nx", type=int, default=128,
                    help="lattice side of the adjoint workload (0: skip it)")
    ap.add_argument("--adj-steps", type=int, default=50)
    ap.add_argument("--adj-deform-fd", type=int, default=6,
                    help="coordinates of the deformed adjoint's FD check (0: no deformed run)")
    ap.add_argument("--adj-fd", type=int, default=16,
                    help="coordinates checked by FP64 central differences")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU; this build has no multi-GPU path "
              "(--gpus %d requested)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import torch
    device = torch.device("cuda:0")
    head = ensemble_headline(args, device) if args.ens_sims > 0 else None
    data = head.pop("_data") if head else None
    sim, plain = plain_section(args, device)
    fma_rate = fma_probe()
    log("measured unpacked FP32 FMA rate %.3e FLOP/s" % fma_rate)
    if args.copy_probe:
        copy_probe(args.copy_probe << 20)
    deform = deform_bench(args, device, sim) if args.deform_steps > 0 else None
    del sim
    torch.cuda.empty_cache()
    large = large_bench(args, device) if args.large_Nx > 0 else None
    adj = adjoint_bench(args, device) if args.adj_Nx > 0 else None
    cpu = cpu1 = None
    if not args.no_cpu:
        if data:
            cpu = cpu_baseline_ensemble(*data, args.ens_N)
        cpu1 = cpu_config1_full()
    for sec in (plain, head):
        if sec:
            sec["roofline"]["measured_unpacked_fma_flops"] = fma_rate
            sec["roofline"]["frac_of_measured_unpacked"] = sec["roofline"]["achieved"] / fma_rate
    if head:
        rf = dict(head["roofline"])
        rf["traffic"] = read_traffic("k_ens_eval", "ensemble64", "N1=N2=%d cells_per_hw=%d spl=%d "
                                     "eval_threads=%d" % (args.ens_N, head["cells_per_hw"],
                                                          head["spl"], head["eval_threads"]))
        value, K, ms = head["batched_particle_steps_per_s"], args.steps, \
            head["batched_ms_per_step_graph"]
        config = {"workload": "batched ensemble: " + head["workload"] + "; one launch per stage "
                              "for all simulations, step replayed from a HIP graph",
                  "simulations": head["simulations"], "particles": head["particles"],
                  "N1": args.ens_N, "N2": args.ens_N, "hwr": 8,
                  "launch_shape": {"cells_per_hw": head["cells_per_hw"], "spl": head["spl"],
                                   "eval_threads": head["eval_threads"]},
                  "parallelism": "single GPU"}
    else:   # --ens-sims 0: the plain workload carries the line
        rf = plain["roofline"]
        value, K, ms = plain["particle_steps_per_s"], args.plain_steps, plain["ms_per_step"]
        config = {"workload": "plain GPU path: " + plain["workload"], "parallelism": "single GPU"}
    line = {
        "metric": "particle-steps/s at fixed RMS vel. error vs FP64 direct sum; batched vs plain",
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (FP64 CDF/Omega)",
        "data": "synthetic",
        "config": config,
        "rms_err_vs_direct": head["rms_err_vs_direct"] if head else
        (plain["error"] or {}).get("rms_err_vs_direct"),
        "batched_over_plain": head["batched_over_plain"] if head else None,
        "batched_over_reference_launch": head["batched_over_reference_launch"] if head else None,
        "fixed_error": plain["fixed_error"],
        "roofline": rf,
        "batched": head,
        "plain": plain,
        "deform": deform,
        "large": large,
        "adjoint": adj,
        "cpu_baseline": cpu,
        "cpu_config1_full": cpu1,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
