"""Ensemble of independent periodic simulations advanced together.

The reference runs one Simulation per process; this class drives the
library's ensemble entry points (vpmc_ensemble_*): every stage of a step is
one launch for all simulations, and a sequence of steps can be captured into
a HIP graph and replayed.  Simulation s evolves exactly as a plain
Simulation(sim=s) with the same seed, cells_per_hw and spl would.

Every step, replay and read runs on torch's current stream.  All of them
share one workspace and the particle state, so work issued on one stream
waits for what the ensemble issued earlier on another (a graph replayed on a
side stream after eager steps on the default stream, say): two steps that
overlapped would sort into the same buffers.
"""
import ctypes as C

import numpy as np
import torch

from . import ops
from ._capi import MODE_CONSISTENT, EnsembleDesc, check, lib


class Ensemble:
    def __init__(self, positions, strengths, h_w, L=1.0, N1=64, N2=64, dt=5e-4,
                 mode=MODE_CONSISTENT, seed=0, cells_per_hw=8, spl=0, eval_threads=0,
                 device="cuda", deform=False):
        """positions[s]: (N_s, 2) array, strengths[s]: (N_s,), h_w: per-simulation
        kernel radius (scalar or sequence), L: box side (scalar or sequence);
        deform: the CUDA step's kernel deformation (F per particle, self.jac,
        starting at I; aux targets each step)"""
        self.device = torch.device(device)
        ns = len(positions)
        assert ns == len(strengths) and ns > 0
        sizes = [int(np.asarray(p).shape[0]) for p in positions]
        self.offsets = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int32)
        self.h_w = np.broadcast_to(np.asarray(h_w, np.float32), (ns,)).copy()
        self.L = np.broadcast_to(np.asarray(L, np.float32), (ns,)).copy()
        pos = np.concatenate([np.asarray(p, np.float32).reshape(-1, 2) for p in positions])
        om = np.concatenate([np.asarray(w, np.float32).ravel() for w in strengths])
        self.pos = torch.as_tensor(pos, device=self.device).contiguous()
        self.omega = torch.as_tensor(om, device=self.device).contiguous()
        self.vel = torch.zeros_like(self.pos)
        self.n_sim = ns
        self.N = int(self.offsets[-1])
        d = EnsembleDesc()
        d.n_sim = ns
        d.h_offset = self.offsets.ctypes.data
        d.h_h_w = self.h_w.ctypes.data
        d.h_L = self.L.ctypes.data
        d.N1, d.N2 = int(N1), int(N2)
        d.mode = int(mode)
        d.cells_per_hw = int(cells_per_hw)
        d.spl = int(spl)
        d.eval_threads = int(eval_threads)
        d.seed = int(seed)
        d.dt = float(dt)
        d.pos_xy = ops._p(self.pos)
        d.omega = ops._p(self.omega)
        d.vel_xy = ops._p(self.vel)
        self.jac = None
        if deform:
            self.jac = torch.zeros((max(self.N, 1), 4), dtype=torch.float32, device=self.device)
            self.jac[:, 0] = 1.0
            self.jac[:, 3] = 1.0
            d.jac = ops._p(self.jac)
        self.desc = d
        need = lib().vpmc_ensemble_workspace_bytes(C.byref(d))
        if need == 0:
            raise ValueError("invalid ensemble description (sizes, h_w, grid)")
        self.ws = ops.workspace(need, self.device)
        self.step_index = 0
        self._graphs = {}
        self._stream = None
        self._order()
        check(lib().vpmc_ensemble_init(C.byref(d), ops._p(self.ws), self.ws.numel(),
                                       ops._stream()), "vpmc_ensemble_init")

    def _order(self):
        """orders the current stream after the ensemble's work on another one"""
        cur = torch.cuda.current_stream(self.device)
        if self._stream is not None and self._stream != cur:
            cur.wait_stream(self._stream)
            # the caching allocator must not hand these out while cur uses them
            for t in (self.ws, self.pos, self.omega, self.vel, self.jac):
                if t is not None:
                    t.record_stream(cur)
        self._stream = cur

    @property
    def workspace_bytes(self):
        return self.ws.numel()

    def set_step(self, step):
        self._order()
        check(lib().vpmc_ensemble_set_step(C.byref(self.desc), ops._p(self.ws), self.ws.numel(),
                                           int(step), ops._stream()), "vpmc_ensemble_set_step")
        self.step_index = int(step)

    def step(self, k=1, events=None):
        """k eager steps; events: optional (start, stop) torch.cuda.Event pair
        recorded by the library around the evaluation kernel of the last step"""
        self._order()
        for i in range(k):
            if events is not None and i == k - 1:
                self.desc.eval_events[0] = events[0].cuda_event
                self.desc.eval_events[1] = events[1].cuda_event
            check(lib().vpmc_ensemble_step(C.byref(self.desc), ops._p(self.ws), self.ws.numel(),
                                           ops._stream()), "vpmc_ensemble_step")
            self.step_index += 1
        self.desc.eval_events[0] = None
        self.desc.eval_events[1] = None

    def capture(self, steps):
        """a HIP graph of `steps` steps on the current (non-default) stream"""
        if steps in self._graphs:
            return self._graphs[steps]
        s = torch.cuda.current_stream()
        if s.cuda_stream == 0:
            raise RuntimeError("graph capture needs a non-default stream")
        h = C.c_void_p()
        check(lib().vpmc_ensemble_graph_create(C.byref(self.desc), ops._p(self.ws),
                                               self.ws.numel(), int(steps), C.c_void_p(s.cuda_stream),
                                               C.byref(h)), "vpmc_ensemble_graph_create")
        self._graphs[steps] = h
        return h

    def replay(self, steps):
        """advance by `steps` steps through the captured graph"""
        h = self.capture(steps)
        self._order()
        check(lib().vpmc_graph_launch(h, ops._stream()), "vpmc_graph_launch")
        self.step_index += steps

    def close(self):
        for h in self._graphs.values():
            lib().vpmc_graph_destroy(h)
        self._graphs = {}

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def positions(self, s=None):
        self._order()
        p = self.pos.cpu().numpy()
        if s is None:
            return p
        return p[self.offsets[s]:self.offsets[s + 1]]

    def jacobians(self, s=None):
        """F per particle (N, 4), or of simulation s"""
        self._order()
        F = self.jac[:self.N].cpu().numpy()
        return F if s is None else F[self.offsets[s]:self.offsets[s + 1]]

    def velocities(self, s=None):
        self._order()
        v = self.vel.cpu().numpy()
        if s is None:
            return v
        return v[self.offsets[s]:self.offsets[s + 1]]


# ---- autotuner ------------------------------------------------------------------

# (cells_per_hw, samples per lane, evaluation workgroup size)
# cells_per_hw sets the visited pairs per in-support pair, (pi + 4/c + 1/c^2)/pi
# (1.34 at 4, 1.27 at 5, 1.22 at 6, 1.17 at 8), against more, smaller tiles
# and, past 2^16 tiles per simulation, a third sort pass
TUNE_CANDIDATES = tuple((c, s, t) for c in (8, 6, 5, 4) for s in (1, 2) for t in (128, 256))


def _tune_cache_path():
    import os
    root = os.environ.get("VPMC_TUNE_CACHE",
                          os.path.join(os.path.expanduser("~"), ".cache", "vortexparticlemc_cuda_amd"))
    return os.path.join(root, "tune.json")


def autotune(positions, strengths, h_w, L=1.0, N1=64, N2=64, mode=MODE_CONSISTENT,
             candidates=TUNE_CANDIDATES, sample_sims=8, steps=2, use_cache=True, device="cuda",
             deform=False):
    """(cells_per_hw, spl, eval_threads) with the shortest step on a subset of the ensemble
    (`sample_sims` simulations spread over its sizes, the largest included:
    the largest simulation sets the sort's passes), timed with HIP events
    after one warm-up step; results are cached per GPU and workload shape
    (deform: the deformed step, whose search reaches 2 h_w)"""
    import json
    import os
    all_sizes = [int(np.asarray(p).shape[0]) for p in positions]
    order = np.argsort(all_sizes, kind="stable")
    pick = sorted(set(int(order[int(round(v))])
                      for v in np.linspace(0, len(order) - 1, min(sample_sims, len(order)))))
    positions = [positions[i] for i in pick]
    strengths = [strengths[i] for i in pick]
    hw_all = np.broadcast_to(np.asarray(h_w, np.float32), (len(all_sizes),))
    L_all = np.broadcast_to(np.asarray(L, np.float32), (len(all_sizes),))
    h_w, L = hw_all[pick], L_all[pick]
    sizes = tuple(all_sizes[i] for i in pick)
    key = "%s|sizes=%s|hw=%s|N1=%d|N2=%d|mode=%d" % (
        torch.cuda.get_device_name(device), sizes,
        tuple(round(float(v), 7) for v in h_w), N1, N2, mode) + ("|deform" if deform else "")
    path = _tune_cache_path()
    cache = {}
    if use_cache and os.path.exists(path):
        try:
            cache = json.load(open(path))
        except (OSError, ValueError):
            cache = {}
    if use_cache and key in cache:
        c = cache[key]
        return (int(c["cells_per_hw"]), int(c["spl"]), int(c["eval_threads"]),
                c.get("times_ms", {}))
    times = {}
    for cph, spl, th in candidates:
        try:
            e = Ensemble(positions, strengths, h_w, L=L, N1=N1, N2=N2,
                         mode=mode, cells_per_hw=cph, spl=spl, eval_threads=th, device=device,
                         deform=deform)
        except Exception:   # grid or workgroup invalid for this candidate
            continue
        e.step(1)
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record()
        e.step(steps)
        b.record()
        torch.cuda.synchronize()
        times["%d,%d,%d" % (cph, spl, th)] = a.elapsed_time(b) / steps
        e.close()
        del e
    if not times:
        raise RuntimeError("no valid tuning candidate")
    best = min(times, key=times.get)
    cph, spl, th = (int(v) for v in best.split(","))
    if use_cache:
        try:
            os.makedirs(os.path.dirname(path), exist_ok=True)
            cache[key] = {"cells_per_hw": cph, "spl": spl, "eval_threads": th, "times_ms": times}
            json.dump(cache, open(path, "w"), indent=1)
        except OSError:
            pass
    return cph, spl, th, times
