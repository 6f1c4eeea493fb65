"""Ensemble of simulations: one set of launches per step for all of them.

Each simulation must evolve exactly as the plain single-simulation path does
(same seed, sim index, cells_per_hw, spl): positions and velocities are
compared bit for bit after several steps.  The captured HIP graph must
reproduce eager steps bit for bit on every replay."""
import numpy as np
import pytest
import torch

from conftest import rel_l2


def _say(msg):
    print("[case] " + msg, flush=True)


def _ensemble_data(sizes=(16, 24, 32), hwr=4, seed=5):
    from vortexparticlemc_cuda_amd.sim import SimParam
    from vortexparticlemc_cuda_amd.workloads import gaussian_vortices
    rng = np.random.default_rng(seed)
    out = []
    for nx in sizes:
        P = SimParam(dt=1e-3, resX=nx, hwr=hwr)
        h = 1.0 / nx
        i, j = np.meshgrid(np.arange(nx), np.arange(nx), indexing="ij")
        x = ((i + 0.5) * h + 0.2 * h * rng.standard_normal((nx, nx))).ravel() % 1.0
        y = ((j + 0.5) * h + 0.2 * h * rng.standard_normal((nx, nx))).ravel() % 1.0
        om = gaussian_vortices(x, y, rng.random((4, 2)), rng.normal(0, 4 * np.pi, 4),
                               rng.uniform(0.05, 0.12, 4))
        pos = np.stack([x, y], 1).astype(np.float32)
        out.append((P, pos, (om * h * h).astype(np.float32)))
    return out


@pytest.mark.gpu
@pytest.mark.parametrize("spl,threads", [(1, 256), (2, 256), (4, 256), (2, 64), (1, 128)])
def test_ensemble_equals_plain_path(device, spl, threads):
    from vortexparticlemc_cuda_amd.ensemble import Ensemble
    from vortexparticlemc_cuda_amd.sim import Simulation
    _say("ensemble == plain, spl=%d threads=%d" % (spl, threads))
    N1, N2, steps, seed = 32, 32, 3, 77
    data = _ensemble_data()
    ens = Ensemble([d[1] for d in data], [d[2] for d in data], [d[0].h_w for d in data],
                   N1=N1, N2=N2, dt=float(data[0][0].dt), seed=seed, cells_per_hw=8, spl=spl,
                   eval_threads=threads, device=device)
    ens.step(steps)
    torch.cuda.synchronize()
    for s, (P, pos, w) in enumerate(data):
        P.N1, P.N2 = N1, N2
        sim = Simulation(P, device=device, cells_per_hw=8, seed=seed, sim=s, spl=2)
        sim.set_particles(pos, w)
        for _ in range(steps):
            v = sim.step()
        torch.cuda.synchronize()
        pe, pp = ens.positions(s), sim.pos.cpu().numpy()
        ve, vp = ens.velocities(s), v.cpu().numpy()
        assert np.array_equal(pe, pp), (s, rel_l2(pe, pp))
        assert np.array_equal(ve, vp), (s, rel_l2(ve, vp))
        assert np.abs(vp).max() > 0


@pytest.mark.gpu
@pytest.mark.parametrize("deform", [False, True])
def test_ensemble_graph_replays_match_eager(device, deform):
    """replays of a captured 2-step graph equal eager steps bit for bit (with
    kernel deformation: F, updated inside the graph, as well)"""
    from vortexparticlemc_cuda_amd.ensemble import Ensemble
    _say("ensemble graph replay, deform=%d" % deform)
    data = _ensemble_data(sizes=(20, 28), seed=9)   # 2 h_w < L / 2 for the deformed reach
    args = ([d[1] for d in data], [d[2] for d in data], [d[0].h_w for d in data])
    kw = dict(N1=16, N2=16, dt=1e-3, seed=3, device=device, deform=deform)
    eager = Ensemble(*args, **kw)
    ref, refF = [], []
    for _ in range(3):
        eager.step(2)
        ref.append(eager.positions().copy())
        if deform:
            refF.append(eager.jacobians().copy())
    side = torch.cuda.Stream()
    with torch.cuda.stream(side):
        g = Ensemble(*args, **kw)
        for k in range(3):
            g.replay(2)
            torch.cuda.synchronize()
            got = g.positions()
            assert np.array_equal(got, ref[k]), (k, rel_l2(got, ref[k]))
            if deform:
                assert np.array_equal(g.jacobians(), refF[k]), k
            # eager work of the program between replays
            scratch = torch.randn(4096, device=device).sum()
            torch.cuda.synchronize()
            assert torch.isfinite(scratch)
        g.close()
    assert g.step_index == 6


@pytest.mark.gpu
def test_ensemble_orders_work_across_streams(device):
    """eager steps queued on the default stream, then a graph replayed on a
    side stream and positions read back on the default stream, with no
    synchronisation by the caller: the result equals four eager steps"""
    from vortexparticlemc_cuda_amd.ensemble import Ensemble
    _say("ensemble work ordered across streams")
    data = _ensemble_data(sizes=(20, 28), seed=9)
    args = ([d[1] for d in data], [d[2] for d in data], [d[0].h_w for d in data])
    kw = dict(N1=16, N2=16, dt=1e-3, seed=3, device=device)
    ref = Ensemble(*args, **kw)
    ref.step(4)
    want = ref.positions()
    e = Ensemble(*args, **kw)
    e.step(2)
    side = torch.cuda.Stream()
    with torch.cuda.stream(side):
        e.capture(2)
        e.replay(2)
    got = e.positions()
    e.close()
    assert np.array_equal(got, want), rel_l2(got, want)


@pytest.mark.gpu
@pytest.mark.parametrize("spl,threads", [(2, 256), (1, 128)])
def test_deformed_ensemble_equals_plain_path(device, spl, threads):
    """with kernel deformation (aux targets, F in fields and jitter, F from
    the five estimates each step) every simulation of the ensemble follows
    the single-simulation deformed step bit for bit: positions, velocities, F"""
    from vortexparticlemc_cuda_amd.ensemble import Ensemble
    from vortexparticlemc_cuda_amd.sim import Simulation
    _say("deformed ensemble == plain, spl=%d threads=%d" % (spl, threads))
    N1, N2, steps, seed = 24, 24, 3, 79
    data = _ensemble_data(sizes=(24, 32, 40))   # 2 h_w (the deformed reach) < L / 2
    ens = Ensemble([d[1] for d in data], [d[2] for d in data], [d[0].h_w for d in data],
                   N1=N1, N2=N2, dt=float(data[0][0].dt), seed=seed, cells_per_hw=8, spl=spl,
                   eval_threads=threads, device=device, deform=True)
    ens.step(steps)
    torch.cuda.synchronize()
    for s, (P, pos, w) in enumerate(data):
        P.N1, P.N2 = N1, N2
        sim = Simulation(P, device=device, cells_per_hw=8, seed=seed, sim=s, spl=spl,
                         eval_threads=threads, deform=True)
        sim.set_particles(pos, w)
        for _ in range(steps):
            v = sim.step()
        torch.cuda.synchronize()
        assert np.array_equal(ens.positions(s), sim.pos.cpu().numpy()), s
        assert np.array_equal(ens.velocities(s), v.cpu().numpy()), s
        F = ens.jacobians(s)
        assert np.array_equal(F, sim.jac[:sim.N].cpu().numpy()), s
        assert not np.allclose(F, [1, 0, 0, 1])
