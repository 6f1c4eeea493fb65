"""Owner chunks and owner ranges of the estimator (vpmc_mc_desc.chunk_samples,
owner_base).

A sample's vorticity depends only on its point and the field, its draws on
the global owner index, and an owner's sums on its own samples in sample
order, so splitting the owners (into chunks inside one call, or into ranges
over several calls with owner_base) must not change a single bit.  The
reference step (aux targets + F, src/cuda/simulation_step.cu:51-73) at the
large run's size needs 3.2e9 sample slots, beyond one chunk: it runs in
chunks and its first owners are checked against the oracle."""
import ctypes as C

import numpy as np
import pytest

import oracle as O
from conftest import rel_l2, tg_problem

pytestmark = pytest.mark.gpu


def _say(msg):
    print("[case]", msg, flush=True)


def _sim(device, mode, deform=False, Nx=48, hwr=4, seed=9):
    from vortexparticlemc_cuda_amd.sim import SimParam, Simulation
    x, y, w, _ = tg_problem(Nx, hwr, jitter=0.3, seed=seed)
    P = SimParam(dt=1e-3, resX=Nx, hwr=hwr)
    sim = Simulation(P, device=device, cells_per_hw=8, mode=mode, seed=seed, deform=deform)
    sim.set_particles(np.stack([x, y], 1), w)
    if deform:   # a sheared, rotated F per particle (det 1)
        rng = np.random.default_rng(seed)
        a = 0.15 * rng.standard_normal(sim.N)
        th = 0.3 * rng.standard_normal(sim.N)
        c, s = np.cos(th), np.sin(th)
        F = np.stack([c + a * s, -s + a * c, s, c], 1)
        sim.set_jacobian(F / np.sqrt(np.abs(F[:, 0] * F[:, 3] - F[:, 1] * F[:, 2]))[:, None])
    return sim


@pytest.mark.parametrize("mode,aux,deform", [(1, False, False), (0, False, False),
                                             (1, True, True)])
def test_chunked_estimate_is_bit_identical(device, mode, aux, deform):
    """chunk_samples forcing many owner chunks == one chunk, bit for bit
    (centre and aux estimates), with a workspace bounded by the chunk"""
    from vortexparticlemc_cuda_amd._capi import lib
    _say("chunked estimate mode=%d aux=%d deform=%d" % (mode, aux, deform))
    N1, N2 = 24, 16
    sim = _sim(device, mode, deform)
    v1 = sim.estimate(N1, N2, aux=aux).clone()
    vs1 = sim.vel_sign.clone()
    va1 = sim.vel_aux.clone() if aux else None
    d1 = sim.desc(N1, N2, aux=aux)
    ws1 = lib().vpmc_mc_binned_workspace_bytes(C.byref(d1))
    per_owner = 2 * (N1 + (5 if aux else 1) * N2)
    for owners_per_chunk in (7, 100, 333):
        sim.chunk_samples = per_owner * owners_per_chunk
        sim.vel_sign.zero_()
        v2 = sim.estimate(N1, N2, aux=aux)
        assert np.array_equal(v2.cpu().numpy(), v1.cpu().numpy()), owners_per_chunk
        assert np.array_equal(sim.vel_sign.cpu().numpy(), vs1.cpu().numpy())
        if aux:
            assert np.array_equal(sim.vel_aux.cpu().numpy(), va1.cpu().numpy())
        d2 = sim.desc(N1, N2, aux=aux)
        ws2 = lib().vpmc_mc_binned_workspace_bytes(C.byref(d2))
        assert 0 < ws2 < ws1
    # one owner's samples must fit a chunk
    sim.chunk_samples = per_owner - 1
    with pytest.raises(Exception):
        sim.estimate(N1, N2, aux=aux)
    sim.chunk_samples = 0


@pytest.mark.parametrize("algo", [0, 1])
def test_owner_ranges_with_owner_base(device, algo):
    """owners [b, b + n) estimated on their own (owner_xy + b, owner_base = b)
    == rows b .. b+n-1 of the estimate over all owners, bit for bit, for the
    binned and the direct algorithm"""
    import torch

    from vortexparticlemc_cuda_amd import ops
    _say("owner ranges algo=%d" % algo)
    N1, N2 = 20, 20
    sim = _sim(device, 1)
    sim.algo = algo
    sim.estimate(N1, N2)
    N = sim.N
    full = sim.vel_sign[:2 * N].cpu().numpy().reshape(2, N, 2)
    cuts = [0, 37, 700, N]
    for b, e in zip(cuts[:-1], cuts[1:]):
        n = e - b
        d = sim.desc(N1, N2)
        out = torch.zeros((2 * n, 2), dtype=torch.float32, device=device)
        d.n_owner = n
        d.owner_xy = ops._p(sim.pos[b:e])
        d.owner_base = b
        d.vel_sign = ops._p(out)
        sim.launch_estimate(d)
        part = out.cpu().numpy().reshape(2, n, 2)
        assert np.array_equal(part, full[:, b:e]), (b, e)


def test_chunked_deformed_steps_are_bit_identical(device):
    """two steps of the CUDA step (aux targets + F) with the owners in chunks
    == the same steps in one chunk: positions, velocities and F"""
    _say("chunked deformed steps")
    N1, N2 = 16, 16
    out = []
    for chunk in (0, 2 * (N1 + 5 * N2) * 150):
        sim = _sim(device, 1, deform=True)
        sim.chunk_samples = chunk
        for _ in range(2):
            v = sim.step(N1, N2)
        out.append((sim.pos.cpu().numpy(), v.cpu().numpy(), sim.jac[:sim.N].cpu().numpy()))
    for a, b in zip(out[0], out[1]):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("threads,hwr", [(64, 4), (128, 4), (256, 4), (64, 16), (128, 16)])
def test_split_tail_waves_change_no_bit(device, threads, hwr):
    """A tile's last wave with at most 32 samples spreads each sample over 2
    or 4 lanes; the counting pass evaluates every sample on one lane.  The
    four-class summation order makes both give the same sums: equal
    velocities bit for bit, for every workgroup size, and the same again
    with the owners in chunks (other tiles' tails).  N1, N2 are chosen so
    that tiles end at many different fill levels of their last wave; at
    hwr = 16 a tile's stencil exceeds what 64 and 128 threads stage at once
    (4 records a thread), so the candidates come in several passes."""
    _say("split tail waves, %d threads, hwr %d" % (threads, hwr))
    N1, N2 = 23, 13
    sim = _sim(device, 1, Nx=56, hwr=hwr)
    sim.eval_threads = threads
    sim.spl = 1
    sim.rebuild_fields()
    sim.launch_estimate(sim.desc(N1, N2))
    v1 = sim.vel_sign.cpu().numpy().copy()
    sim.eval_count.zero_()
    sim.vel_sign.zero_()
    sim.launch_estimate(sim.desc(N1, N2, count=True))
    assert np.array_equal(sim.vel_sign.cpu().numpy(), v1)
    n_in, n_vis = (int(v) for v in sim.eval_count.cpu())
    assert 0 < n_in <= n_vis
    sim.chunk_samples = 2 * (N1 + N2) * 211
    sim.vel_sign.zero_()
    sim.launch_estimate(sim.desc(N1, N2))
    assert np.array_equal(sim.vel_sign.cpu().numpy(), v1)
    sim.chunk_samples = 0


def test_tile_chunks_change_no_bit(device):
    """A tile's samples are evaluated in chunks of 8 batches (512 samples
    with 64-thread tiles, 2048 with 256-thread ones), one workgroup each.  On
    a Gaussian-vortex member of the ensemble workload (hwr 8, N1 = N2 = 64)
    branch-1 samples crowd the vortex cores, so over all owners the heaviest
    tiles split into several chunks; with the owners in chunks of 397 a
    tile holds a few hundredth of those samples and rarely splits.  For each
    workgroup size the velocities must agree bit for bit (a sample's terms
    and their order do not depend on the workgroup evaluating it)."""
    from vortexparticlemc_cuda_amd.sim import SimParam, Simulation
    from vortexparticlemc_cuda_amd.workloads import gaussian_ensemble
    _say("tile chunks, Gaussian-vortex member")
    nx, pos, om, _ = gaussian_ensemble(4)[0]
    P = SimParam(dt=5e-4, resX=nx, hwr=8)
    sim = Simulation(P, device=device, cells_per_hw=5, mode=1, seed=3)
    h2 = (1.0 / nx) ** 2
    sim.set_particles(pos, (om * h2).astype(np.float32))
    sim.spl = 1
    for threads in (64, 256):
        out = []
        for chunk_samples in (0, 2 * 128 * 397):
            sim.eval_threads = threads
            sim.chunk_samples = chunk_samples
            sim.rebuild_fields()
            sim.vel_sign.zero_()
            sim.launch_estimate(sim.desc(64, 64))
            out.append(sim.vel_sign.cpu().numpy().copy())
        assert np.abs(out[0]).max() > 0
        assert np.array_equal(out[0], out[1]), threads
    sim.chunk_samples = 0


def test_reference_step_at_large_size_runs_in_chunks(device):
    """the CUDA step's semantics (aux targets + F) on the large run: double
    shear layer, 1,048,576 particles, hwr = 8, N1 = N2 = 256: 3.2e9 sample
    slots per step, more than one chunk (2^30).  The first owners' velocities
    (F = I before the first step) equal the oracle's on the same Philox
    stream; F comes out with det 1 and close to I; positions stay in the box"""
    import torch

    from vortexparticlemc_cuda_amd.sim import SimParam, Simulation, lattice_positions
    from vortexparticlemc_cuda_amd.workloads import double_shear
    _say("reference step at 1024^2, N1 = N2 = 256, aux + F (chunked)")
    Nx, hwr, N1 = 1024, 8, 256
    P = SimParam(dt=2.5e-4, resX=Nx, hwr=hwr)
    pos = lattice_positions(Nx)
    w = (double_shear(pos[:, 0], pos[:, 1]) / (Nx * Nx)).astype(np.float32)
    sim = Simulation(P, device=device, cells_per_hw=8, seed=0, spl=1, deform=True)
    sim.set_particles(pos, w)
    assert sim.N * 2 * (N1 + 5 * N1) > 2 ** 31
    v = sim.step(N1, N1)
    torch.cuda.synchronize()
    n_own = 16
    vs = sim.vel_sign[:2 * sim.N].view(2, sim.N, 2)
    vg = (vs[0, :n_own] - vs[1, :n_own]).cpu().numpy()
    assert np.array_equal(vg, v[:n_own].cpu().numpy())
    x = pos[:, 0].copy()
    y = pos[:, 1].copy()
    rsh = O.rsh_of(Nx, hwr)
    over = dict(cdf_pos=sim.sub["cdf_pos"][:sim.cnt[0]].cpu().numpy(),
                cdf_neg=sim.sub["cdf_neg"][:sim.cnt[1]].cpu().numpy(),
                cum_pos=sim.sub["totals"][0].item(), cum_neg=sim.sub["totals"][1].item())
    prob = O.Problem(x, y, w, rsh, rsh, cdf_override=over)
    vo = prob.velocities(prob.params(O.MODE_CONSISTENT, N1, N1, float(P.h_w), seed=0),
                         owners=n_own)
    err = rel_l2(vg, vo)
    _say("large reference step: first owners rel err %.3g" % err)
    assert err < 1e-5
    F = sim.jac[:sim.N].double()
    det = F[:, 0] * F[:, 3] - F[:, 1] * F[:, 2]
    assert torch.isfinite(F).all()
    assert float((det - 1).abs().max()) < 1e-5
    assert float((F - torch.tensor([1.0, 0.0, 0.0, 1.0], dtype=torch.float64,
                                   device=F.device)).abs().max()) < 0.5
    p1 = sim.pos.cpu().numpy()
    assert np.isfinite(p1).all() and (p1 >= 0).all() and (p1 < 1).all()
