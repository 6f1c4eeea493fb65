"""vpmc_step_noalloc (the step for callers that cannot allocate) against
vpmc_step, through the C ABI: the same positions, velocities (and F) bit for
bit, on two calls in a row (the second with a smaller input), one of them
large enough that the static scratch takes several owner chunks; sizes beyond
the scratch are rejected with VPMC_ERR_WORKSPACE."""
import ctypes as C

import numpy as np
import pytest

from conftest import tg_problem

pytestmark = pytest.mark.gpu


def _say(msg):
    print("[case]", msg, flush=True)


def _pair(device, Nx, hwr, deform):
    from vortexparticlemc_cuda_amd._capi import MODE_CONSISTENT
    from vortexparticlemc_cuda_amd.sim import SimParam, Simulation
    x, y, w, _ = tg_problem(Nx, hwr, jitter=0.3, seed=Nx)
    sims = []
    for _ in range(2):
        s = Simulation(SimParam(dt=1e-3, resX=Nx, hwr=hwr), device=device, cells_per_hw=8,
                       mode=MODE_CONSISTENT, seed=4, deform=deform)
        s.set_particles(np.stack([x, y], 1), w)
        sims.append(s)
    return sims


def _noalloc_step(sim, N1, N2):
    from vortexparticlemc_cuda_amd import ops
    from vortexparticlemc_cuda_amd._capi import lib
    d = sim.desc(N1, N2, aux=sim.deform)
    st = lib().vpmc_step_noalloc(C.byref(d), sim.N, ops._p(sim.pos), ops._p(sim.omega),
                                 ops._p(sim.sub["sub_pos"]), ops._p(sim.sub["sub_neg"]),
                                 ops._p(sim.vel), float(sim.P.dt), ops._stream())
    sim.step_index += 1
    return st


@pytest.mark.parametrize("deform", [False, True])
def test_step_noalloc_equals_step(device, deform):
    import torch
    # first call: 4096 owners x 2 signs x (512 + 512) samples = 8.4e6 sample
    # slots (about 480 MB of sample buffers: several chunks of the 256 MiB
    # scratch); second call smaller
    cases = ((64, 4, 512), (24, 4, 32)) if not deform else ((64, 4, 160), (24, 4, 16))
    for Nx, hwr, N1 in cases:
        _say("noalloc step deform=%d Nx=%d N1=N2=%d" % (deform, Nx, N1))
        a, b = _pair(device, Nx, hwr, deform)
        for _ in range(2):
            va = a.step(N1, N1)
            assert _noalloc_step(b, N1, N1) == 0
        torch.cuda.synchronize()
        assert np.array_equal(a.pos.cpu().numpy(), b.pos.cpu().numpy())
        assert np.array_equal(va.cpu().numpy(), b.vel[:b.N].cpu().numpy())
        assert np.abs(va.cpu().numpy()).max() > 0
        if deform:
            assert np.array_equal(a.jac[:a.N].cpu().numpy(), b.jac[:b.N].cpu().numpy())


def test_step_noalloc_rejects_sizes_beyond_its_scratch(device):
    # one owner's 2 x (N1 + N2) samples need more than the 256 MiB scratch
    _, b = _pair(device, 16, 4, False)
    p0 = b.pos.clone()
    assert _noalloc_step(b, 3_000_000, 3_000_000) == 3   # VPMC_ERR_WORKSPACE
    assert np.array_equal(b.pos.cpu().numpy(), p0.cpu().numpy())   # nothing ran
