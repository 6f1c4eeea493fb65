"""GPU tests of the reinitialisation the reference's README describes
(README.md:319-321, 416-420; no code in the reference): particles to a
uniform lattice, F to I, strengths re-fitted to the current field."""
import numpy as np
import pytest

import oracle as O
from conftest import rel_l2, tg_problem
from test_gpu_deform import _brute_deformed, _rand_F

pytestmark = pytest.mark.gpu


def test_reinit_of_the_lattice_state_returns_it(device):
    """TG strengths on the lattice (closed form w = b / lambda11): the field at
    the same lattice is A w, so the refit returns w and the lattice itself"""
    import torch

    from vortexparticlemc_cuda_amd import ops
    print("[case] reinit lattice identity", flush=True)
    Nx, hwr = 32, 4
    x, y = O.lattice(Nx)
    w = (O.tg_vorticity(x, y) / O.tg_lambda11(Nx, hwr)).astype(np.float32)
    pos = torch.as_tensor(np.stack([x, y], 1).astype(np.float32), device=device)
    p2, w2, it, rr, st = ops.reinit(pos, torch.as_tensor(w, device=device), hwr / Nx, Nx, Nx,
                                    tol=1e-10)
    assert st == 0 and rr < 1e-10, (st, rr)
    np.testing.assert_allclose(p2.cpu().numpy(), np.stack([x, y], 1), rtol=0, atol=1e-7)
    assert rel_l2(w2.cpu().numpy(), w) < 1e-6


@pytest.mark.parametrize("deformed", [False, True])
def test_reinit_fits_the_current_field(device, deformed):
    """jittered particles (and random det-1 F) onto a 40 x 40 lattice: the new
    particles' field at the lattice == the old field there (FP64 brute force
    of sum_j w_j W(|F_j^-1 d|)), to the solve's residual (the lattice Gram
    matrix is indefinite and MINRES slows near 1e-6 on such a field)"""
    import torch

    from vortexparticlemc_cuda_amd import ops
    print("[case] reinit fit deformed=%d" % deformed, flush=True)
    x, y, w, hw = tg_problem(32, 4, seed=31)
    F = _rand_F(x.size, np.random.default_rng(12), 1.5) if deformed else None
    pos = torch.as_tensor(np.stack([x, y], 1), device=device)
    n2 = 40
    p2, w2, it, rr, st = ops.reinit(pos, torch.as_tensor(w, device=device), hw, n2, n2,
                                    jac=None if F is None else torch.as_tensor(F, device=device),
                                    tol=1e-5, max_iter=3000)
    assert st == 0 and rr < 1e-5, (st, rr, it)
    q = p2.cpu().numpy()
    Fb = F if deformed else np.tile(np.array([1, 0, 0, 1], np.float32), (x.size, 1))
    b = _brute_deformed(q[:, 0], q[:, 1], x, y, w, Fb, hw, signed=True)
    # the refit field at the lattice, by the FP64 brute force of its own particles
    I = np.tile(np.array([1, 0, 0, 1], np.float32), (n2 * n2, 1))
    wn = w2.cpu().numpy()
    got = _brute_deformed(q[:, 0], q[:, 1], q[:, 0], q[:, 1], wn, I, hw, signed=True)
    assert rel_l2(got, b) < 3e-5, rel_l2(got, b)


def test_simulation_reinit_resets_F_and_steps(device):
    """Simulation.reinit with deformation: lattice positions, F = I, the next
    deformed step runs and refreshes F"""
    from vortexparticlemc_cuda_amd._capi import MODE_CONSISTENT
    from vortexparticlemc_cuda_amd.sim import SimParam, Simulation
    print("[case] simulation reinit", flush=True)
    x, y, w, hw = tg_problem(32, 4, seed=32)
    sim = Simulation(SimParam(dt=2e-3, resX=32, hwr=4), device=device, mode=MODE_CONSISTENT,
                     deform=True)
    sim.set_particles(np.stack([x, y], 1), w)
    for _ in range(3):
        sim.step(16, 16)
    assert not np.allclose(sim.jac[:sim.N].cpu().numpy(), [1, 0, 0, 1])
    it, rr, st = sim.reinit()
    assert st == 0 and rr < 1e-6
    assert sim.N == 32 * 32
    np.testing.assert_array_equal(sim.jac[:sim.N].cpu().numpy(),
                                  np.tile(np.array([1, 0, 0, 1], np.float32), (sim.N, 1)))
    lx, ly = O.lattice(32)
    np.testing.assert_allclose(sim.pos.cpu().numpy(), np.stack([lx, ly], 1), atol=1e-7)
    v = sim.step(16, 16).cpu().numpy()
    assert np.isfinite(v).all()
    assert not np.allclose(sim.jac[:sim.N].cpu().numpy(), [1, 0, 0, 1])


@pytest.mark.parametrize("Nx,Ny,hwr", [(32, 32, 4), (30, 24, 4), (48, 48, 8)])
def test_lattice_preconditioned_solve_matches_dense(device, Nx, Ny, hwr):
    """vpmc_cg_rbf_lattice on a smooth FP64 field with a sharp layer (double
    shear): FP64 residual 1e-11 in few iterations, strengths == the dense
    FP64 solve of the Gram matrix of the same FP32 positions"""
    import torch

    from vortexparticlemc_cuda_amd import ops
    from vortexparticlemc_cuda_amd.sim import lattice_positions
    from vortexparticlemc_cuda_amd.workloads import double_shear
    print("[case] lattice solve %dx%d hwr=%d" % (Nx, Ny, hwr), flush=True)
    ix, iy = np.meshgrid(np.arange(Nx), np.arange(Ny), indexing="ij")
    pos = np.stack([((ix + 0.5) / Nx).ravel(), ((iy + 0.5) / Ny).ravel()], 1).astype(np.float32)
    if Nx == Ny:
        assert np.array_equal(pos, lattice_positions(Nx))
    hw = hwr / Nx
    b = double_shear(pos[:, 0], pos[:, 1])
    pt = torch.as_tensor(pos, device=device)
    w, it, rr, st = ops.cg_rbf(pt, torch.as_tensor(b, device=device), hw, tol=1e-11,
                               max_iter=500, lattice=(Nx, Ny))
    assert st == 0 and rr < 1e-11, (st, rr, it)
    assert it < 100, it
    w0, it0, rr0, st0 = ops.cg_rbf(pt, torch.as_tensor(b, device=device), hw, tol=1e-11,
                                   max_iter=500)
    assert rr0 > 1e-11 or it0 > it   # the plain solve is slower or stalls
    p64 = pos.astype(np.float64)
    d = p64[:, None, :] - p64[None, :, :]
    d -= np.round(d)
    q = np.hypot(d[..., 0], d[..., 1]) / hw
    sig = 40.0 / (7 * np.pi * hw * hw)
    A = np.where(q <= 0.5, sig * (1 - 6 * q ** 2 + 6 * q ** 3),
                 np.where(q < 1, 2 * sig * (1 - q) ** 3, 0.0))
    wd = np.linalg.solve(A, b)
    # the strengths come back rounded to FP32 (the FP64 iterate's residual is
    # rr); on the 48 x 48, hwr = 8 lattice the shear layer is thinner than a
    # blob and the strengths oscillate at ~15 |b|, so the rounding alone
    # leaves |A w32 - b| / |b| ~ 4e-6
    wg = w.cpu().numpy().astype(np.float64)
    assert rel_l2(wg, wd) < 1e-5, rel_l2(wg, wd)
