"""The bindings module phash (the reference's bindings/phash.cpp:6-14, a
pybind11 module) on the device: its Simulation steps exactly like the Python
mirror (the same library entry points and Philox streams), and its rollout,
a torch.autograd.Function written in C++ over the hand-written step adjoint,
gives the same gradients as the Python autograd path."""
import numpy as np
import pytest

import oracle as O
from conftest import rel_l2

pytestmark = pytest.mark.gpu


def _say(msg):
    print("[case]", msg, flush=True)


def _problem(Nx=32, hwr=4):
    x, y = O.lattice(Nx)
    w = (O.tg_vorticity(x, y) / O.tg_lambda11(Nx, hwr)).astype(np.float32)
    return np.stack([x, y], 1).astype(np.float32), w


def test_phash_simulation_steps_like_the_mirror(device):
    import vortexparticlemc_cuda_amd as v
    from vortexparticlemc_cuda_amd._capi import MODE_CONSISTENT
    from vortexparticlemc_cuda_amd.sim import SimParam, Simulation
    _say("phash Simulation.step_cuda == Python mirror")
    phash = v.phash()
    assert phash.fill_dummy() == "phash imported OK"
    pos, w = _problem()
    P = phash.SimParam(dt=1e-3, resX=32, hwr=4)
    P.N1 = P.N2 = 32
    s = phash.Simulation(P)
    s.mode = phash.EstimatorMode.Consistent
    s.seed = 7
    s.set_particles(pos, w)
    Pm = SimParam(dt=1e-3, resX=32, hwr=4)
    m = Simulation(Pm, device=device, cells_per_hw=8, mode=MODE_CONSISTENT, seed=7)
    m.set_particles(pos, w)
    for _ in range(3):
        s.step_cuda(True)
        vm = m.step(32, 32)
    assert s.step_index == 3
    assert np.array_equal(s.positions(), m.pos.cpu().numpy())
    assert np.array_equal(s.velocities(), vm.cpu().numpy())
    # the reference's eval_vorticity_cpu on the views hashed at these positions
    s.do_spatial_hashing()
    p = s.positions()
    rsh = O.rsh_of(32, 4)
    prob = O.Problem(p[:, 0].copy(), p[:, 1].copy(), w, rsh, rsh)
    hw = float(P.h_w)
    for q in ((0.3, 0.7), (0.05, 0.95), (0.5, 0.5)):
        ref = prob.eval([q[0]], [q[1]], +1, hw)[0] - prob.eval([q[0]], [q[1]], -1, hw)[0]
        assert abs(s.eval_vorticity_cpu(q[0], q[1], True) - ref) <= 1e-5 * abs(ref) + 1e-6
    # copies are independent simulations
    import copy
    c = copy.copy(s)
    c.step_cuda(True)
    assert c.step_index == 4 and s.step_index == 3
    assert not np.array_equal(c.positions(), s.positions())


@pytest.mark.parametrize("traj,deform", [("f64", False), ("f32", False), ("f64", True),
                                         ("f32", True)])
def test_phash_rollout_gradient_matches_python_autograd(device, traj, deform):
    import torch

    import vortexparticlemc_cuda_amd as v
    from vortexparticlemc_cuda_amd import adjoint as A
    from vortexparticlemc_cuda_amd._capi import MODE_CONSISTENT
    from vortexparticlemc_cuda_amd.sim import SimParam, Simulation
    _say("phash.rollout backward == adjoint.rollout backward, %s trajectory, deform=%s"
         % (traj, deform))
    phash = v.phash()
    pos, w = _problem()
    T, K = 5, 2
    P = phash.SimParam(dt=2e-3, resX=32, hwr=4)
    P.N1 = P.N2 = 16
    s = phash.Simulation(P)
    s.seed = 3
    s.kernel_deformation = deform   # the state is (positions, F), F = I at the start
    s.set_particles(pos, w)
    Pm = SimParam(dt=2e-3, resX=32, hwr=4)
    Pm.N1 = Pm.N2 = 16
    m = Simulation(Pm, device=device, cells_per_hw=8, mode=MODE_CONSISTENT, seed=3,
                   deform=deform)
    m.set_particles(pos, w)
    rng = np.random.default_rng(1)
    target = torch.as_tensor(rng.standard_normal((pos.shape[0], 2)).astype(np.float32),
                             device=device)
    grads = []
    for run in ("phash", "python"):
        p = torch.as_tensor(pos, device=device).requires_grad_(True)
        ww = torch.as_tensor(w, device=device).requires_grad_(True)
        if run == "phash":
            X = phash.rollout(s, p, ww, T, K) if traj == "f64" else phash.rollout(s, p, ww, T, K, "f32")
        else:
            X = A.rollout(m, p, ww, T, K) if traj == "f64" else A.rollout(m, p, ww, T, K, "f32")
        # the default trajectory is FP64 and returns FP64 positions
        assert X.dtype == (torch.float64 if traj == "f64" else torch.float32)
        loss = (X * target).sum()
        loss.backward()
        grads.append((X.detach().cpu().numpy(), p.grad.cpu().numpy(), ww.grad.cpu().numpy()))
    (xa, gpa, gwa), (xb, gpb, gwb) = grads
    assert np.array_equal(xa, xb)
    assert rel_l2(gpa, gpb) < 1e-6 and rel_l2(gwa, gwb) < 1e-6
    assert np.abs(gwa).max() > 0
