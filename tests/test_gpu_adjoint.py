"""GPU tests of the adjoint (vpmc_mc_adjoint / vpmc_step_adjoint /
vpmc_grid_vorticity_adjoint behind vortexparticlemc_cuda_amd.adjoint).

The FP32 adjoint gradients are checked against central finite differences of
the FP64 forward of the same frozen-sample functions (vpmc_step_f64,
vpmc_grid_vorticity_f64), and for determinism.  Tolerances are relative to
the largest gradient component of the same kind (positions or strengths).
"""
import numpy as np
import pytest

from conftest import tg_problem

pytestmark = pytest.mark.gpu


def _sim(device, Nx=32, hwr=4, N1=16, N2=16, dt=2e-3, seed=3, jitter=0.25):
    from vortexparticlemc_cuda_amd.sim import SimParam, Simulation
    x, y, w, hw = tg_problem(Nx, hwr, jitter=jitter, seed=seed)
    P = SimParam(dt=dt, resX=Nx, hwr=hwr)
    P.N1, P.N2 = N1, N2
    sim = Simulation(P, device=device, cells_per_hw=8, seed=seed)
    pos = np.stack([x, y], 1).astype(np.float32)
    sim.set_particles(pos, w)
    return sim, pos, w


def _fd_pick(rng, n, k, w, min_frac=0.2):
    """k particle indices whose |w| is not small (no sign flip under eps)"""
    big = np.nonzero(np.abs(w) > min_frac * np.abs(w).max())[0]
    return rng.choice(big, size=k, replace=False)


def test_grid_vorticity_forward_and_adjoint(device):
    import torch

    from vortexparticlemc_cuda_amd import adjoint as A
    sim, pos, w = _sim(device)
    h = float(sim.P.h_w)
    G = 16
    p = torch.tensor(pos, device=device, requires_grad=True)
    wt = torch.tensor(w, device=device, requires_grad=True)
    om = A.grid_vorticity(p, wt, h, G, G)
    om64 = A.grid_vorticity_f64(p.detach().double(), wt.detach().double(), h, G, G)
    scale = float(om64.abs().max())
    assert float((om.detach().double() - om64).abs().max()) < 2e-5 * scale
    rng = np.random.default_rng(0)
    c = torch.tensor(rng.standard_normal(G * G), device=device)
    (om.double() * c).sum().backward()
    gp = p.grad.double().cpu().numpy()
    gw = wt.grad.double().cpu().numpy()

    def L(pp, ww):
        return float((A.grid_vorticity_f64(pp, ww, h, G, G) * c).sum())

    p64 = p.detach().double()
    w64 = wt.detach().double()
    eps_x, eps_w = 1e-6, 1e-6 * float(np.abs(w).max())
    for i in _fd_pick(rng, len(w), 6, w):
        for k in range(2):
            a = p64.clone()
            b = p64.clone()
            a[i, k] += eps_x
            b[i, k] -= eps_x
            fd = (L(a, w64) - L(b, w64)) / (2 * eps_x)
            assert abs(fd - gp[i, k]) < 1e-4 * np.abs(gp).max(), (i, k, fd, gp[i, k])
        a = w64.clone()
        b = w64.clone()
        a[i] += eps_w
        b[i] -= eps_w
        fd = (L(p64, a) - L(p64, b)) / (2 * eps_w)
        assert abs(fd - gw[i]) < 1e-4 * np.abs(gw).max(), (i, fd, gw[i])


def _loss_setup(device, steps):
    import torch

    from vortexparticlemc_cuda_amd import adjoint as A
    sim, pos, w = _sim(device)
    h = float(sim.P.h_w)
    G = 16
    rng = np.random.default_rng(1)
    ref = A.grid_vorticity_f64(torch.tensor(pos, device=device).double(),
                               torch.tensor(w * (1 + 0.2 * rng.standard_normal(w.size)),
                                            device=device).double(), h, G, G)
    return sim, pos, w, h, G, ref


@pytest.mark.parametrize("traj", ["f32", "f64"])
def test_rollout_adjoint_matches_fp64_finite_differences(device, traj):
    """dL/dx0, dL/dw of a checkpointed rollout vs central differences of the
    FP64 forward; along the FP64 trajectory the bound is 10x tighter"""
    import torch

    from vortexparticlemc_cuda_amd import adjoint as A
    steps = 4
    sim, pos, w, h, G, ref = _loss_setup(device, steps)
    p = torch.tensor(pos, device=device, requires_grad=True)
    wt = torch.tensor(w, device=device, requires_grad=True)
    XT = A.rollout(sim, p, wt, steps, checkpoint=3, trajectory=traj)
    om = A.grid_vorticity(XT, wt, h, G, G, f64=(traj == "f64"))
    loss = ((om.double() - ref) ** 2).mean()
    loss.backward()
    gp = p.grad.double().cpu().numpy()
    gw = wt.grad.double().cpu().numpy()

    # FP64 forward of the same frozen-sample rollout (sampling tables of w)
    def L(pp, ww):
        X = A.rollout_f64(sim, pp, ww, steps)
        o = A.grid_vorticity_f64(X, ww, h, G, G)
        return float(((o - ref) ** 2).mean())

    p64 = torch.tensor(pos, device=device).double()
    w64 = torch.tensor(w, device=device).double()
    L0 = L(p64, w64)
    assert abs(L0 - float(loss.detach())) < 1e-4 * abs(L0)
    rng = np.random.default_rng(2)
    eps_x, eps_w = 1e-6, 1e-6 * float(np.abs(w).max())
    errs_x, errs_w = [], []
    for i in _fd_pick(rng, len(w), 6, w):
        k = int(rng.integers(2))
        a = p64.clone()
        b = p64.clone()
        a[i, k] += eps_x
        b[i, k] -= eps_x
        fd = (L(a, w64) - L(b, w64)) / (2 * eps_x)
        errs_x.append(abs(fd - gp[i, k]) / np.abs(gp).max())
        a = w64.clone()
        b = w64.clone()
        a[i] += eps_w
        b[i] -= eps_w
        fd = (L(p64, a) - L(p64, b)) / (2 * eps_w)
        errs_w.append(abs(fd - gw[i]) / np.abs(gw).max())
    print("rollout adjoint (%s trajectory) vs FD: max rel err positions %.3g strengths %.3g"
          % (traj, max(errs_x), max(errs_w)))
    tol = 1e-4 if traj == "f32" else 1e-5
    assert max(errs_x) < tol and max(errs_w) < tol


@pytest.mark.parametrize("hwr,N1,N2", [(4, 16, 16), (6, 24, 8)])
def test_f64_step_matches_fp32_step(device, hwr, N1, N2):
    """one FP64 frozen-sample step (samples evaluated in cell order) against
    the FP32 step on the same Philox stream and the same sampling tables: the
    velocities agree to FP32 rounding, sample by sample the same estimate (a
    sample evaluated for another sample's owner or sign would move them by
    the velocity's own size); a second call on the same inputs repeats every
    bit"""
    import torch

    from vortexparticlemc_cuda_amd import adjoint as A
    sim, pos, w = _sim(device, Nx=32, hwr=hwr, N1=N1, N2=N2)
    dt = float(sim.P.dt)
    sim.step_index = 3
    u32 = sim.estimate(N1, N2).double().cpu().numpy()
    w64 = torch.tensor(w, device=device, dtype=torch.float64)
    out = []
    for _ in range(2):
        X = torch.tensor(pos, device=device, dtype=torch.float64)
        A._step_f64(sim, X, w64, 3, None)
        torch.cuda.synchronize()
        out.append(X.cpu().numpy())
    assert np.array_equal(out[0], out[1])
    d = out[0] - pos.astype(np.float64)
    d -= np.round(d)   # the wrap into [0, 1)
    u64 = d / dt
    err = np.abs(u64 - u32).max() / np.abs(u32).max()
    print("FP64 step vs FP32 step: max |du| / max |u| = %.3g" % err)
    assert np.abs(u32).max() > 0
    assert err < 1e-3


def test_adjoint_is_deterministic(device):
    import torch

    from vortexparticlemc_cuda_amd import adjoint as A
    sim, pos, w = _sim(device)
    rng = np.random.default_rng(4)
    c = torch.tensor(rng.standard_normal((len(w), 2)).astype(np.float32), device=device)
    out = []
    for _ in range(2):
        p = torch.tensor(pos, device=device, requires_grad=True)
        wt = torch.tensor(w, device=device, requires_grad=True)
        sim.step_index = 5
        u = A.mc_velocity(sim, p, wt)
        (u * c).sum().backward()
        out.append((p.grad.cpu().numpy(), wt.grad.cpu().numpy()))
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])
    assert np.abs(out[0][0]).max() > 0 and np.abs(out[0][1]).max() > 0


def _torch_blob_velocity(pos, w, h_w, kc, L=1.0):
    """FP64 velocity of the blob field at the particles, differentiable (torch
    autograd): u_i = sum_j w_j [K(r) M(|r|/h_w) + K_s(r)], r = minimg(x_i -
    x_j), M the spline's radial mass fraction, K_s the smooth part of the
    doubly periodic kernel from the lattice coefficients kc"""
    import torch
    d = pos[:, None, :] - pos[None, :, :]
    d = d - L * torch.round(d / L)
    rx, ry = d[..., 0], d[..., 1]
    n = pos.shape[0]
    eye = torch.eye(n, dtype=torch.bool, device=pos.device)
    r2 = torch.where(eye, torch.ones_like(rx), rx * rx + ry * ry)
    q = torch.sqrt(r2) / h_w
    q2 = q * q
    m_in = (80.0 / 7.0) * (0.5 * q2 - 1.5 * q2 * q2 + 1.2 * q2 * q2 * q)
    m_out = (80.0 / 7.0) * (q2 - 2.0 * q2 * q + 1.5 * q2 * q2 - 0.4 * q2 * q2 * q - 0.0125)
    M = torch.where(q <= 0.5, m_in, torch.where(q < 1.0, m_out, torch.ones_like(q)))
    f = torch.where(eye, torch.zeros_like(rx), M / (2 * np.pi * r2))
    s = torch.complex(rx, ry) / L
    s4 = s ** 4
    hsum = torch.zeros_like(s)
    for c in reversed(list(kc)):
        hsum = hsum * s4 + float(c)
    A = hsum * s ** 3 + np.pi * torch.conj(s)
    kx = -ry * f - A.imag / (2 * np.pi * L)
    ky = rx * f - A.real / (2 * np.pi * L)
    return torch.stack([(kx * w[None, :]).sum(1), (ky * w[None, :]).sum(1)], 1)


def test_adjoint_mean_matches_fp64_autograd_of_direct_sum(device):
    """E[MC adjoint] over Philox seeds = the gradient of the estimator's
    expectation, the blob direct sum, taken by FP64 torch autograd.  The
    strength gradient has finite variance (weights K/D are bounded); the
    position gradient's variance diverges logarithmically (the direction of K
    at branch-1 samples next to the owner), so it gets a looser bound."""
    import torch

    import oracle as O
    from vortexparticlemc_cuda_amd import adjoint as A
    sim, pos, w = _sim(device, Nx=24, hwr=4, N1=32, N2=32, jitter=0.25, seed=11)
    h = float(sim.P.h_w)
    n = len(w)
    # the torch restatement against the oracle's FP64 blob sum
    kc = O.kper_coeffs(24)
    p64 = torch.tensor(pos, device=device, dtype=torch.float64, requires_grad=True)
    w64 = torch.tensor(w, device=device, dtype=torch.float64, requires_grad=True)
    u = _torch_blob_velocity(p64, w64, h, kc)
    uo = O.direct_blob(pos[:, 0].astype(np.float64), pos[:, 1].astype(np.float64),
                       pos[:, 0].astype(np.float64), pos[:, 1].astype(np.float64),
                       w.astype(np.float64), h)
    assert np.abs(u.detach().cpu().numpy() - uo).max() < 1e-9 * np.abs(uo).max()
    rng = np.random.default_rng(7)
    c = torch.tensor(rng.standard_normal((n, 2)), device=device)
    (u * c).sum().backward()
    gx_ref = p64.grad.cpu().numpy()
    gw_ref = w64.grad.cpu().numpy()
    K = 96
    gx = np.zeros((K, n, 2))
    gw = np.zeros((K, n))
    c32 = c.float()
    for k in range(K):
        sim.seed = 1000 + k
        p = torch.tensor(pos, device=device, requires_grad=True)
        wt = torch.tensor(w, device=device, requires_grad=True)
        uu = A.mc_velocity(sim, p, wt)
        (uu * c32).sum().backward()
        gx[k] = p.grad.double().cpu().numpy()
        gw[k] = wt.grad.double().cpu().numpy()

    def ratio(samples, ref):
        err2 = ((samples.mean(0) - ref) ** 2).sum()
        var = samples.var(0, ddof=1).sum() / samples.shape[0]
        return float(err2 / var)

    rw, rx = ratio(gw, gw_ref), ratio(gx, gx_ref)
    print("MC adjoint mean vs FP64 autograd: err^2 / expected err^2 = %.3f (strengths), "
          "%.3f (positions)" % (rw, rx))
    assert rw < 2.0
    assert rx < 4.0
