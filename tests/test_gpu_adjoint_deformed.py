"""GPU tests of the deformed step's adjoint (vpmc_step_adjoint_deformed) and
of its FP64 forward (vpmc_step_f64_deformed).

The deformed step is step_cuda with kernel deformation: four auxiliary
targets per owner and F per particle (src/cuda/simulation_step.cu:51-73,
src/cuda/monte_carlo_bs.cu:166-169, 210-212, 274-288,
include/vorticity_query.cuh:52-66).  A particle's blob is W(|F^-1 d|) in
every query, its branch-1 jitter is F u, and the step sets
F' = renorm(I + dt/(2 dx) [u_R - u_L | u_T - u_B]) from the aux estimates.

The FP32 adjoint is checked against central finite differences of the FP64
forward of the same frozen-sample step, with respect to positions,
strengths and F; tolerances are relative to the largest gradient component
of the same kind.
"""
import numpy as np
import pytest

from conftest import tg_problem

pytestmark = pytest.mark.gpu


def _rand_F(n, rng, smax):
    """det-1 deformations R(a) diag(s, 1/s) R(b), sigma_max = s <= smax"""
    a, b = rng.uniform(0, 2 * np.pi, n), rng.uniform(0, 2 * np.pi, n)
    s = rng.uniform(1.0, smax, n)
    ca, sa, cb, sb = np.cos(a), np.sin(a), np.cos(b), np.sin(b)
    F = np.empty((n, 4))
    F[:, 0] = ca * s * cb - sa * sb / s
    F[:, 1] = -ca * s * sb - sa * cb / s
    F[:, 2] = sa * s * cb + ca * sb / s
    F[:, 3] = -sa * s * sb + ca * cb / s
    return F.astype(np.float32)


def _sim(device, Nx=32, hwr=4, N1=16, N2=16, dt=2e-3, seed=3, jitter=0.25):
    from vortexparticlemc_cuda_amd.sim import SimParam, Simulation
    x, y, w, hw = tg_problem(Nx, hwr, jitter=jitter, seed=seed)
    P = SimParam(dt=dt, resX=Nx, hwr=hwr)
    P.N1, P.N2 = N1, N2
    sim = Simulation(P, device=device, cells_per_hw=8, seed=seed, deform=True)
    pos = np.stack([x, y], 1).astype(np.float32)
    sim.set_particles(pos, w)
    return sim, pos, w


def _fd_pick(rng, k, w, min_frac=0.2):
    big = np.nonzero(np.abs(w) > min_frac * np.abs(w).max())[0]
    return rng.choice(big, size=k, replace=False)


def _unwrap(d):
    return d - np.round(d)


def test_f64_deformed_step_matches_fp32_step(device):
    """one FP64 deformed step against the FP32 deformed step (vpmc_step) from
    the same state on the same Philox stream: the move and F' agree to FP32
    rounding of the estimates"""
    import torch

    from vortexparticlemc_cuda_amd import adjoint as A
    sim, pos, w = _sim(device)
    rng = np.random.default_rng(5)
    F0 = _rand_F(len(w), rng, 1.3)
    sim.set_jacobian(F0)
    t = 2
    sim.step_index = t
    sim.step()
    x32 = sim.pos[:sim.N].double().cpu().numpy()
    F32 = sim.jac[:sim.N].double().cpu().numpy()
    X, F = A.step_f64(sim, torch.tensor(pos, device=device).double(),
                      torch.tensor(w, device=device).double(), t,
                      torch.tensor(F0, device=device).double())
    dt = float(sim.P.dt)
    u32 = _unwrap(x32 - pos) / dt
    u64 = _unwrap(X.cpu().numpy() - pos) / dt
    err_u = np.abs(u64 - u32).max() / np.abs(u32).max()
    dF32 = F32 - np.array([1, 0, 0, 1.0])
    dF64 = F.cpu().numpy() - np.array([1, 0, 0, 1.0])
    err_F = np.abs(dF64 - dF32).max() / np.abs(dF32).max()
    print("FP64 vs FP32 deformed step: move %.3g, F' - I %.3g" % (err_u, err_F))
    assert np.abs(dF32).max() > 1e-4
    assert err_u < 1e-3 and err_F < 1e-3


def test_deformed_step_vjp_matches_fp64_finite_differences(device):
    """L = sum c.(x' - x) + sum e.F' of one deformed step from a deformed
    state: dL/dx, dL/dw and dL/dF of the FP32 adjoint vs central differences
    of the FP64 forward.  The F part covers the quadratic forms in the
    queries, the jitter F u and the F update's aux estimates."""
    import torch

    from vortexparticlemc_cuda_amd import adjoint as A
    sim, pos, w = _sim(device)
    n = len(w)
    rng = np.random.default_rng(6)
    F0 = _rand_F(n, rng, 1.3)
    sim.set_jacobian(F0)
    t = 1
    c = rng.standard_normal((n, 2))
    e = rng.standard_normal((n, 4))
    ct = torch.tensor(c, device=device)
    et = torch.tensor(e, device=device)
    gx, gw, gF = A.step_vjp(sim, torch.tensor(pos, device=device),
                            torch.tensor(F0, device=device), t, ct.float(), et.float())
    gx = gx.double().cpu().numpy() - c   # the identity part of x' = x + dt u
    gw = gw.double().cpu().numpy()
    gF = gF.double().cpu().numpy()
    p64 = torch.tensor(pos, device=device).double()
    w64 = torch.tensor(w, device=device).double()
    F64 = torch.tensor(F0, device=device).double()

    def L(pp, ww, FF):
        X, F1 = A.step_f64(sim, pp, ww, t, FF)
        d = X - pp
        d = d - torch.round(d)
        return float((d * ct).sum() + (F1 * et).sum())

    errs = {"x": [], "w": [], "F": []}
    eps_x, eps_w, eps_F = 1e-6, 1e-6 * float(np.abs(w).max()), 1e-6
    for i in _fd_pick(rng, 6, w):
        k = int(rng.integers(2))
        a, b = p64.clone(), p64.clone()
        a[i, k] += eps_x
        b[i, k] -= eps_x
        fd = (L(a, w64, F64) - L(b, w64, F64)) / (2 * eps_x)
        errs["x"].append(abs(fd - gx[i, k]) / np.abs(gx).max())
        a, b = w64.clone(), w64.clone()
        a[i] += eps_w
        b[i] -= eps_w
        fd = (L(p64, a, F64) - L(p64, b, F64)) / (2 * eps_w)
        errs["w"].append(abs(fd - gw[i]) / np.abs(gw).max())
        k = int(rng.integers(4))
        a, b = F64.clone(), F64.clone()
        a[i, k] += eps_F
        b[i, k] -= eps_F
        fd = (L(p64, w64, a) - L(p64, w64, b)) / (2 * eps_F)
        errs["F"].append(abs(fd - gF[i, k]) / np.abs(gF).max())
    print("deformed step VJP vs FD: max rel err x %.3g w %.3g F %.3g"
          % (max(errs["x"]), max(errs["w"]), max(errs["F"])))
    assert np.abs(gx).max() > 0 and np.abs(gF).max() > 0
    for k in errs:
        assert max(errs[k]) < 1e-4, (k, errs[k])


def test_deformed_rollout_adjoint_matches_fp64_finite_differences(device):
    """a deformed rollout from F = I (F becomes part of the state after the
    first step), grid-vorticity loss: dL/dx0 and dL/dw along the FP64
    trajectory vs central differences of the FP64 rollout"""
    import torch

    from vortexparticlemc_cuda_amd import adjoint as A
    steps = 3
    sim, pos, w = _sim(device)
    h = float(sim.P.h_w)
    G = 16
    rng = np.random.default_rng(1)
    ref = A.grid_vorticity_f64(torch.tensor(pos, device=device).double(),
                               torch.tensor(w * (1 + 0.2 * rng.standard_normal(w.size)),
                                            device=device).double(), h, G, G)
    I = torch.tensor([1.0, 0.0, 0.0, 1.0], device=device).repeat(len(w), 1)
    sim.set_jacobian(I.cpu().numpy())
    p = torch.tensor(pos, device=device, requires_grad=True)
    wt = torch.tensor(w, device=device, requires_grad=True)
    XT = A.rollout(sim, p, wt, steps, checkpoint=2, trajectory="f64")
    FT = sim.jac[:sim.N].double()
    assert float((FT - I.double()).abs().max()) > 1e-4   # the rollout deformed
    om = A.grid_vorticity(XT, wt, h, G, G, f64=True)
    loss = ((om.double() - ref) ** 2).mean()
    loss.backward()
    gp = p.grad.double().cpu().numpy()
    gw = wt.grad.double().cpu().numpy()

    def L(pp, ww):
        X = A.rollout_f64(sim, pp, ww, steps, F0=I)
        o = A.grid_vorticity_f64(X, ww, h, G, G)
        return float(((o - ref) ** 2).mean())

    p64 = torch.tensor(pos, device=device).double()
    w64 = torch.tensor(w, device=device).double()
    L0 = L(p64, w64)
    assert abs(L0 - float(loss.detach())) < 1e-6 * abs(L0)
    eps_x, eps_w = 1e-6, 1e-6 * float(np.abs(w).max())
    errs_x, errs_w = [], []
    for i in _fd_pick(rng, 6, w):
        k = int(rng.integers(2))
        a, b = p64.clone(), p64.clone()
        a[i, k] += eps_x
        b[i, k] -= eps_x
        fd = (L(a, w64) - L(b, w64)) / (2 * eps_x)
        errs_x.append(abs(fd - gp[i, k]) / np.abs(gp).max())
        a, b = w64.clone(), w64.clone()
        a[i] += eps_w
        b[i] -= eps_w
        fd = (L(p64, a) - L(p64, b)) / (2 * eps_w)
        errs_w.append(abs(fd - gw[i]) / np.abs(gw).max())
    print("deformed rollout adjoint vs FD: max rel err positions %.3g strengths %.3g"
          % (max(errs_x), max(errs_w)))
    assert max(errs_x) < 1e-4 and max(errs_w) < 1e-4


def test_deformed_adjoint_is_deterministic(device):
    import torch

    from vortexparticlemc_cuda_amd import adjoint as A
    sim, pos, w = _sim(device)
    rng = np.random.default_rng(7)
    F0 = torch.tensor(_rand_F(len(w), rng, 1.2), device=device)
    lam = torch.tensor(rng.standard_normal((len(w), 2)).astype(np.float32), device=device)
    jb = torch.tensor(rng.standard_normal((len(w), 4)).astype(np.float32), device=device)
    out = [A.step_vjp(sim, torch.tensor(pos, device=device), F0, 4, lam, jb) for _ in range(2)]
    for a, b in zip(out[0], out[1]):
        assert np.array_equal(a.cpu().numpy(), b.cpu().numpy())
