"""The benchmark configurations at their stated sizes, against the oracle.

config 1  plain method: Taylor-Green, Nx = Ny = 100 (10,000 particles at the
          cell centres), hwr = 16, N1 = N2 = 64.  RBF strengths to 1e-8 against
          the lattice's closed form; the reference-mode estimator on injected
          uniforms (the draws monte_carlo_bs_cpu consumes,
          src/cpu/monte_carlo_bs.cpp:28-128) and the consistent estimator on
          Philox streams against the oracle for every owner; one step of
          each against the oracle's step; ten steps with the invariants.
config 3  the 64-member Gaussian ensemble (hwr = 8, N1 = N2 = 64): the first
          owners of members of three different Nx against the oracle on the
          member's own Philox stream (its sim index); whole-ensemble invariants.
config 4  the adjoint workload (Nx = 128, hwr = 16, T = 50, checkpoint every 10):
          the gradient along the FP64 trajectory against FP64 central
          differences (eps = 1e-4) on 8 coordinates.
grid      the B-spline vorticity on a 64 x 64 grid behind the adjoint's loss
          (vpmc_grid_vorticity, _f64 and the adjoint) against a numpy FP64
          brute-force sum over every particle and periodic image.
"""
import numpy as np
import pytest

import oracle as O
from conftest import rel_l2

pytestmark = pytest.mark.gpu


def _say(msg):
    print("[case]", msg, flush=True)


def _gpu_cdfs(sim):
    return dict(cdf_pos=sim.sub["cdf_pos"][:sim.cnt[0]].cpu().numpy(),
                cdf_neg=sim.sub["cdf_neg"][:sim.cnt[1]].cpu().numpy(),
                cum_pos=sim.sub["totals"][0].item(), cum_neg=sim.sub["totals"][1].item())


# --------------------------------------------------------------------------- config 1

def _config1(device, mode):
    import torch

    from vortexparticlemc_cuda_amd import ops
    from vortexparticlemc_cuda_amd.sim import SimParam, Simulation, lattice_positions
    Nx, hwr = 100, 16
    P = SimParam(dt=1e-3, resX=Nx, hwr=hwr)
    P.N1 = P.N2 = 64
    pos = lattice_positions(Nx)
    b = O.tg_vorticity(pos[:, 0].astype(np.float64), pos[:, 1].astype(np.float64))
    w, it, rr, st = ops.cg_rbf(torch.as_tensor(pos, device=device),
                               torch.as_tensor(b, device=device), float(P.h_w), tol=1e-8,
                               refine=True, lattice=(Nx, Nx))
    w = w.cpu().numpy()
    sim = Simulation(P, device=device, cells_per_hw=8, mode=mode, seed=0)
    sim.set_particles(pos, w)
    return Nx, hwr, P, pos, b, w, rr, st, sim


def test_config1_strengths_and_reference_replay(device):
    """RBF strengths to 1e-8 == b / lambda_11; the reference-mode estimator
    of every owner on injected uniforms == the oracle; one step == the
    oracle's step (x + u dt, wrapped)"""
    _say("config 1: strengths + reference-mode replay, 10,000 owners")
    Nx, hwr, P, pos, b, w, rr, st, sim = _config1(device, O.MODE_REFERENCE)
    assert st == 0 and rr < 1e-8
    # the FP64 Gram operator of the oracle (diagonal included,
    # src/cuda/cg_rbf_init.cu:11-31) at the FP32 positions: the FP32 strengths
    # leave a residual of their own rounding only (~3e-8 relative)
    x64, y64 = pos[:, 0].astype(np.float64), pos[:, 1].astype(np.float64)
    res = O.rbf_matvec(x64, y64, w.astype(np.float64), float(P.h_w)) - b
    assert np.linalg.norm(res) / np.linalg.norm(b) < 1e-7
    # the exact lattice's closed form w = b / lambda_11: the FP32 positions
    # perturb the circulant Gram matrix, and a solve to 1e-8 resolves that
    # perturbation in its near-null modes (a 1e-6 solve agrees to 2e-7,
    # test_cg_taylor_green_closed_form)
    w_closed = b / O.tg_lambda11(Nx, hwr)
    assert rel_l2(w, w_closed) < 1e-5
    N, N1, N2 = pos.shape[0], 64, 64
    rng = np.random.default_rng(11)
    U = rng.random(N * 2 * (N1 * 3 + N2 * 2), dtype=np.float32)
    vg = sim.estimate(N1, N2, uniforms=U).cpu().numpy()
    x = pos[:, 0].copy()
    y = pos[:, 1].copy()
    rsh = O.rsh_of(Nx, hwr)
    prob = O.Problem(x, y, w, rsh, rsh, cdf_override=_gpu_cdfs(sim))
    Pm = prob.params(O.MODE_REFERENCE, N1, N2, float(P.h_w), U=U)
    vo = prob.velocities(Pm)
    err = rel_l2(vg, vo)
    _say("config 1 reference replay rel err %.3g" % err)
    assert err < 1e-5 and np.abs(vo).max() > 0
    # one step on the same draws: the oracle's step_cpu restatement
    sim.step(N1, N2, uniforms=U)
    vel = np.zeros(2 * N, np.float32)
    O.lib().orc_step(O.C.byref(Pm), O.C.byref(prob.sets[0]), O.C.byref(prob.sets[1]), N, x, y,
                     float(P.dt), vel, 0)
    d = sim.pos.cpu().numpy() - np.stack([x, y], 1)
    d -= np.rint(d)
    assert np.abs(d).max() < 1e-6


def test_config1_consistent_steps(device):
    """consistent estimator on Philox streams (seed 0, step 0) == the oracle
    for every owner; one step == the oracle's step; ten steps keep the
    particles in the box with finite velocities and match their own eager
    velocities (x_{t+1} - x_t = u_t dt, wrapped)"""
    _say("config 1: consistent estimator + steps, 10,000 owners")
    Nx, hwr, P, pos, b, w, rr, st, sim = _config1(device, O.MODE_CONSISTENT)
    N, N1, N2 = pos.shape[0], 64, 64
    x = pos[:, 0].copy()
    y = pos[:, 1].copy()
    rsh = O.rsh_of(Nx, hwr)
    prob = O.Problem(x, y, w, rsh, rsh, cdf_override=_gpu_cdfs(sim))
    Pm = prob.params(O.MODE_CONSISTENT, N1, N2, float(P.h_w), seed=0, step=0)
    vg = sim.estimate(N1, N2).cpu().numpy()
    vo = prob.velocities(Pm)
    err = rel_l2(vg, vo)
    _say("config 1 consistent rel err %.3g" % err)
    assert err < 1e-5
    # the estimate converges to the blob field's periodic velocity; on the
    # Taylor-Green lattice that is c u_TG (c = 1.14395 for the min-image unit
    # cell): 64 + 64 samples leave a statistical error of order 0.3
    ut = O.tg_velocity(x, y)
    assert rel_l2(vo, 1.14395 * ut) < 0.6
    sim.step(N1, N2)
    vel = np.zeros(2 * N, np.float32)
    O.lib().orc_step(O.C.byref(Pm), O.C.byref(prob.sets[0]), O.C.byref(prob.sets[1]), N, x, y,
                     float(P.dt), vel, 0)
    d = sim.pos.cpu().numpy() - np.stack([x, y], 1)
    d -= np.rint(d)
    assert np.abs(d).max() < 1e-6
    for _ in range(9):   # steps 1 .. 9
        p0 = sim.pos.cpu().numpy().astype(np.float64)
        v = sim.step(N1, N2).cpu().numpy().astype(np.float64)
        p1 = sim.pos.cpu().numpy().astype(np.float64)
        assert np.isfinite(v).all() and (p1 >= 0).all() and (p1 < 1).all()
        dd = p1 - (p0 + v * float(P.dt))
        dd -= np.rint(dd)
        assert np.abs(dd).max() < 1e-6
    assert sim.step_index == 10


# --------------------------------------------------------------------------- config 3

def test_config3_ensemble_members_match_oracle(device):
    """64 Gaussian-vortex simulations (Nx from {64,96,128,192,256}, hwr = 8,
    N1 = N2 = 64, dt = 5e-4) in one ensemble: after one step the first owners
    of the first member of three different sizes equal the oracle on the
    member's Philox stream (seed 0, step 0, sim = member index); every member's
    particles moved by their velocity times dt and stayed in the box"""
    import torch

    from vortexparticlemc_cuda_amd import ops
    from vortexparticlemc_cuda_amd.ensemble import Ensemble
    from vortexparticlemc_cuda_amd.sim import SimParam
    from vortexparticlemc_cuda_amd.workloads import gaussian_ensemble
    _say("config 3: 64-member ensemble")
    data = gaussian_ensemble(64)
    hwr, N1, dt = 8, 64, 5e-4
    pos, w, hw, nxs = [], [], [], []
    for nx, p, om, _ in data:
        P = SimParam(dt=dt, resX=nx, hwr=hwr)
        wt, it, rr, st = ops.cg_rbf(torch.as_tensor(p, device=device),
                                    torch.as_tensor(om.astype(np.float32), device=device),
                                    float(P.h_w), tol=5e-3, max_iter=300, refine=True)
        pos.append(p)
        w.append(wt.cpu().numpy())
        hw.append(float(P.h_w))
        nxs.append(nx)
    assert len(set(nxs)) >= 3
    ens = Ensemble(pos, w, hw, N1=N1, N2=N1, dt=dt, seed=0, cells_per_hw=4, spl=1,
                   eval_threads=256, device=device)
    assert ens.N == sum(p.shape[0] for p in pos) and ens.n_sim == 64
    ens.step(1)
    torch.cuda.synchronize()
    picked = []
    for s, nx in enumerate(nxs):
        if nx not in [nxs[q] for q in picked]:
            picked.append(s)
    picked = picked[:3]
    n_own = 48
    for s in picked:
        nx = nxs[s]
        x = pos[s][:, 0].copy()
        y = pos[s][:, 1].copy()
        rsh = O.rsh_of(nx, hwr)
        prob = O.Problem(x, y, w[s], rsh, rsh)
        Pm = prob.params(O.MODE_CONSISTENT, N1, N1, hw[s], seed=0, step=0, sim=s)
        vo = prob.velocities(Pm, owners=n_own)
        ve = ens.velocities(s)[:n_own]
        err = rel_l2(ve, vo)
        _say("config 3 member %d (Nx=%d) rel err %.3g" % (s, nx, err))
        assert err < 1e-5 and np.abs(vo).max() > 0
    for s in range(64):   # whole-ensemble invariants
        p1 = ens.positions(s).astype(np.float64)
        v = ens.velocities(s).astype(np.float64)
        assert np.isfinite(v).all() and (p1 >= 0).all() and (p1 < 1).all()
        d = p1 - (pos[s].astype(np.float64) + v * dt)
        d -= np.rint(d)
        assert np.abs(d).max() < 1e-6, s
    ens.close()


# --------------------------------------------------------------------------- config 4

def test_config4_adjoint_fp64_trajectory_vs_central_differences(device):
    """Nx = 128 Taylor-Green + 0.1 N(0,1) strength noise, hwr = 16,
    N1 = N2 = 64, T = 50, checkpoint 10, loss = MSE of the final vorticity on
    a 64 x 64 grid against the unperturbed run: dL/dx0 and dL/dw along the
    FP64 trajectory == FP64 central differences (eps = 1e-4) on 8 coordinates
    to 1e-4 of the largest gradient entry"""
    import torch

    from vortexparticlemc_cuda_amd import adjoint as A
    from vortexparticlemc_cuda_amd import ops
    from vortexparticlemc_cuda_amd.sim import SimParam, Simulation, lattice_positions
    from vortexparticlemc_cuda_amd.workloads import taylor_green
    _say("config 4: adjoint, T=50, FD on 8 coordinates")
    Nx, hwr, N, T, K, G = 128, 16, 64, 50, 10, 64
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
        ref = A.grid_vorticity(A.rollout(sim, pos_t, w_tg, T, checkpoint=K, trajectory="f32"),
                               w_tg, h, G, G, f64=True)
    p = pos_t.clone().requires_grad_(True)
    wt = w0.clone().requires_grad_(True)
    XT = A.rollout(sim, p, wt, T, checkpoint=K, trajectory="f64")
    loss = ((A.grid_vorticity(XT, wt, h, G, G, f64=True) - ref) ** 2).mean()
    loss.backward()
    gp = p.grad.double().cpu().numpy()
    gw = wt.grad.double().cpu().numpy()
    p64 = pos_t.double()
    w64 = w0.double()

    def L(pp, ww):
        X = A.rollout_f64(sim, pp, ww, T)
        return float(((A.grid_vorticity_f64(X, ww, h, G, G) - ref) ** 2).mean())

    L0 = L(p64, w64)
    assert abs(L0 - float(loss.detach())) < 1e-6 * abs(L0)
    wn = w0.abs().cpu().numpy()
    pick = rng.choice(np.nonzero(wn > 0.2 * wn.max())[0], size=8, replace=False)
    eps_w = 1e-4 * float(w0.abs().max())

    def dpos(i, k, e):
        a, b = p64.clone(), p64.clone()
        a[i, k] += e
        b[i, k] -= e
        return (L(a, w64) - L(b, w64)) / (2 * e)

    # Positions: a branch-1 sample at distance r from another owner puts a
    # 1/r Biot-Savart term in the frozen-sample velocity, whose central
    # difference at eps carries a truncation error ~ (eps / r)^2 of itself;
    # one Richardson step from eps = 1e-4 and 5e-5 cancels that term, and
    # eps = 1e-6 is compared as well.
    ex_r, ex_6, ex_raw, ew = [], [], [], []
    for c, i in enumerate(pick):
        if c % 2 == 0:
            k = c // 2 % 2
            d4, d45 = dpos(i, k, 1e-4), dpos(i, k, 5e-5)
            rich = (4.0 * d45 - d4) / 3.0
            gm = np.abs(gp).max()
            ex_raw.append(abs(d4 - gp[i, k]) / gm)
            ex_r.append(abs(rich - gp[i, k]) / gm)
            ex_6.append(abs(dpos(i, k, 1e-6) - gp[i, k]) / gm)
        else:
            a, b = w64.clone(), w64.clone()
            a[i] += eps_w
            b[i] -= eps_w
            fd = (L(p64, a) - L(p64, b)) / (2 * eps_w)
            ew.append(abs(fd - gw[i]) / np.abs(gw).max())
    _say("config 4: max |g - fd| / max|g|: positions eps=1e-4 %.3g, Richardson(1e-4) %.3g, "
         "eps=1e-6 %.3g; strengths eps=1e-4 %.3g" % (max(ex_raw), max(ex_r), max(ex_6), max(ew)))
    assert max(ex_r) < 1e-4 and max(ex_6) < 1e-5 and max(ew) < 1e-4


# --------------------------------------------------------------------------- grid field

def _spline_np(r, h):
    """include/utilities.hpp:7-17 in FP64, and its derivative dW/dr"""
    sig2 = 40.0 / (7.0 * np.pi * h * h)
    q = r / h
    W = np.where(q <= 0.5, 6.0 * (q ** 3 - q ** 2) + 1.0,
                 np.where(q < 1.0, 2.0 * (1.0 - q) ** 3, 0.0)) * sig2
    dW = np.where(q <= 0.5, 18.0 * q ** 2 - 12.0 * q,
                  np.where(q < 1.0, -6.0 * (1.0 - q) ** 2, 0.0)) * sig2 / h
    return W, dW


def test_grid_vorticity_vs_numpy_brute_force(device):
    """omega_h(z_ab) = sum_j w_j W(|z_ab - x_j|) over every particle by
    minimum image (h_w < L/2), z_ab = ((a + 1/2)/64, (b + 1/2)/64): the FP64
    kernel to 1e-12, the FP32 fine-grid kernel to 1e-5, and the adjoint (dL/dw_j
    = sum_g g_g W, dL/dx_j = sum_g g_g w_j W'(r) (x_j - z_g)/r) to 1e-5, against
    numpy in FP64 on the adjoint workload's size (16,384 jittered particles)"""
    import torch

    from vortexparticlemc_cuda_amd import adjoint as A
    _say("grid field vs numpy FP64 brute force")
    Nx, hwr, G = 128, 16, 64
    h_w = np.float32(hwr / Nx)
    rng = np.random.default_rng(5)
    x, y = O.lattice(Nx)
    x = ((x + 0.3 / Nx * rng.standard_normal(x.size)) % 1.0).astype(np.float32)
    y = ((y + 0.3 / Nx * rng.standard_normal(y.size)) % 1.0).astype(np.float32)
    w = (O.tg_vorticity(x, y) / Nx ** 2 + 0.05 * rng.standard_normal(x.size) / Nx ** 2)
    w = w.astype(np.float32)
    go = rng.standard_normal(G * G)
    ga, gb = np.meshgrid((np.arange(G) + 0.5) / G, (np.arange(G) + 0.5) / G, indexing="ij")
    zx, zy = ga.ravel(), gb.ravel()
    X, Y, Wt = x.astype(np.float64), y.astype(np.float64), w.astype(np.float64)
    om = np.zeros(G * G)
    gw = np.zeros(x.size)
    gpx = np.zeros(x.size)
    gpy = np.zeros(x.size)
    for a in range(0, G * G, 256):
        dx = X[None, :] - zx[a:a + 256, None]
        dy = Y[None, :] - zy[a:a + 256, None]
        dx -= np.rint(dx)
        dy -= np.rint(dy)
        r = np.sqrt(dx * dx + dy * dy)
        Wv, dW = _spline_np(r, float(h_w))
        om[a:a + 256] = Wv @ Wt
        gw += go[a:a + 256] @ Wv
        f = np.where(r > 0, dW / np.where(r > 0, r, 1.0), 0.0) * go[a:a + 256, None]
        gpx += (f * dx).sum(0) * Wt
        gpy += (f * dy).sum(0) * Wt
    pos = torch.as_tensor(np.stack([x, y], 1), device=device)
    wt = torch.as_tensor(w, device=device)
    o64 = A.grid_vorticity_f64(pos.double(), wt.double(), float(h_w), G, G).cpu().numpy()
    assert rel_l2(o64, om) < 1e-12
    p = pos.clone().requires_grad_(True)
    ww = wt.clone().requires_grad_(True)
    o32 = A.grid_vorticity(p, ww, float(h_w), G, G)
    assert rel_l2(o32.detach().cpu().numpy(), om) < 1e-5
    (o32 * torch.as_tensor(go, dtype=torch.float32, device=device)).sum().backward()
    assert rel_l2(ww.grad.cpu().numpy(), gw) < 1e-5
    assert rel_l2(p.grad.cpu().numpy(), np.stack([gpx, gpy], 1)) < 1e-5
