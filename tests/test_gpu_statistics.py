"""Statistical checks of the estimator on the GPU:
- the consistent estimator's error against the FP64 direct blob sum falls
  as 1/sqrt(N1 + N2);
- in reference mode, the per-owner mean and variance of the GPU estimator
  (Philox draws) agree with the reference's own monte_carlo_bs_cpu
  (src/cpu/monte_carlo_bs.cpp:28-128, compiled from its sources into
  oracle/_ref, std::mt19937 draws) within their 95 % intervals;
- sampled particle indices follow |omega_j| / Omega (chi-square).
"""
import numpy as np
import pytest

import oracle as O
from conftest import tg_problem

pytestmark = pytest.mark.gpu


def _say(msg):
    print("[case]", msg, flush=True)


def _sim(device, x, y, w, Nx, hwr, mode, seed=0):
    from vortexparticlemc_cuda_amd.sim import SimParam, Simulation
    P = SimParam(dt=1e-3, resX=Nx, hwr=hwr)
    sim = Simulation(P, device=device, cells_per_hw=8, mode=mode, seed=seed)
    sim.set_particles(np.stack([x, y], 1), w)
    return sim


def test_consistent_error_falls_as_inverse_sqrt_samples(device):
    """RMS |u_MC - u_direct| at N1 = N2 = 16, 64, 256 (averaged over 8 seeds):
    each 4x in samples halves the error (ratio 2 +- 0.25)"""
    import torch

    from vortexparticlemc_cuda_amd import ops
    from vortexparticlemc_cuda_amd._capi import MODE_CONSISTENT
    _say("convergence rate")
    x, y, w, hw = tg_problem(32, 4, seed=41)
    sim = _sim(device, x, y, w, 32, 4, MODE_CONSISTENT)
    pos = torch.as_tensor(np.stack([x, y], 1).astype(np.float64), device=device)
    ud = ops.direct_sum_blob(pos, pos, torch.as_tensor(w.astype(np.float64), device=device),
                             hw).cpu().numpy()
    err = []
    for n in (16, 64, 256):
        e2 = 0.0
        for k in range(8):
            sim.seed = 100 + k
            u = sim.estimate(n, n).double().cpu().numpy()
            e2 += ((u - ud) ** 2).sum(1).mean()
        err.append(np.sqrt(e2 / 8))
    r1, r2 = err[0] / err[1], err[1] / err[2]
    print("rms error at N1=N2=16/64/256:", err, "ratios", r1, r2, flush=True)
    assert 1.75 < r1 < 2.25 and 1.75 < r2 < 2.25, (err, r1, r2)


def test_reference_mode_mean_and_variance_match_the_reference_cpu(device):
    """per-owner mean and variance over 24 independent runs each: the GPU's
    reference mode (Philox) vs the reference's monte_carlo_bs_cpu (mt19937)
    on the same particles: the 95 % bootstrap interval of the variance ratio
    holds 1, and the per-owner mean differences are standard-error sized"""
    from vortexparticlemc_cuda_amd._capi import MODE_REFERENCE
    R = O.ref_lib()
    if R is None:
        pytest.fail("oracle/_ref/libvpmc_ref.so not built (run build())")
    _say("variance vs reference cpu")
    Nx, hwr, N1, N2, n_own, K = 24, 4, 32, 32, 160, 24
    x, y, w, hw = tg_problem(Nx, hwr, seed=42)
    rsh = O.rsh_of(Nx, hwr)
    sim = _sim(device, x, y, w, Nx, hwr, MODE_REFERENCE)
    g = []
    for k in range(K):
        sim.seed = 1000 + k
        g.append(sim.estimate(N1, N2).double().cpu().numpy()[:n_own])
    g = np.array(g)                                           # (K, n_own, 2)
    c = []
    for k in range(K):
        v = np.zeros(2 * n_own, np.float32)
        R.ref_mc_velocities(x.size, x, y, w, rsh, rsh, 1.0, 1.0, np.float32(hw), 1, N1, N2, 0,
                            7919 * (k + 1), 0, n_own, v)
        c.append(v.reshape(-1, 2).astype(np.float64))
    c = np.array(c)
    vg, vc = g.var(0, ddof=1), c.var(0, ddof=1)              # (n_own, 2)
    ratio = vg.sum() / vc.sum()
    # 95 % bootstrap interval of the ratio (runs resampled with replacement,
    # all owners together, so owner-to-owner differences and tails count)
    rng = np.random.default_rng(45)
    boot = []
    for _ in range(400):
        a, b = rng.integers(0, K, K), rng.integers(0, K, K)
        boot.append(g[a].var(0, ddof=1).sum() / c[b].var(0, ddof=1).sum())
    lo, hi = np.percentile(boot, [2.5, 97.5])
    print("variance ratio gpu/reference %.4f, 95%% interval [%.4f, %.4f]" % (ratio, lo, hi),
          flush=True)
    assert lo < 1.0 < hi, (ratio, lo, hi)
    se = np.sqrt((vg + vc) / K)
    z = (g.mean(0) - c.mean(0)) / se
    assert np.abs(z).mean() < 1.0 and (np.abs(z) > 4).mean() < 0.01, (np.abs(z).mean(),)


def test_sampled_indices_follow_abs_omega(device):
    """vpmc_sample_idx on 2^20 uniform draws: the index histogram matches
    |omega_j| / Omega (chi-square, p > 1e-3)"""
    import torch
    from scipy import stats

    from vortexparticlemc_cuda_amd import ops
    _say("index chi-square")
    x, y, w, hw = tg_problem(24, 4, seed=43)
    s = ops.build_subsets(torch.as_tensor(w, device=device))
    cnt = s["counts"].cpu().numpy()
    m = ops.guide_log2_for(w.size)
    rng = np.random.default_rng(44)
    for k, key in enumerate(("pos", "neg")):
        g = ops.build_guide(s["cdf_" + key], s["counts"][k:k + 1], m)
        u = torch.as_tensor(rng.random(1 << 20).astype(np.float32), device=device)
        idx = ops.sample_idx(u, s["cdf_" + key], s["counts"][k:k + 1], g, m).cpu().numpy()
        sub = s["sub_" + key][:cnt[k]].cpu().numpy()
        p = np.abs(w[sub].astype(np.float64))
        p /= p.sum()
        h = np.bincount(idx, minlength=cnt[k])
        assert stats.chisquare(h, p * idx.size).pvalue > 1e-3
