"""HIP library vs the oracle (CPU restatement of the reference's src/cpu path).

Bars: bit-exact for keys, orders and indices; <= 1e-5 relative (norm-wise)
for FP32 sums; FP64 oracles to 1e-12.  Sizes keep the oracle to seconds.
"""
import numpy as np
import pytest

import oracle as O
from conftest import rel_l2, tg_problem

pytestmark = pytest.mark.gpu


def _t(a, device, dtype=None):
    import torch
    t = torch.as_tensor(np.ascontiguousarray(a), device=device)
    return t if dtype is None else t.to(dtype)


def _say(msg):
    print("[case]", msg, flush=True)


# --------------------------------------------------------------------------- hash

@pytest.mark.parametrize("n,rsh,seed", [(10000, 128, 123), (11999, 128, 7), (5000, 6, 1),
                                        (1, 3, 0), (3000, 1, 2)])
def test_hash_bit_exact(device, n, rsh, seed):
    """FillCells/CountingSort on the device == the CPU path (keys, histogram,
    scan and the stable within-cell order), and == std::stable_sort by key;
    the reference's invariants (apps/test_spatial_hashing.cu:159-167) hold."""
    from vortexparticlemc_cuda_amd import ops
    _say("hash n=%d rsh=%d seed=%d" % (n, rsh, seed))
    rng = np.random.default_rng(seed)
    px = rng.random(n, dtype=np.float32)
    py = rng.random(n, dtype=np.float32)
    px[::97] = 1.0          # on the upper edge -> clamped
    py[::89] = 0.0
    sub = np.arange(n, dtype=np.int32)
    before, acc, vpc, vps, after = O.hash_build(rsh, rsh, 1.0, 1.0, sub, px, py)
    pos = _t(np.stack([px, py], 1), device)
    num, acc_g, vpc_g = ops.fill_cells(rsh, rsh, 1.0, 1.0, pos, _t(sub, device))
    assert np.array_equal(num.cpu().numpy(), before)
    assert np.array_equal(acc_g.cpu().numpy(), acc)
    assert np.array_equal(vpc_g.cpu().numpy(), vpc)
    vps_g = ops.counting_sort(num, acc_g, vpc_g)
    assert np.array_equal(vps_g.cpu().numpy(), vps)
    assert np.array_equal(vps, np.argsort(vpc, kind="stable"))
    assert num.sum().item() == 0 and np.array_equal(num.cpu().numpy(), after)


def test_hash_subset_and_empty(device):
    from vortexparticlemc_cuda_amd import ops
    _say("hash subset/empty")
    rng = np.random.default_rng(3)
    px = rng.random(4000, dtype=np.float32)
    py = rng.random(4000, dtype=np.float32)
    sub = np.sort(rng.choice(4000, 1500, replace=False)).astype(np.int32)
    before, acc, vpc, vps, _ = O.hash_build(10, 7, 1.0, 1.0, sub, px, py)
    num, acc_g, vpc_g = ops.fill_cells(10, 7, 1.0, 1.0, _t(np.stack([px, py], 1), device),
                                       _t(sub, device))
    assert np.array_equal(vpc_g.cpu().numpy(), vpc)
    assert np.array_equal(ops.counting_sort(num, acc_g, vpc_g).cpu().numpy(), vps)
    num, acc_g, vpc_g = ops.fill_cells(4, 4, 1.0, 1.0, _t(np.zeros((1, 2), np.float32), device),
                                       _t(np.zeros(0, np.int32), device))
    assert num.sum().item() == 0 and acc_g.sum().item() == 0


# --------------------------------------------------------------------------- subsets

@pytest.mark.parametrize("case", ["tg", "allpos", "zeros", "tiny"])
def test_subsets_cdf_and_sampler(device, case):
    import torch
    from vortexparticlemc_cuda_amd import ops
    _say("subsets " + case)
    rng = np.random.default_rng(5)
    if case == "tg":
        x, y, w, _ = tg_problem(48, 4)
    elif case == "allpos":
        w = rng.random(3000).astype(np.float32) + 0.01
    elif case == "zeros":
        w = np.zeros(500, np.float32)
        w[::7] = rng.standard_normal(w[::7].size).astype(np.float32)
    else:
        w = np.array([-1.0, 2.0], np.float32)
    ref = O.build_subsets(w)
    s = ops.build_subsets(_t(w, device))
    cnt = s["counts"].cpu().numpy()
    assert cnt[0] == ref["sub_pos"].size and cnt[1] == ref["sub_neg"].size
    assert np.array_equal(s["sub_pos"][:cnt[0]].cpu().numpy(), ref["sub_pos"])
    assert np.array_equal(s["sub_neg"][:cnt[1]].cpu().numpy(), ref["sub_neg"])
    for key, k in (("pos", 0), ("neg", 1)):
        c = s["cdf_" + key][:cnt[k]].cpu().numpy()
        if cnt[k]:
            np.testing.assert_allclose(c, ref["cdf_" + key], rtol=1e-13, atol=1e-15)
            assert c[-1] == 1.0
        tot = s["totals"][k].item()
        np.testing.assert_allclose(tot, ref["cum_" + key], rtol=1e-13)
        # sampler: guided search == numpy.searchsorted(side='right') on the device CDF
        m = ops.guide_log2_for(w.size)
        g = ops.build_guide(s["cdf_" + key], s["counts"][k:k + 1], m)
        u = rng.random(20000).astype(np.float32)
        # edge draws, including out-of-range injected values (a float CDF value
        # cast to FP32 can round to exactly 1.0)
        u[:8] = [0.0, np.float32(1 - 2 ** -24), 0.5, 0.25, 1.0, 1.5, -0.25, np.nan]
        if cnt[k]:
            u[8:8 + min(cnt[k], 100)] = c[:min(cnt[k], 100)].astype(np.float32)
        idx = ops.sample_idx(_t(u, device), s["cdf_" + key], s["counts"][k:k + 1], g, m)
        if cnt[k]:
            exp = np.minimum(np.searchsorted(c, u.astype(np.float64), side="right"), cnt[k] - 1)
            assert np.array_equal(idx.cpu().numpy(), exp)
            exp_o = np.array([O.lib().orc_sample_idx(float(v), c, int(cnt[k])) for v in u[:500]])
            assert np.array_equal(idx.cpu().numpy()[:500], exp_o)
        else:
            assert (idx.cpu().numpy() == -1).all()
        torch.cuda.synchronize()


# --------------------------------------------------------------------------- query

@pytest.mark.parametrize("Nx,hwr", [(32, 4), (40, 4), (24, 4)])
def test_query_view_matches_cpu(device, Nx, hwr):
    """queryVorticityAbs on the reference view == evalVorticityAbs_cpu"""
    from vortexparticlemc_cuda_amd import ops
    _say("query_view Nx=%d" % Nx)
    x, y, w, hw = tg_problem(Nx, hwr)
    rsh = O.rsh_of(Nx, hwr)
    prob = O.Problem(x, y, w, rsh, rsh)
    rng = np.random.default_rng(11)
    q = (rng.random((3000, 2)) * 1.6 - 0.3).astype(np.float32)
    pos = _t(np.stack([x, y], 1), device)
    wt = _t(w, device)
    for sign in (1, -1, 0):
        for use_signed in (0, 1):
            key = "pos" if sign >= 0 else "neg"
            if sign == 0:   # all-particle view (makeAllView)
                sub = np.arange(x.size, dtype=np.int32)
                _, acc, _, vps, _ = O.hash_build(rsh, rsh, 1.0, 1.0, sub, x, y)
                v = O.View(x.ctypes.data, y.ctypes.data, w.ctypes.data, x.size, vps.ctypes.data,
                           acc.ctypes.data, rsh, rsh, 1.0, 1.0, x.size, sub.ctypes.data)
                exp = np.array([O.lib().orc_eval_vorticity(O.C.byref(v), float(a), float(b), 0,
                                                           use_signed, 0, hw, 1) for a, b in q])
            else:
                sub = prob.sub["sub_" + key]
                _, acc, _, vps, _ = O.hash_build(rsh, rsh, 1.0, 1.0, sub, x, y)
                exp = prob.eval(q[:, 0], q[:, 1], sign, hw, use_signed=bool(use_signed))
            got = ops.query_view(pos, wt, _t(vps, device), _t(acc, device), _t(sub, device),
                                 rsh, rsh, 1.0, 1.0, sub.size, _t(q, device), sign, hw,
                                 use_signed=bool(use_signed)).cpu().numpy()
            np.testing.assert_allclose(got, exp, rtol=2e-6, atol=1e-6 * np.abs(exp).max())


def _brute_periodic(qx, qy, px, py, w, hw):
    """FP64 sum_j |w_j| W(|q - x_j - image|) over the 3 x 3 periodic images"""
    qx = qx.astype(np.float64)[:, None]
    qy = qy.astype(np.float64)[:, None]
    px = px.astype(np.float64)[None, :]
    py = py.astype(np.float64)[None, :]
    aw = np.abs(w.astype(np.float64))
    sig = 40.0 / (7 * np.pi * hw * hw)
    out = np.zeros(qx.shape[0])
    for ix in (-1, 0, 1):
        for iy in (-1, 0, 1):
            q = np.hypot(px + ix - qx, py + iy - qy) / hw
            W = np.where(q <= 0.5, sig * (1 - 6 * q ** 2 + 6 * q ** 3),
                         np.where(q < 1, 2 * sig * (1 - q) ** 3, 0.0))
            out += W @ aw
    return out


@pytest.mark.parametrize("Nx,hwr,cph", [(32, 4, 8), (64, 8, 4), (50, 16, 8), (30, 4, 2),
                                         (64, 16, 2)])
def test_field_query_matches_cpu(device, Nx, hwr, cph):
    """fine-grid field sum == the FP64 brute-force periodic sum, and == the
    CPU query on the reference grid (same in-support set, other summation
    order) where that grid has >= 3 cells a side; the field's cell order is
    stable for both of the build's sorts (64 x 64, hwr 16 on an 8 x 8 grid:
    32 particles a cell, the radix-sort build; the others, the atomic one).
    With fewer than 3 reference cells a side, the
    reference's 3 x 3 stencil visits a cell more than once
    (src/cpu/evaluate_vorticity.cpp:22-40): at Nx=50, hwr=16 the grid is 1 x 1
    and its sum counts particles several times (9 times at the box centre);
    the field does not reproduce that."""
    from vortexparticlemc_cuda_amd import ops
    _say("field_query Nx=%d hwr=%d" % (Nx, hwr))
    x, y, w, hw = tg_problem(Nx, hwr)
    rsh = O.rsh_of(Nx, hwr)
    prob = O.Problem(x, y, w, rsh, rsh)
    rng = np.random.default_rng(12)
    q = (rng.random((4000, 2)) * 1.5 - 0.25).astype(np.float32)
    q[:200] = np.stack([x[:200], y[:200]], 1)   # exactly at particles
    pos = _t(np.stack([x, y], 1), device)
    nx, ny = ops.field_grid(hw, 1.0, 1.0, cph)
    for sign, key in ((1, "sub_pos"), (-1, "sub_neg")):
        sub = prob.sub[key]
        fb = ops.FieldBuffers(nx, ny, 1.0, 1.0, sub.size, device)
        ops.field_build(fb, pos, _t(w, device), sub=_t(sub, device), abs_w=True)
        # stable cell order of the field
        cs = fb.cell_start.cpu().numpy()
        order = fb.order[:sub.size].cpu().numpy()
        assert cs[-1] == sub.size
        keys = np.minimum((x[sub] * np.float32(nx)).astype(np.int64), nx - 1) * ny + \
            np.minimum((y[sub] * np.float32(ny)).astype(np.int64), ny - 1)
        assert np.array_equal(order, np.argsort(keys, kind="stable"))
        got = ops.field_query(fb, hw, _t(q, device)).cpu().numpy()
        bf = _brute_periodic(q[:, 0], q[:, 1], x[sub], y[sub], w[sub], hw)
        assert rel_l2(got, bf) < 1e-6
        np.testing.assert_allclose(got, bf, rtol=1e-5, atol=2e-6 * np.abs(bf).max())
        if rsh >= 3:
            exp = prob.eval(q[:, 0], q[:, 1], sign, hw)
            assert rel_l2(got, exp) < 1e-6
            np.testing.assert_allclose(got, exp, rtol=1e-5, atol=2e-6 * np.abs(exp).max())


# --------------------------------------------------------------------------- estimator

def _gpu_sim(x, y, w, hw, Nx, hwr, mode, cph=8, seed=0, algo=0):
    from vortexparticlemc_cuda_amd.sim import SimParam, Simulation
    P = SimParam(dt=1e-3, resX=Nx, hwr=hwr)
    assert abs(float(P.h_w) - hw) < 1e-7
    sim = Simulation(P, cells_per_hw=cph, mode=mode, seed=seed, algo=algo)
    sim.set_particles(np.stack([x, y], 1), w)
    return sim


@pytest.mark.parametrize("algo", [0, 1])
@pytest.mark.parametrize("mode", [O.MODE_REFERENCE, O.MODE_CONSISTENT])
@pytest.mark.parametrize("aux", [0, 1])
def test_estimator_replay(device, mode, aux, algo):
    """velocities on injected uniforms: GPU == oracle to 1e-5 (norm-wise), for
    the binned (0) and the direct (1) algorithm"""
    _say("estimator replay mode=%d aux=%d algo=%d" % (mode, aux, algo))
    Nx, hwr, N1, N2 = 32, 4, 32, 32
    x, y, w, hw = tg_problem(Nx, hwr)
    rsh = O.rsh_of(Nx, hwr)
    sim = _gpu_sim(x, y, w, hw, Nx, hwr, mode, algo=algo)
    # the oracle sees the device's FP64 CDF and totals (parity of the sampler
    # is tested separately; this isolates the estimator)
    over = dict(cdf_pos=sim.sub["cdf_pos"][:sim.cnt[0]].cpu().numpy(),
                cdf_neg=sim.sub["cdf_neg"][:sim.cnt[1]].cpu().numpy(),
                cum_pos=sim.sub["totals"][0].item(), cum_neg=sim.sub["totals"][1].item())
    prob = O.Problem(x, y, w, rsh, rsh, cdf_override=over)
    b1 = 3
    per = N1 * b1 + N2 * (10 if aux else 2)
    rng = np.random.default_rng(21)
    U = rng.random(x.size * 2 * per).astype(np.float32)
    vg = sim.estimate(N1, N2, uniforms=U, aux_draws=aux).cpu().numpy()
    P = prob.params(mode, N1, N2, hw, U=U, aux_draws=aux)
    vo = prob.velocities(P)
    assert rel_l2(vg, vo) < 1e-5, rel_l2(vg, vo)


@pytest.mark.parametrize("algo", [0, 1])
@pytest.mark.parametrize("mode", [O.MODE_REFERENCE, O.MODE_CONSISTENT])
def test_estimator_philox_stream(device, mode, algo):
    """the device Philox stream equals the oracle's (same counters, same key)"""
    _say("estimator philox mode=%d algo=%d" % (mode, algo))
    Nx, hwr, N1, N2 = 32, 4, 64, 64
    x, y, w, hw = tg_problem(Nx, hwr, seed=4)
    rsh = O.rsh_of(Nx, hwr)
    sim = _gpu_sim(x, y, w, hw, Nx, hwr, mode, seed=0x1234ABCD5678, algo=algo)
    sim.step_index = 7
    sim.sim = 3
    over = dict(cdf_pos=sim.sub["cdf_pos"][:sim.cnt[0]].cpu().numpy(),
                cdf_neg=sim.sub["cdf_neg"][:sim.cnt[1]].cpu().numpy(),
                cum_pos=sim.sub["totals"][0].item(), cum_neg=sim.sub["totals"][1].item())
    prob = O.Problem(x, y, w, rsh, rsh, cdf_override=over)
    vg = sim.estimate(N1, N2).cpu().numpy()
    P = prob.params(mode, N1, N2, hw, seed=0x1234ABCD5678, step=7, sim=3)
    vo = prob.velocities(P)
    assert rel_l2(vg, vo) < 1e-5, rel_l2(vg, vo)


def test_step_replay_three_steps(device):
    """three full steps (field rebuild, estimate, Euler + wrap) == oracle steps"""
    _say("step replay")
    Nx, hwr, N1, N2 = 32, 4, 16, 16
    x, y, w, hw = tg_problem(Nx, hwr, seed=9)
    rsh = O.rsh_of(Nx, hwr)
    sim = _gpu_sim(x, y, w, hw, Nx, hwr, O.MODE_REFERENCE)
    over = dict(cdf_pos=sim.sub["cdf_pos"][:sim.cnt[0]].cpu().numpy(),
                cdf_neg=sim.sub["cdf_neg"][:sim.cnt[1]].cpu().numpy(),
                cum_pos=sim.sub["totals"][0].item(), cum_neg=sim.sub["totals"][1].item())
    per = N1 * 3 + N2 * 2
    U = np.random.default_rng(2).random(x.size * 2 * per).astype(np.float32)
    px, py = x.copy(), y.copy()
    dt = 1e-2
    for s in range(3):
        sim.step(N1, N2, dt=dt, uniforms=U)
        prob = O.Problem(px, py, w, rsh, rsh, cdf_override=over)
        P = prob.params(O.MODE_REFERENCE, N1, N2, hw, U=U)
        vel = np.zeros(2 * px.size, np.float32)
        O.lib().orc_step(O.C.byref(P), O.C.byref(prob.sets[0]), O.C.byref(prob.sets[1]),
                         px.size, px, py, dt, vel, 0)
        pg = sim.pos.cpu().numpy()
        d = np.stack([px, py], 1) - pg
        d -= np.rint(d)
        assert np.abs(d).max() < 1e-6, (s, np.abs(d).max())
        # continue both from identical positions
        sim.pos.copy_(_t(np.stack([px, py], 1), device))


# --------------------------------------------------------------------------- CG / direct

def test_cg_taylor_green_closed_form(device):
    """TG on the periodic lattice: A b = lambda_11 b, so w = b / lambda_11.
    b rounded to FP32 is that eigenvector plus ~3e-8 noise over every mode;
    most Gram eigenvalues are tiny, so the meaningful tolerance is ~1e-6 (an
    exact solve would amplify the rounding noise).  MINRES from zero stops
    after one iteration with w = b / lambda_11."""
    from vortexparticlemc_cuda_amd import ops
    _say("cg tg closed form")
    for Nx, hwr in ((64, 8), (100, 16)):
        x, y = O.lattice(Nx)
        hw = hwr / Nx
        b = O.tg_vorticity(x, y).astype(np.float32)
        pos = _t(np.stack([x, y], 1).astype(np.float32), device)
        lam = O.tg_lambda11(Nx, hwr)
        w, it, rr, st = ops.cg_rbf(pos, _t(b, device), hw, tol=1e-6, refine=True)
        assert st == 0 and it == 1 and rr < 1e-6, (st, it, rr)
        assert rel_l2(w.cpu().numpy(), b / lam) < 2e-7
        # the reference's FP32 CG at its tolerance (src/cuda/cg_rbf_init.cu:53):
        # one iteration reaches rr ~ 3e-6 (Nx=64) / 1.4e-5 (Nx=100), after
        # which FP32 CG drifts on the rounding noise (rr grows)
        w, it, rr, st = ops.cg_rbf(pos, _t(b, device), hw, tol=5e-3, refine=False)
        assert st == 0 and it == 1 and rr < 5e-5, (st, it, rr)
        assert rel_l2(w.cpu().numpy(), b / lam) < 1e-4


def _dense_gram(x, y, hw):
    dx = x[:, None].astype(np.float64) - x[None, :]
    dy = y[:, None].astype(np.float64) - y[None, :]
    dx -= np.rint(dx)
    dy -= np.rint(dy)
    q = np.hypot(dx, dy) / hw
    sig = 40.0 / (7 * np.pi * hw * hw)
    return np.where(q <= 0.5, sig * (6 * (q ** 3 - q ** 2) + 1),
                    np.where(q < 1, 2 * sig * (1 - q) ** 3, 0))


def test_cg_matches_dense_solve(device):
    """several Fourier modes on the lattice (distinct eigenvalues): CG needs
    several iterations and must agree with a dense FP64 solve"""
    from vortexparticlemc_cuda_amd import ops
    _say("cg dense")
    Nx, hwr = 24, 4
    x, y = O.lattice(Nx)
    x = x.astype(np.float32)
    y = y.astype(np.float32)
    hw = hwr / Nx
    tp = 2 * np.pi
    b = (np.sin(tp * x) * np.sin(tp * y) + 0.5 * np.cos(2 * tp * x) * np.sin(tp * y)
         + 0.25 * np.cos(3 * tp * x) + 0.1 * np.sin(tp * (x + 2 * y))).astype(np.float32)
    A = _dense_gram(x, y, hw)
    wd = np.linalg.solve(A, b.astype(np.float64))
    pos = _t(np.stack([x, y], 1), device)
    w, it, rr, st = ops.cg_rbf(pos, _t(b, device), hw, tol=1e-10, refine=True)
    assert st == 0 and rr < 1e-10, (st, it, rr)
    assert it >= 3
    assert rel_l2(w.cpu().numpy(), wd) < 1e-5
    # reference semantics (no refinement, tol 5e-3, cg_rbf_init.cu:53-116)
    w2, it2, rr2, st2 = ops.cg_rbf(pos, _t(b, device), hw, tol=5e-3, refine=False)
    wo, ito, rro = O.cg_rbf(x, y, b, hw, tol=5e-3)
    assert st2 == 0 and rr2 < 6e-3 and abs(it2 - ito) <= 1


def test_cg_reports_breakdown(device):
    """the lattice Gram matrix is indefinite (negative eigenvalues beyond
    |k| h_w ~ 1.81); a generic right-hand side meets p.Ap <= 0, which the
    reference's CG does not check (src/cuda/cg_rbf_init.cu:91-92)"""
    from vortexparticlemc_cuda_amd import ops
    _say("cg breakdown")
    x, y, _, hw = tg_problem(32, 4, jitter=0.2, seed=4)
    assert np.linalg.eigvalsh(_dense_gram(x, y, hw)).min() < 0
    b = O.tg_vorticity(x, y).astype(np.float32)
    pos = _t(np.stack([x, y], 1), device)
    w, it, rr, st = ops.cg_rbf(pos, _t(b, device), hw, tol=1e-12, refine=False)
    _, ito, _ = O.cg_rbf(x, y, b, hw, tol=1e-12)
    assert ito < 0 and st == 5


def test_direct_sums_fp64(device):
    from vortexparticlemc_cuda_amd import ops
    _say("direct sums")
    x, y, w, hw = tg_problem(32, 4, seed=3)
    rng = np.random.default_rng(1)
    q = rng.random((300, 2))
    pos = np.stack([x, y], 1).astype(np.float64)
    ug = ops.direct_sum_eps(_t(q, device), _t(pos, device), _t(w.astype(np.float64), device),
                            eps=0.5 * hw).cpu().numpy()
    uo = O.direct_eps(q[:, 0], q[:, 1], x, y, w, eps=0.5 * hw)
    assert rel_l2(ug, uo) < 1e-12
    ub = ops.direct_sum_blob(_t(q, device), _t(pos, device), _t(w.astype(np.float64), device),
                             hw).cpu().numpy()
    uo = O.direct_blob(q[:, 0], q[:, 1], x, y, w, hw)
    assert rel_l2(ub, uo) < 1e-12
