"""Pins the oracle (CPU restatement) -- CPU only.

Each check is independent of the formula it checks: the reference's own
outputs (golden vectors written from its compiled src/cpu sources), closed
forms, quadrature, brute force, library routines (numpy), published
known-answer vectors, and the analytic Taylor-Green velocity.
"""
import json
import os

import numpy as np
import pytest
from scipy import integrate

import oracle as O
from conftest import rel_l2, tg_problem

GOLD = os.path.join(os.path.dirname(__file__), "golden")


@pytest.fixture(scope="module")
def gold():
    return np.load(os.path.join(GOLD, "ref_cpu_vectors.npz"))


# --------------------------------------------------------------------- kernel W

@pytest.mark.parametrize("h", [0.16, 0.08, 0.0078125])
def test_spline_closed_forms(h):
    L = O.lib()
    # W(0) = 40 / (7 pi h^2)
    assert np.isclose(L.orc_spline_d(0.0, h), 40 / (7 * np.pi * h * h), rtol=1e-14)
    # continuity at q = 1/2 (value sig2 / 4) and zero at q >= 1
    s = 40 / (7 * np.pi * h * h)
    assert np.isclose(L.orc_spline_d(0.5 * h * (1 - 1e-12), h), s / 4, rtol=1e-9)
    assert np.isclose(L.orc_spline_d(0.5 * h * (1 + 1e-12), h), s / 4, rtol=1e-9)
    assert L.orc_spline_d(h, h) == 0.0 and L.orc_spline_d(1.5 * h, h) == 0.0
    # unit mass: int 2 pi r W(r) dr = 1 (the radial integral is 7/80)
    m, _ = integrate.quad(lambda r: 2 * np.pi * r * L.orc_spline_d(r, h), 0, h,
                          points=[0.5 * h], epsabs=1e-14, epsrel=1e-13)
    assert abs(m - 1.0) < 1e-12
    # the FP32 restatement agrees with the FP64 one
    for r in np.linspace(0, 1.1 * h, 37):
        assert np.isclose(L.orc_spline_f(r, h), L.orc_spline_d(r, h), rtol=2e-6,
                          atol=1e-6 * s)


def test_mass_fraction_and_inverse():
    L = O.lib()
    h = 1.0
    for q in np.linspace(0, 1, 41):
        m, _ = integrate.quad(lambda r: 2 * np.pi * r * L.orc_spline_d(r, h), 0, q,
                              points=[0.5] if q > 0.5 else None, epsabs=1e-14)
        assert abs(L.orc_mass_fraction(q) - m) < 1e-12
    for u in np.concatenate([np.linspace(0, 1, 101), [1e-12, 1e-7, 1 - 1e-7, 1 - 2 ** -24]]):
        q = L.orc_mass_inverse(u)
        assert 0 <= q <= 1
        assert abs(L.orc_mass_fraction(q) - u) < 1e-13
    # sampled radii follow 2 pi r W(r): Kolmogorov-Smirnov against the quadrature CDF
    from scipy import stats
    rng = np.random.default_rng(0)
    qs = np.array([L.orc_mass_inverse(u) for u in rng.random(20000)])
    ks = stats.kstest(qs, lambda t: np.array([L.orc_mass_fraction(v) for v in np.atleast_1d(t)]))
    assert ks.pvalue > 1e-3


# --------------------------------------------------------------------- hash

@pytest.mark.parametrize("tag", ["a", "b", "c"])
def test_hash_matches_reference_vectors(gold, tag):
    px, py = gold["hash_%s_px" % tag], gold["hash_%s_py" % tag]
    rsh = int(gold["hash_%s_rsh" % tag])
    sub = np.arange(px.size, dtype=np.int32)
    before, acc, vpc, vps, after = O.hash_build(rsh, rsh, 1.0, 1.0, sub, px, py)
    assert np.array_equal(before, gold["hash_%s_num" % tag])
    assert np.array_equal(acc, gold["hash_%s_acc" % tag])
    assert np.array_equal(vpc, gold["hash_%s_cell" % tag])
    assert np.array_equal(vps, gold["hash_%s_map" % tag])
    # std::stable_sort by key, and the reference driver's invariants
    assert np.array_equal(vps, np.argsort(vpc, kind="stable"))
    assert after.sum() == 0
    for c in range(rsh * rsh):
        assert (vpc[vps[acc[c]:acc[c] + before[c]]] == c).all()


def test_hash_against_live_reference():
    R = O.ref_lib()
    if R is None:
        pytest.skip("oracle/_ref not built here")
    rng = np.random.default_rng(77)
    for n, rsh in ((10000, 128), (11999, 128), (64, 2)):
        px = rng.random(n, dtype=np.float32)
        py = rng.random(n, dtype=np.float32)
        sub = np.arange(n, dtype=np.int32)
        num = np.zeros(rsh * rsh, np.int32)
        acc = np.zeros(rsh * rsh, np.int32)
        vpc = np.zeros(n, np.int32)
        vpm = np.zeros(n, np.int32)
        R.ref_fill_cells(rsh, rsh, 1.0, 1.0, n, sub, n, px, py, num, acc, vpc)
        R.ref_counting_sort(n, rsh * rsh, num, acc, vpc, vpm)
        _, acc2, vpc2, vps2, _ = O.hash_build(rsh, rsh, 1.0, 1.0, sub, px, py)
        assert np.array_equal(acc, acc2) and np.array_equal(vpc, vpc2)
        assert np.array_equal(vpm, vps2)
        # the OpenMP fill agrees with the serial one (its sort is the defect at
        # src/omp/hash_omp.cpp:65 and is not used)
        num3 = np.zeros(rsh * rsh, np.int32)
        acc3 = np.zeros(rsh * rsh, np.int32)
        vpc3 = np.zeros(n, np.int32)
        R.ref_fill_cells_omp(rsh, rsh, 1.0, 1.0, n, sub, n, px, py, num3, acc3, vpc3)
        assert np.array_equal(acc3, acc) and np.array_equal(vpc3, vpc)


# --------------------------------------------------------------------- query

def _gold_problem(gold, use_f32_cdf=True):
    x, y, w = gold["px"], gold["py"], gold["w"]
    rsh = int(gold["rsh"])
    over = None
    if use_f32_cdf:
        over = dict(cdf_pos=gold["cdf_pos_f32"].astype(np.float64),
                    cdf_neg=gold["cdf_neg_f32"].astype(np.float64),
                    cum_pos=float(gold["cum_f32"][0]), cum_neg=float(gold["cum_f32"][1]))
    return O.Problem(x, y, w, rsh, rsh, cdf_override=over), float(gold["h_w"])


def test_query_matches_reference_vectors(gold):
    prob, hw = _gold_problem(gold)
    q = gold["q"]
    for sign in (1, -1):
        got = prob.eval(q[:, 0], q[:, 1], sign, hw)
        assert np.array_equal(got, gold["query_%+d" % sign])


def test_query_against_brute_force():
    """hashed query == all-particle min-image sum (queryVorticityAbsNaive,
    include/vorticity_query.cuh:82-103, with min image) in FP64"""
    x, y, w, hw = tg_problem(40, 4, seed=8)
    rsh = O.rsh_of(40, 4)
    prob = O.Problem(x, y, w, rsh, rsh)
    rng = np.random.default_rng(1)
    q = rng.random((300, 2))
    for sign in (1, -1):
        got = prob.eval(q[:, 0], q[:, 1], sign, hw)
        dx = x[None, :].astype(np.float64) - q[:, :1]
        dy = y[None, :].astype(np.float64) - q[:, 1:]
        dx -= np.rint(dx)
        dy -= np.rint(dy)
        r = np.hypot(dx, dy) / hw
        f = np.where(r <= 0.5, 6 * (r ** 3 - r ** 2) + 1, np.where(r < 1, 2 * (1 - r) ** 3, 0.0))
        mask = (w > 0) if sign > 0 else (w < 0)
        exp = (np.abs(w) * mask * f).sum(1) * 40 / (7 * np.pi * hw * hw)
        assert rel_l2(got, exp) < 1e-6


# --------------------------------------------------------------------- subsets / sampler

def test_subsets_and_sampler():
    rng = np.random.default_rng(3)
    w = rng.standard_normal(5000).astype(np.float32)
    w[::13] = 0
    s = O.build_subsets(w)
    assert np.array_equal(s["sub_pos"], np.nonzero(w > 0)[0])
    assert np.array_equal(s["sub_neg"], np.nonzero(w < 0)[0])
    cp = np.cumsum(w[w > 0].astype(np.float64))
    np.testing.assert_allclose(s["cdf_pos"], cp / cp[-1], rtol=1e-14)
    assert s["cdf_pos"][-1] == 1.0 and np.isclose(s["cum_pos"], cp[-1], rtol=1e-14)
    u = rng.random(3000).astype(np.float32)
    u[:3] = [0.0, np.float32(1 - 2 ** -24), s["cdf_pos"][10]]
    got = np.array([O.lib().orc_sample_idx(float(v), s["cdf_pos"], s["cdf_pos"].size) for v in u])
    exp = np.minimum(np.searchsorted(s["cdf_pos"], u.astype(np.float64), side="right"),
                     s["cdf_pos"].size - 1)
    assert np.array_equal(got, exp)


# --------------------------------------------------------------------- Philox

def test_philox_known_answers():
    d = json.load(open(os.path.join(GOLD, "philox_kat.json")))
    for c in d["cases"]:
        assert O.philox(c["ctr"], c["key"]).tolist() == c["out"]
    # Random123 kat_vectors for philox4x32_10
    assert O.philox([0, 0, 0, 0], [0, 0]).tolist() == [0x6627e8d5, 0xe169c58d, 0xbc57ac4c,
                                                     0x9b00dbd8]
    assert O.philox([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344],
                    [0xa4093822, 0x299f31d0]).tolist() == [0xd16cfe09, 0x94fdcceb, 0x5001e420,
                                                          0x24126ea1]


def test_philox_stream_uniformity():
    """the estimator's counter layout (k, owner, step, sim/sign/branch tag):
    the four words' uniforms pass chi-square (256 bins, and 16 x 16 bins on
    pairs of words) and Kolmogorov-Smirnov tests at p > 1e-3"""
    from scipy import stats
    L = O.lib()
    words = []
    for owner in range(2):
        for k in range(16384):
            words.append(O.philox([k, owner, 7, (3 << 8) | (1 << 7) | (2 << 4)],
                                  [0x1234ABCD, 0x5678]))
    w = np.array(words, np.uint32)                     # (32768, 4)
    u = np.vectorize(lambda v: L.orc_u01(int(v)), otypes=[np.float32])(w)
    assert (u >= 0).all() and (u < 1).all()
    for c in range(4):
        h = np.histogram(u[:, c], bins=256, range=(0, 1))[0]
        assert stats.chisquare(h).pvalue > 1e-3, c
        assert stats.kstest(u[:, c], "uniform").pvalue > 1e-3, c
    h2 = np.histogram2d(u[:, 0], u[:, 1], bins=16, range=[[0, 1], [0, 1]])[0].ravel()
    assert stats.chisquare(h2).pvalue > 1e-3


def test_uniform_conversion():
    L = O.lib()
    assert L.orc_u01(0) == 0.0
    assert L.orc_u01(0xffffffff) == np.float32(1 - 2 ** -24)
    assert L.orc_u01(1 << 31) == 0.5


# --------------------------------------------------------------------- estimator

@pytest.mark.parametrize("aux", [0, 1])
def test_mc_matches_reference_vectors(gold, aux):
    """monte_carlo_bs_cpu on the reference's own mt19937 draws, replayed"""
    prob, hw = _gold_problem(gold)
    N1, N2, n_own = int(gold["mc_N1"]), int(gold["mc_N2"]), int(gold["mc_owners"])
    P = prob.params(O.MODE_REFERENCE, N1, N2, hw, U=gold["mc_aux%d_U" % aux], aux_draws=aux)
    v = prob.velocities(P, owners=n_own)
    assert np.array_equal(v, gold["mc_aux%d_vel" % aux])


def test_mc_against_live_reference():
    R = O.ref_lib()
    if R is None:
        pytest.skip("oracle/_ref not built here")
    x, y, w, hw = tg_problem(40, 4, seed=5)
    rsh = O.rsh_of(40, 4)
    cnt = np.zeros(2, np.int32)
    cp = np.zeros(x.size, np.float32)
    cn = np.zeros(x.size, np.float32)
    cum = np.zeros(2, np.float32)
    R.ref_subsets_f32(x.size, w, cnt, cp, cn, cum)
    over = dict(cdf_pos=cp[:cnt[0]].astype(np.float64), cdf_neg=cn[:cnt[1]].astype(np.float64),
                cum_pos=float(cum[0]), cum_neg=float(cum[1]))
    prob = O.Problem(x, y, w, rsh, rsh, cdf_override=over)
    N1, N2, n_own, seed = 24, 40, 64, 99
    vel = np.zeros(2 * n_own, np.float32)
    R.ref_mc_velocities(x.size, x, y, w, rsh, rsh, 1.0, 1.0, hw, 1, N1, N2, 0, seed, 0, n_own,
                        vel)
    U = np.zeros(n_own * 2 * (3 * N1 + 2 * N2), np.float32)
    R.ref_uniform_stream(seed, U.size, U)
    v = prob.velocities(prob.params(O.MODE_REFERENCE, N1, N2, hw, U=U), owners=n_own)
    assert np.array_equal(v, vel.reshape(-1, 2))
    # the multi-threaded timing harness on one thread is the same call sequence
    vm = np.zeros(2 * n_own, np.float32)
    R.ref_mc_velocities_mt(x.size, x, y, w, rsh, rsh, 1.0, 1.0, hw, 1, N1, N2, seed, 0, n_own,
                           1, vm)
    assert np.array_equal(vm, vel)


def test_consistent_estimator_is_unbiased():
    """the consistent-mode MC mean converges to the exact periodic blob
    velocity; the reference formulas do not (their disk density does not
    match the sqrt(u) radius draw)"""
    x, y, w, hw = tg_problem(32, 4, jitter=0.0)
    rsh = O.rsh_of(32, 4)
    prob = O.Problem(x, y, w, rsh, rsh)
    owners = 24
    exact = O.direct_blob(x[:owners], y[:owners], x, y, w, hw)
    runs = []
    for seed in range(6):
        P = prob.params(O.MODE_CONSISTENT, 1024, 1024, hw, seed=seed + 1)
        runs.append(prob.velocities(P, owners=owners))
    runs = np.array(runs, np.float64)
    mean = runs.mean(0)
    sem = runs.std(0, ddof=1) / np.sqrt(len(runs))
    z = np.abs(mean - exact) / np.maximum(sem, 1e-9)
    assert np.mean(z < 3.5) > 0.9, z
    # analytic Taylor-Green velocity of the lattice field (aliasing of the
    # 4-spacing blob leaves ~4e-5 at this resolution)
    ut = O.tg_velocity(x[:owners], y[:owners])
    assert rel_l2(exact, ut) < 1e-4
    assert rel_l2(mean, ut) < 0.05
    Pr = prob.params(O.MODE_REFERENCE, 1024, 1024, hw, seed=3)
    vr = prob.velocities(Pr, owners=owners)
    assert rel_l2(vr, ut) > 0.3


# --------------------------------------------------------------------- periodic kernel, sums

def test_periodic_kernel():
    import math
    c = O.kper_coeffs(24)
    assert abs(c[0] - math.gamma(0.25) ** 8 / (960 * math.pi ** 2)) < 1e-13
    # brute-force lattice sums for G_8, G_12
    a = np.arange(-200, 201)
    X, Y = np.meshgrid(a, a, indexing="ij")
    z = (X + 1j * Y).ravel()
    z = z[z != 0]
    for m in (2, 3):
        assert abs(np.sum(z ** (-4 * m)).real - c[m - 1]) < 1e-9
    # K_per is periodic across the cell edges
    def kper(rx, ry):
        kx, ky = O.kernel_smooth(rx, ry)
        r2 = rx * rx + ry * ry
        return np.array([-ry / (2 * np.pi * r2) + kx, rx / (2 * np.pi * r2) + ky])
    for t in np.linspace(-0.45, 0.45, 7):
        np.testing.assert_allclose(kper(0.5 - 1e-10, t), kper(-0.5 + 1e-10, t), atol=1e-8)
        np.testing.assert_allclose(kper(t, 0.5 - 1e-10), kper(t, -0.5 + 1e-10), atol=1e-8)


@pytest.mark.parametrize("Nx,hwr", [(64, 8), (100, 16), (48, 8)])
def test_blob_direct_sum_is_taylor_green(Nx, hwr):
    x, y = O.lattice(Nx)
    w = O.tg_vorticity(x, y) / O.tg_lambda11(Nx, hwr)
    rng = np.random.default_rng(2)
    q = rng.random((40, 2))
    u = O.direct_blob(q[:, 0], q[:, 1], x, y, w, hwr / Nx)
    assert rel_l2(u, O.tg_velocity(q[:, 0], q[:, 1])) < 1e-5


def test_eps_direct_sum_brute_force():
    x, y, w, hw = tg_problem(20, 4, seed=1)
    q = np.stack([x, y], 1).astype(np.float64)[:50]
    u = O.direct_eps(q[:, 0], q[:, 1], x, y, w, eps=0.5 * hw, skip_self=True)
    d = q[:, None, :] - np.stack([x, y], 1)[None].astype(np.float64)
    d -= np.rint(d)
    den = 2 * np.pi * ((d ** 2).sum(-1) + (0.5 * hw) ** 2)
    K = np.stack([-d[..., 1], d[..., 0]], -1) / den[..., None]
    mask = np.ones((50, x.size))
    mask[np.arange(50), np.arange(50)] = 0
    exp = (K * (w[None, :, None] * mask[..., None])).sum(1)
    assert rel_l2(u, exp) < 1e-13


# --------------------------------------------------------------------- CG

def test_lambda11_known_values():
    """lambda_11 h^2 of the Taylor-Green mode on the 100 x 100 and 200 x 200
    lattices at hwr = 16 (independently computed values)"""
    assert abs(O.tg_lambda11(100, 16) * 1e-4 - 0.92289) < 1e-5
    assert abs(O.tg_lambda11(200, 16) * 2.5e-5 - 0.98020) < 1e-5


def test_cg_closed_form_and_dense():
    Nx, hwr = 32, 4
    x, y = O.lattice(Nx)
    hw = hwr / Nx
    b = O.tg_vorticity(x, y)
    w, it, rr = O.cg_rbf(x, y, b, hw, tol=1e-12)
    assert it >= 0 and rr < 1e-12
    assert rel_l2(w, b / O.tg_lambda11(Nx, hwr)) < 1e-10
    # the operator equals the dense Gram matrix (min image, diagonal kept)
    rng = np.random.default_rng(4)
    xj = (x + 0.2 / Nx * rng.standard_normal(x.size)) % 1
    yj = (y + 0.2 / Nx * rng.standard_normal(y.size)) % 1
    dx = xj[:, None] - xj[None]
    dy = yj[:, None] - yj[None]
    dx -= np.rint(dx)
    dy -= np.rint(dy)
    r = np.hypot(dx, dy) / hw
    s = 40 / (7 * np.pi * hw * hw)
    A = np.where(r <= 0.5, s * (6 * (r ** 3 - r ** 2) + 1), np.where(r < 1, 2 * s * (1 - r) ** 3, 0))
    v = rng.standard_normal(x.size)
    assert rel_l2(O.rbf_matvec(xj, yj, v, hw), A @ v) < 1e-13
    # the Gram matrix is indefinite on the lattice: a generic right-hand
    # side makes CG meet p.Ap <= 0, which the oracle reports (negative count)
    assert np.linalg.eigvalsh(A).min() < 0
    bj = O.tg_vorticity(xj, yj)
    w, it, rr = O.cg_rbf(xj, yj, bj, hw, tol=1e-12)
    assert it < 0 or rel_l2(A @ w, bj) < 1e-11


# --------------------------------------------------------------------- deformed query

def _det1_F(n, rng, smax):
    a, b = rng.uniform(0, 2 * np.pi, n), rng.uniform(0, 2 * np.pi, n)
    s = rng.uniform(1.0, smax, n)
    ca, sa, cb, sb = np.cos(a), np.sin(a), np.cos(b), np.sin(b)
    return np.stack([ca * s * cb - sa * sb / s, -ca * s * sb - sa * cb / s,
                     sa * s * cb + ca * sb / s, -sa * s * sb + ca * cb / s], 1).astype(np.float32)


def _W64(r, h):
    q = r / h
    sig = 40.0 / (7 * np.pi * h * h)
    return np.where(q <= 0.5, sig * (1 - 6 * q ** 2 + 6 * q ** 3),
                    np.where(q < 1, 2 * sig * (1 - q) ** 3, 0.0))


def test_deformed_query_reduces_to_the_undeformed_one():
    """the oracle's deformed query (include/vorticity_query.cuh:52-66) with F = I
    and with rotations == the FP64 min-image brute force of the undeformed sum"""
    x, y, w, hw = tg_problem(32, 4, seed=61)
    rsh = O.rsh_of(32, 4)
    rng = np.random.default_rng(62)
    q = rng.random((300, 2)).astype(np.float32)
    d = x[None, :].astype(np.float64) - q[:, :1]
    e = y[None, :].astype(np.float64) - q[:, 1:]
    d -= np.round(d)
    e -= np.round(e)
    Wm = _W64(np.hypot(d, e), hw)
    th = rng.uniform(0, 2 * np.pi, x.size)
    for F in (np.tile(np.array([1, 0, 0, 1], np.float32), (x.size, 1)),
              np.stack([np.cos(th), -np.sin(th), np.sin(th), np.cos(th)], 1)):
        prob = O.Problem(x, y, w, rsh, rsh, F=F)
        for sign, sel in ((1, w > 0), (-1, w < 0)):
            got = prob.eval(q[:, 0], q[:, 1], sign, hw)
            exp = Wm[:, sel] @ np.abs(w[sel].astype(np.float64))
            assert rel_l2(got, exp) < 2e-6, rel_l2(got, exp)


def test_deformed_query_single_blob_closed_form():
    """one particle with F: the field at x_j + F u is |w| W(|u|)"""
    rng = np.random.default_rng(63)
    hw = np.float32(0.125)
    for trial in range(5):
        F = _det1_F(1, rng, 1.9)
        xj = rng.random(2).astype(np.float32)
        prob = O.Problem(xj[:1], xj[1:], np.array([0.8], np.float32), 4, 4, F=F)
        u = rng.uniform(-1, 1, (400, 2)) * float(hw)
        Fm = F.reshape(2, 2).astype(np.float64)
        p = (xj[None, :] + u @ Fm.T) % 1.0
        got = prob.eval(p[:, 0], p[:, 1], 1, hw)
        exp = 0.8 * _W64(np.hypot(u[:, 0], u[:, 1]), float(hw))
        np.testing.assert_allclose(got, exp, rtol=0, atol=2e-5 * exp.max())
