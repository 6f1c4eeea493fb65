"""The C++ API (the reference's interface, include/*.hpp) driven by a C++
program on the device, its outputs checked against the oracle."""
import os
import subprocess

import numpy as np
import pytest

import oracle as O
from conftest import rel_l2

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cxx", "bin", "vpmc_cxx_check")


def test_cxx_api_end_to_end(device, tmp_path):
    print("[case] cxx api", flush=True)
    if not os.path.exists(BIN):
        pytest.fail("tests/cxx/bin/vpmc_cxx_check not built (run build())")
    Nx, hwr = 32, 4
    x, y = O.lattice(Nx)
    x = x.astype(np.float32)
    y = y.astype(np.float32)
    N = x.size
    b = O.tg_vorticity(x, y).astype(np.float32)
    inp = tmp_path / "in.bin"
    with open(inp, "wb") as f:
        np.array([Nx, hwr, N], np.int32).tofile(f)
        np.stack([x, y], 1).astype(np.float32).tofile(f)
        b.tofile(f)
    r = subprocess.run([BIN, str(inp), str(tmp_path)], capture_output=True, text=True,
                       timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0 and "cxx api ok" in r.stdout
    rd = lambda n, t=np.float32: np.fromfile(tmp_path / n, dtype=t)  # noqa: E731
    hw = hwr / Nx
    # CG strengths: Taylor-Green closed form on the lattice
    w = rd("omega.bin")
    assert rel_l2(w, b / O.tg_lambda11(Nx, hwr)) < 2e-6
    # SpatialHasher == CPU path
    rsh = Nx // hwr // 2
    sub = np.arange(N, dtype=np.int32)
    before, acc, vpc, vps, after = O.hash_build(rsh, rsh, 1.0, 1.0, sub, x, y)
    assert np.array_equal(rd("vp_cell.bin", np.int32)[:N], vpc)
    assert np.array_equal(rd("vp_sort.bin", np.int32)[:N], vps)
    assert np.array_equal(rd("cell_acc.bin", np.int32), acc)
    assert (rd("cell_num_after.bin", np.int32) == 0).all()
    # estimator (consistent mode, Philox seed 7, step 0) == oracle
    prob = O.Problem(x, y, w, rsh, rsh)
    vo = prob.velocities(prob.params(O.MODE_CONSISTENT, 64, 64, hw, seed=7))
    assert rel_l2(rd("vel0.bin").reshape(-1, 2), vo) < 1e-5
    # Ax == FP64 Gram operator
    assert rel_l2(rd("Ax.bin"), O.rbf_matvec(x, y, w, hw)) < 1e-5
    # naive direct sum == FP64 eps-kernel sum (src/cuda/naive_biot_savart.cu)
    un = O.direct_eps(x, y, x, y, w, eps=0.5 * hw, skip_self=True)
    assert rel_l2(rd("vel_naive.bin").reshape(-1, 2), un) < 1e-6
    # two plain steps == oracle steps on the same Philox streams
    px, py = x.copy(), y.copy()
    for step in range(2):
        pr = O.Problem(px, py, w, rsh, rsh)
        P = pr.params(O.MODE_CONSISTENT, 64, 64, hw, seed=7, step=step)
        vel = np.zeros(2 * N, np.float32)
        O.lib().orc_step(O.C.byref(P), O.C.byref(pr.sets[0]), O.C.byref(pr.sets[1]), N, px, py,
                         1e-3, vel, 0)
    d = rd("pos2.bin").reshape(-1, 2) - np.stack([px, py], 1)
    d -= np.rint(d)
    assert np.abs(d).max() < 1e-6
    # the driver's output files of that state (include/viz_utils.hpp) == the
    # reference's per-pixel formula in FP64 and its arrow geometry
    from test_viz import _ref_image
    from vortexparticlemc_cuda_amd import viz
    p2 = rd("pos2.bin").reshape(-1, 2)[:N].astype(np.float64)
    for name, per in (("vort_per.png", True), ("vort.png", False)):
        img = viz.decode_png((tmp_path / name).read_bytes())
        ref = _ref_image(p2[:, 0], p2[:, 1], w.astype(np.float64), hw, 48, periodic=per)
        diff = np.abs(img.astype(int) - ref.astype(int))
        assert img.shape == (48, 48) and diff.max() <= 1 and (diff > 0).mean() < 0.01, name
    # monte_carlo_bs.cuh's device helpers: the header's query agrees with the
    # library's (separate FP32 code of one sum: rounding only), the math with numpy
    hv = rd("cuh_helpers.bin")
    h, hq = hv[:160].reshape(16, 10), hv[160:]
    pts = rd("cuh_points.bin").reshape(-1, 2).astype(np.float64)
    assert np.allclose(h[:, 0], hq, rtol=1e-6, atol=1e-9) and (hq > 0).any()
    F = np.array([[1.2, 0.3], [-0.1, 0.9]])
    assert np.allclose(h[:, 1:3], pts @ F.T, atol=1e-6)
    dmi = pts[:, 0] - 0.9
    assert np.allclose(h[:, 3], dmi - np.rint(dmi), atol=1e-6)
    assert np.allclose(h[:, 4:8], np.tile(F.ravel() / np.sqrt(np.linalg.det(F)), (16, 1)),
                       atol=1e-6)
    d = pts - 0.5
    c = 1.0 / (2 * np.pi * ((d ** 2).sum(1) + 1e-4))
    assert np.allclose(h[:, 8:10], np.stack([-d[:, 1] * c, d[:, 0] * c], 1), rtol=1e-5)
    svg = (tmp_path / "svg" / "vel.svg").read_text().splitlines()
    ref_svg = viz.velocity_svg(rd("pos2.bin").reshape(-1, 2)[:N],
                               rd("vel2.bin").reshape(-1, 2)[:N], 0.2, True).splitlines()
    assert len(svg) == len(ref_svg) and svg[:4] == ref_svg[:4]
    num = lambda ln: [float(t) for t in ln.split('"')[3:11:2]]  # noqa: E731
    for a, b in zip(svg[4:-1], ref_svg[4:-1]):
        assert np.allclose(num(a), num(b), atol=2e-6), (a, b)
    # kernel_deformation: step_cuda with aux targets and F == the Python mirror
    # (same library entry point, same Philox streams) bit for bit
    from vortexparticlemc_cuda_amd._capi import MODE_CONSISTENT
    from vortexparticlemc_cuda_amd.sim import SimParam, Simulation
    sd = Simulation(SimParam(dt=1e-3, resX=Nx, hwr=hwr), device=device, cells_per_hw=8,
                    mode=MODE_CONSISTENT, seed=7, deform=True)
    sd.set_particles(np.stack([x, y], 1), w)
    sd.step(64, 64)
    sd.step(64, 64)
    assert np.array_equal(rd("pos_def2.bin").reshape(-1, 2), sd.pos.cpu().numpy())
    F = rd("F_def2.bin").reshape(-1, 4)[:N]
    assert np.array_equal(F, sd.jac[:N].cpu().numpy())
    assert not np.array_equal(F, np.tile([1, 0, 0, 1], (N, 1)).astype(np.float32))
    # Simulation::eval_vorticity_cpu == the oracle's w+ - w- query on the
    # reference grid (src/cpu/evaluate_vorticity.cpp:8-75) at the two-step state
    p2f = rd("pos2.bin").reshape(-1, 2)[:N]
    pr2 = O.Problem(p2f[:, 0].copy(), p2f[:, 1].copy(), w, rsh, rsh)
    ev = rd("eval_vort.bin")
    ep = pr2.eval(pts[:, 0], pts[:, 1], +1, hw) - pr2.eval(pts[:, 0], pts[:, 1], -1, hw)
    assert rel_l2(ev, ep) < 1e-5 and np.abs(ep).max() > 0
    # a bounded box (step_cuda(false), estimate_velocity(false): the direct
    # estimator, no wrap) == the oracle with periodic = 0, both modes
    for md, mode in ((0, O.MODE_REFERENCE), (1, O.MODE_CONSISTENT)):
        prn = O.Problem(x, y, w, rsh, rsh)
        Pn = prn.params(mode, 64, 64, hw, periodic=False, seed=7)
        vn = prn.velocities(Pn)
        assert rel_l2(rd("vel_np%d.bin" % md).reshape(-1, 2)[:N], vn) < 1e-5, md
        assert np.abs(vn).max() > 0
        pn = rd("pos_np%d.bin" % md).reshape(-1, 2)[:N].astype(np.float64)
        pe = np.stack([x, y], 1) + vn.astype(np.float64) * 1e-3
        assert np.abs(pn - pe).max() < 1e-6, md
    # copy-constructed and copy-assigned simulations step like the original
    assert list(rd("copy_ok.bin", np.int32)) == [1, 1]
    # reinitialize() == the Python mirror's reinit (same entry point, same state)
    sd.reinit(tol=1e-6)
    assert np.array_equal(rd("pos_reinit.bin").reshape(-1, 2)[:sd.N], sd.pos.cpu().numpy())
    assert np.array_equal(rd("omega_reinit.bin")[:sd.N], sd.omega.cpu().numpy())


def test_cxx_cpu_backend_entry_points_replay_the_reference(device, tmp_path):
    """monte_carlo_bs_cpu, the std::mt19937 samplers, evalVorticityAbs_cpu and
    step_cpu through the C++ API: the same generator draws as the reference's
    src/cpu code, the sums on the GPU.  Compared with the reference's own
    outputs (tests/golden, written from oracle/_ref: ref_mc_aux,
    ref_mc_velocities with aux, ref_eval_vorticity) on mt19937(1234)."""
    print("[case] cxx cpu-backend entry points", flush=True)
    if not os.path.exists(BIN):
        pytest.fail("tests/cxx/bin/vpmc_cxx_check not built (run build())")
    G = np.load(os.path.join(ROOT, "tests", "golden", "ref_cpu_vectors.npz"))
    x, y, w = G["px"], G["py"], G["w"]
    N1, N2, n_own = int(G["mc_N1"]), int(G["mc_N2"]), int(G["mc_owners"])
    Nx, hwr, seed = 32, 4, 1234
    assert float(G["h_w"]) == hwr / Nx
    cp, cn, cum = G["cdf_pos_f32"], G["cdf_neg_f32"], G["cum_f32"]
    q = G["q"]
    inp = tmp_path / "mcbs.bin"
    with open(inp, "wb") as f:
        np.array([Nx, hwr, x.size, N1, N2, n_own, seed], np.int32).tofile(f)
        np.array([G["mcaux_dx"]], np.float32).tofile(f)
        for a in (x, y, w):
            a.astype(np.float32).tofile(f)
        np.array([cp.size, cn.size], np.int32).tofile(f)
        for a in (cp, cn, cum):
            a.astype(np.float32).tofile(f)
        np.array([len(q)], np.int32).tofile(f)
        q.astype(np.float32).tofile(f)
    r = subprocess.run([BIN, str(inp), str(tmp_path), "mcbs"], capture_output=True, text=True,
                       timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0 and "cxx mcbs ok" in r.stdout
    rd = lambda n, t=np.float32: np.fromfile(tmp_path / n, dtype=t)  # noqa: E731
    # five estimates per owner and sign == the reference's monte_carlo_bs_cpu
    est = rd("mcbs_est.bin").reshape(n_own, 2, 5, 2)
    ref = G["mcaux_est"]
    for si in range(2):
        for t in range(5):
            e = rel_l2(est[:, si, t], ref[:, si, t])
            assert e < 1e-5, (si, t, e)
    assert rd("mcbs_rng.bin", np.int32)[0] == 1   # exactly the reference's draws taken
    # samplers on the generator's first draws (mt19937(1234), as golden mc_aux0_U)
    U = G["mc_aux0_U"].astype(np.float64)
    smp = rd("samplers.bin")
    assert int(smp[0]) == min(int(np.searchsorted(cp, np.float32(U[0]), side="right")),
                              cp.size - 1)
    phi, rr = 2 * np.pi * U[1], 0.7 * np.sqrt(U[2])
    assert np.allclose(smp[1:3], [0.3 + rr * np.cos(phi), 0.6 + rr * np.sin(phi)], atol=2e-6)
    pos4 = U[3] * 4
    i0 = int(pos4)
    tab = [0.0, 0.3, 0.5, 0.7, 1.0]
    kr = (tab[i0] * (1 - (pos4 - i0)) + tab[min(i0 + 1, 4)] * (pos4 - i0)) * 0.25
    assert abs(smp[3] - kr) < 1e-6
    # evalVorticityAbs_cpu == the reference's query at the golden points
    qv = rd("evalq.bin").reshape(2, -1)
    assert rel_l2(qv[0], G["query_+1"]) < 1e-5 and rel_l2(qv[1], G["query_-1"]) < 1e-5
    # step_cpu: particle velocities of the first owners == the reference's step
    # velocities on mt19937(1234) with aux draws (its FP32 CDFs pick the same
    # particles as the simulation's FP64 ones for these draws)
    v = rd("step_vel.bin").reshape(-1, 2)
    assert rel_l2(v[:n_own], G["mc_aux1_vel"]) < 1e-5, rel_l2(v[:n_own], G["mc_aux1_vel"])
    p1 = rd("step_pos.bin").reshape(-1, 2)[:x.size]
    moved = np.stack([x, y], 1).astype(np.float64) + 1e-3 * v[:x.size].astype(np.float64)
    d = p1 - moved
    d -= np.rint(d)
    assert np.abs(d).max() < 1e-6 and (p1 >= 0).all() and (p1 < 1).all()
    assert np.abs(rd("step_left.bin").reshape(-1, 2)[:x.size]).max() > 0   # aux points moved
    assert rd("step_rng_member.bin", np.int32)[0] == 1   # the member generator is not advanced
    # step_cpu over two replay chunks == monte_carlo_bs_cpu per particle (particles
    # 0..699 cross the first chunk's end at 671), bit for bit
    vc, vl = rd("step_chunks_vel.bin"), rd("step_chunks_loop.bin")
    assert vc.size == 1400 and np.array_equal(vc, vl) and np.abs(vc).max() > 0


def test_cxx_host_span_hash_matches_reference(tmp_path):
    """FillCells_cpu/_omp + CountingSort_cpu/_omp (host spans, run on the GPU)
    == the reference's CPU hash on its golden cases, bit for bit (the OMP
    variant in the CPU order: the reference's OMP sort writes from a wrong
    base, src/omp/hash_omp.cpp:65)"""
    print("[case] cxx host-span hash", flush=True)
    if not os.path.exists(BIN):
        pytest.fail("tests/cxx/bin/vpmc_cxx_check not built (run build())")
    G = np.load(os.path.join(ROOT, "tests", "golden", "ref_cpu_vectors.npz"))
    tags = ("a", "b", "c")
    inp = tmp_path / "hash.bin"
    with open(inp, "wb") as f:
        np.array([len(tags)], np.int32).tofile(f)
        for t in tags:
            px, py = G["hash_%s_px" % t], G["hash_%s_py" % t]
            np.array([px.size, int(G["hash_%s_rsh" % t])], np.int32).tofile(f)
            px.astype(np.float32).tofile(f)
            py.astype(np.float32).tofile(f)
    r = subprocess.run([BIN, str(inp), str(tmp_path), "hash"], capture_output=True, text=True,
                       timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0 and "cxx hash ok" in r.stdout
    for c, t in enumerate(tags):
        n, cells = G["hash_%s_px" % t].size, int(G["hash_%s_rsh" % t]) ** 2
        for b in ("cpu", "omp"):
            res = np.fromfile(tmp_path / ("hash_%s_%d.bin" % (b, c)), dtype=np.int32)
            num, acc = res[:cells], res[cells:2 * cells]
            cell, mp = res[2 * cells:2 * cells + n], res[2 * cells + n:2 * cells + 2 * n]
            after = res[2 * cells + 2 * n:]
            assert np.array_equal(num, G["hash_%s_num" % t]), (t, b)
            assert np.array_equal(acc, G["hash_%s_acc" % t]), (t, b)
            assert np.array_equal(cell, G["hash_%s_cell" % t]), (t, b)
            assert np.array_equal(mp, G["hash_%s_map" % t]), (t, b)
            assert (after == 0).all(), (t, b)
