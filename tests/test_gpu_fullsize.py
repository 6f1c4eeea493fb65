"""GPU parity at the workloads' full sizes: the plain GPU path (Taylor-Green,
200 x 200, hwr 16, N1 = N2 = 128) and the large run (double shear layer,
1024 x 1024 = 1,048,576 particles, hwr 8, N1 = N2 = 256).

Hash, subsets and fields are compared whole (bit-exact, or through their
invariants: sortedness, permutation, stable order, CDF end value and
Omega+ + Omega- = sum |omega|); the estimator runs against the full fields
for the first owners and is compared with the oracle on the same Philox
stream (the oracle's per-owner cost keeps the owner count small)."""
import numpy as np
import pytest

import oracle as O
from conftest import rel_l2

pytestmark = pytest.mark.gpu


def _say(msg):
    print("[case]", msg, flush=True)


def _workload(name):
    from vortexparticlemc_cuda_amd.workloads import double_shear
    if name == "tg200":
        Nx, hwr, N1 = 200, 16, 128
        x, y = O.lattice(Nx)
        w = O.tg_vorticity(x, y) / O.tg_lambda11(Nx, hwr)
    else:
        Nx, hwr, N1 = 1024, 8, 256
        x, y = O.lattice(Nx)
        w = double_shear(x, y) / (Nx * Nx)   # strengths ~ h^2 omega (no fit needed here)
    return Nx, hwr, N1, x.astype(np.float32), y.astype(np.float32), w.astype(np.float32)


@pytest.mark.parametrize("name", ["tg200", "shear1024"])
def test_hash_full_size(device, name):
    """FillCells / CountingSort over every particle on the reference grid ==
    the oracle, bit for bit; the reference's invariants hold"""
    import torch

    from vortexparticlemc_cuda_amd import ops
    Nx, hwr, _, x, y, _ = _workload(name)
    rsh = O.rsh_of(Nx, hwr)
    _say("hash full size %s rsh=%d n=%d" % (name, rsh, x.size))
    sub = np.arange(x.size, dtype=np.int32)
    before, acc, vpc, vps, after = O.hash_build(rsh, rsh, 1.0, 1.0, sub, x, y)
    pos = torch.as_tensor(np.stack([x, y], 1), device=device)
    num, acc_g, vpc_g = ops.fill_cells(rsh, rsh, 1.0, 1.0, pos,
                                       torch.as_tensor(sub, device=device))
    assert np.array_equal(num.cpu().numpy(), before)
    assert np.array_equal(acc_g.cpu().numpy(), acc)
    assert np.array_equal(vpc_g.cpu().numpy(), vpc)
    vps_g = ops.counting_sort(num, acc_g, vpc_g).cpu().numpy()
    assert np.array_equal(vps_g, vps)
    assert np.array_equal(np.sort(vps_g), sub)                       # a permutation
    assert (np.diff(vpc[vps_g]) >= 0).all()                          # sorted by cell
    assert num.sum().item() == 0


@pytest.mark.parametrize("name", ["tg200", "shear1024"])
def test_subsets_and_fields_full_size(device, name):
    """subsets, FP64 CDFs and Omega at full size == the oracle; the fine-grid
    fields are cell-sorted, stable and complete"""
    import torch

    from vortexparticlemc_cuda_amd import ops
    Nx, hwr, _, x, y, w = _workload(name)
    _say("subsets/fields full size %s" % name)
    ref = O.build_subsets(w)
    s = ops.build_subsets(torch.as_tensor(w, device=device))
    cnt = s["counts"].cpu().numpy()
    aw = np.abs(w.astype(np.float64))
    tot = s["totals"].cpu().numpy()
    np.testing.assert_allclose(tot.sum(), aw.sum(), rtol=1e-12)
    for k, key in enumerate(("pos", "neg")):
        sub = s["sub_" + key][:cnt[k]].cpu().numpy()
        assert np.array_equal(sub, ref["sub_" + key])
        c = s["cdf_" + key][:cnt[k]].cpu().numpy()
        assert c[-1] == 1.0 and (np.diff(c) >= 0).all()
        np.testing.assert_allclose(c, ref["cdf_" + key], rtol=1e-12, atol=1e-15)
        np.testing.assert_allclose(tot[k], aw[sub].sum(), rtol=1e-12)
        # the sign's fine-grid field
        hw = np.float32(hwr / Nx)
        nx, ny = ops.field_grid(float(hw), 1.0, 1.0, 8)
        fb = ops.FieldBuffers(nx, ny, 1.0, 1.0, int(cnt[k]), device)
        ops.field_build(fb, torch.as_tensor(np.stack([x, y], 1), device=device),
                        torch.as_tensor(w, device=device), sub=s["sub_" + key], abs_w=True)
        cs = fb.cell_start.cpu().numpy().astype(np.int64)
        order = fb.order[:cnt[k]].cpu().numpy()
        pts = fb.pts[:cnt[k]].cpu().numpy()
        assert cs[0] == 0 and cs[-1] == cnt[k] and (np.diff(cs) >= 0).all()
        assert np.array_equal(np.sort(order), np.arange(cnt[k]))
        px, py = x[sub][order], y[sub][order]
        assert np.array_equal(pts[:, 0], px) and np.array_equal(pts[:, 1], py)
        assert np.array_equal(pts[:, 2], np.abs(w[sub][order]))
        kx = np.minimum((px * np.float32(nx)).astype(np.int64), nx - 1)
        ky = np.minimum((py * np.float32(ny)).astype(np.int64), ny - 1)
        key_ = kx * ny + ky
        assert (np.diff(key_) >= 0).all()                          # cell order
        same = np.diff(key_) == 0
        assert (np.diff(order)[same] > 0).all()                    # stable within a cell
        assert np.array_equal(cs[key_] <= np.arange(cnt[k]), np.ones(cnt[k], bool))
        assert (np.arange(cnt[k]) < cs[key_ + 1]).all()


@pytest.mark.parametrize("name", ["tg200", "shear1024"])
def test_estimator_full_size_owners(device, name):
    """consistent estimator on the full fields (Philox seed 0, step 0), the
    first 48 owners: GPU == oracle to 1e-5"""
    from vortexparticlemc_cuda_amd._capi import MODE_CONSISTENT
    from vortexparticlemc_cuda_amd.sim import SimParam, Simulation
    Nx, hwr, N1, x, y, w = _workload(name)
    _say("estimator full size %s" % name)
    P = SimParam(dt=1e-3, resX=Nx, hwr=hwr)
    hw = float(P.h_w)
    sim = Simulation(P, cells_per_hw=8, mode=MODE_CONSISTENT, seed=0)
    sim.set_particles(np.stack([x, y], 1), w)
    sim.rebuild_fields()
    n_own = 48
    d = sim.desc(N1, N1)
    d.n_owner = n_own
    sim.launch_estimate(d)
    vs = sim.vel_sign[:2 * n_own].cpu().numpy().reshape(2, n_own, 2)
    vg = vs[0] - vs[1]
    over = dict(cdf_pos=sim.sub["cdf_pos"][:sim.cnt[0]].cpu().numpy(),
                cdf_neg=sim.sub["cdf_neg"][:sim.cnt[1]].cpu().numpy(),
                cum_pos=sim.sub["totals"][0].item(), cum_neg=sim.sub["totals"][1].item())
    rsh = O.rsh_of(Nx, hwr)
    prob = O.Problem(x, y, w, rsh, rsh, cdf_override=over)
    vo = prob.velocities(prob.params(O.MODE_CONSISTENT, N1, N1, hw, seed=0), owners=n_own)
    assert rel_l2(vg, vo) < 1e-5, rel_l2(vg, vo)
