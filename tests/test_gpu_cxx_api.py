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
    # reinitialize() == the Python mirror's reinit (same entry point, same state)
    sd.reinit(tol=1e-6)
    assert np.array_equal(rd("pos_reinit.bin").reshape(-1, 2)[:sd.N], sd.pos.cpu().numpy())
    assert np.array_equal(rd("omega_reinit.bin")[:sd.N], sd.omega.cpu().numpy())
