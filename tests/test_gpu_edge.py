"""Edge cases of the estimator on the GPU, each against the oracle on the same
Philox stream (consistent and reference modes, binned and per-owner
algorithms): an empty sign set, N1 = 0 (disk samples only), N2 = 0 (particle
samples only), a single particle, and particles on the box edges."""
import numpy as np
import pytest

import oracle as O
from conftest import rel_l2, tg_problem

pytestmark = pytest.mark.gpu


def _case(name):
    x, y, w, hw = tg_problem(24, 4, seed=81)
    N1 = N2 = 24
    if name == "empty_neg":
        w = np.abs(w) + np.float32(1e-3)
    elif name == "n1_zero":
        N1 = 0
    elif name == "n2_zero":
        N2 = 0
    elif name == "single":
        x, y, w = x[:1] * 0 + 0.3, y[:1] * 0 + 0.7, np.array([0.05], np.float32)
    elif name == "edges":
        x = x.copy()
        y = y.copy()
        x[:24] = 0.0                               # on the lower edge
        y[24:48] = np.float32(1 - 2 ** -24)        # just below the upper edge
    return x.astype(np.float32), y.astype(np.float32), w.astype(np.float32), hw, N1, N2


@pytest.mark.parametrize("algo", [0, 1])
@pytest.mark.parametrize("mode", [O.MODE_REFERENCE, O.MODE_CONSISTENT])
@pytest.mark.parametrize("name", ["empty_neg", "n1_zero", "n2_zero", "single", "edges"])
def test_estimator_edge_cases_match_oracle(device, name, mode, algo):
    from vortexparticlemc_cuda_amd.sim import SimParam, Simulation
    print("[case] edge %s mode=%d algo=%d" % (name, mode, algo), flush=True)
    x, y, w, hw, N1, N2 = _case(name)
    P = SimParam(dt=1e-3, resX=24, hwr=4)
    sim = Simulation(P, device=device, cells_per_hw=8, mode=mode, seed=99, algo=algo)
    sim.set_particles(np.stack([x, y], 1), w)
    sim.step_index = 2
    vg = sim.estimate(N1, N2).cpu().numpy()
    over = dict(cdf_pos=sim.sub["cdf_pos"][:sim.cnt[0]].cpu().numpy(),
                cdf_neg=sim.sub["cdf_neg"][:sim.cnt[1]].cpu().numpy(),
                cum_pos=sim.sub["totals"][0].item(), cum_neg=sim.sub["totals"][1].item())
    rsh = O.rsh_of(24, 4)
    prob = O.Problem(x, y, w, rsh, rsh, cdf_override=over)
    vo = prob.velocities(prob.params(mode, N1, N2, hw, seed=99, step=2))
    assert np.isfinite(vg).all()
    if np.abs(vo).max() == 0.0:
        assert np.abs(vg).max() == 0.0
    else:
        assert rel_l2(vg, vo) < 1e-5, rel_l2(vg, vo)
    if name == "empty_neg":
        assert sim.cnt[1] == 0
