"""Output formats (vortexparticlemc_cuda_amd/viz.py): the PNG writer, the SVG
arrows (include/viz_utils.hpp:84-132) and, on the GPU, the vorticity image
against the reference's per-pixel formula (viz_utils.hpp:26-82) evaluated in
FP64 with numpy."""
import numpy as np
import pytest

from conftest import tg_problem


def test_png_round_trip():
    from vortexparticlemc_cuda_amd import viz
    rng = np.random.default_rng(0)
    for h, w in ((1, 1), (7, 300), (64, 64)):
        img = rng.integers(0, 256, size=(h, w)).astype(np.uint8)
        data = viz.encode_png(img)
        assert data[:8] == b"\x89PNG\r\n\x1a\n" and data[12:16] == b"IHDR"
        assert np.array_equal(viz.decode_png(data), img)


def test_svg_arrows_follow_the_reference_geometry():
    from vortexparticlemc_cuda_amd import viz
    pos = np.array([[0.5, 0.5], [0.99, 0.1], [0.2, 0.3]], np.float32)
    vel = np.array([[1.0, 0.0], [0.0, 2.0], [0.0, 0.0]], np.float32)
    text = viz.velocity_svg(pos, vel, vel_scale=0.25, periodic=True)
    lines = [ln for ln in text.splitlines() if ln.startswith("<line")]
    assert len(lines) == 2   # zero velocity: no arrow
    # centred on the particle, length 0.25 |v|, y flipped, ends wrapped
    assert lines[0] == '<line class="arrow" x1="0.375" y1="0.5" x2="0.625" y2="0.5"/>'
    assert lines[1] == '<line class="arrow" x1="0.99" y1="0.15" x2="0.99" y2="0.65"/>'
    wrapped = viz.velocity_svg(np.array([[0.95, 0.5]], np.float32),
                               np.array([[1.0, 0.0]], np.float32), vel_scale=0.2)
    x1, x2 = [float(v) for v in wrapped.splitlines()[-2].split('"')[3:8:4]]
    assert abs(x1 - 0.85) < 1e-6 and abs(x2 - 0.05) < 1e-6


def _ref_image(x, y, w, hw, res, periodic=True):
    """viz_utils.hpp:42-80 in FP64: |w|-weighted spline sum, min image on the
    unit box when periodic, r^2 < 1e-12 skipped, 128 + int(value * 127 / max)"""
    sig2 = 40.0 / (7.0 * np.pi * hw * hw)
    c = (np.arange(res) + 0.5) / res
    X, Y = np.meshgrid(c, c, indexing="xy")          # [j, i]: y row, x column
    val = np.zeros_like(X)
    for xi, yi, wi in zip(x, y, np.abs(w)):
        dx = xi - X
        dy = yi - Y
        if periodic:
            dx -= np.rint(dx)
            dy -= np.rint(dy)
        r2 = dx * dx + dy * dy
        q = np.sqrt(r2) / hw
        k = np.where(q <= 0.5, 6 * (q ** 3 - q ** 2) + 1, np.where(q < 1, 2 * (1 - q) ** 3, 0.0))
        k[(r2 > hw * hw) | (r2 < 1e-12)] = 0.0
        val += wi * sig2 * k
    inv = 127.0 / np.abs(val).max()
    return np.clip(128 + np.trunc(val * inv), 0, 255).astype(np.uint8)


@pytest.mark.gpu
def test_vorticity_png_matches_reference_formula(device, tmp_path):
    from vortexparticlemc_cuda_amd import viz
    from vortexparticlemc_cuda_amd.sim import SimParam, Simulation
    x, y, w, hw = tg_problem(32, 4, jitter=0.3, seed=5)
    P = SimParam(dt=1e-3, resX=32, hwr=4)
    sim = Simulation(P, device=device)
    sim.set_particles(np.stack([x, y], 1), w)
    res = 64
    img = viz.dump_vorticity_png(sim, str(tmp_path / "w.png"), res=res)
    back = viz.decode_png((tmp_path / "w.png").read_bytes())
    assert np.array_equal(back, img) and img.shape == (res, res)
    ref = _ref_image(x.astype(np.float64), y.astype(np.float64), w.astype(np.float64), float(hw),
                     res)
    diff = np.abs(img.astype(int) - ref.astype(int))
    # FP32 sums in a different order may cross a quantisation boundary
    assert diff.max() <= 1 and (diff > 0).mean() < 0.01
