"""Synthetic workloads (no datasets: every field is analytic).

taylor_green        omega = 4 pi sin(2 pi x) sin(2 pi y)  (include/analytical_flow.hpp:24-31)
gaussian_ensemble   independent periodic simulations on [0,1]^2, each with
                    Nx drawn from {64, 96, 128, 192, 256} and vorticity the sum
                    of 8 Gaussian vortices: centres U[0,1]^2, peak N(0, (4 pi)^2),
                    core sigma U(0.03, 0.10) in peak * exp(-d^2 / (2 sigma^2)),
                    d the minimum-image distance
double_shear        u = tanh((y - 1/4)/rho) (y <= 1/2), tanh((3/4 - y)/rho)
                    (y > 1/2), v = delta sin(2 pi x); omega = dv/dx - du/dy
"""
import numpy as np

from .sim import lattice_positions

ENSEMBLE_SIZES = (64, 96, 128, 192, 256)


def taylor_green(x, y):
    x = np.asarray(x, np.float64)
    y = np.asarray(y, np.float64)
    return 4 * np.pi * np.sin(2 * np.pi * x) * np.sin(2 * np.pi * y)


def gaussian_vortices(x, y, centres, peaks, sigmas, L=1.0):
    x = np.asarray(x, np.float64)
    y = np.asarray(y, np.float64)
    out = np.zeros_like(x)
    for (cx, cy), a, s in zip(centres, peaks, sigmas):
        dx = x - cx
        dy = y - cy
        dx -= L * np.rint(dx / L)
        dy -= L * np.rint(dy / L)
        out += a * np.exp(-(dx * dx + dy * dy) / (2 * s * s))
    return out


def gaussian_ensemble(n_sim=64, seed=20261203, sizes=ENSEMBLE_SIZES, n_vortex=8):
    """per simulation: (Nx, lattice positions (N, 2) float32, vorticity at
    the particles (N,) float64, the vortex parameters)"""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n_sim):
        nx = int(rng.choice(sizes))
        centres = rng.random((n_vortex, 2))
        peaks = rng.normal(0.0, 4 * np.pi, n_vortex)
        sigmas = rng.uniform(0.03, 0.10, n_vortex)
        pos = lattice_positions(nx)
        om = gaussian_vortices(pos[:, 0], pos[:, 1], centres, peaks, sigmas)
        out.append((nx, pos, om, (centres, peaks, sigmas)))
    return out


def double_shear(x, y, rho=1.0 / 30.0, delta=0.05):
    x = np.asarray(x, np.float64)
    y = np.asarray(y, np.float64)
    lower = y <= 0.5
    s1 = 1.0 / np.cosh((y - 0.25) / rho) ** 2
    s2 = 1.0 / np.cosh((0.75 - y) / rho) ** 2
    du_dy = np.where(lower, s1 / rho, -s2 / rho)
    dv_dx = 2 * np.pi * delta * np.cos(2 * np.pi * x)
    return dv_dx - du_dy
