"""Writes tests/golden/ref_cpu_vectors.npz from the reference's own CPU path.

Uses oracle/_ref/libvpmc_ref.so (the reference's src/cpu sources compiled by
oracle/build.py in this container) to record inputs and outputs:
  - FillCells_cpu / CountingSort_cpu on random points (the shape of
    apps/test_spatial_hashing.cu:66-75, smaller N);
  - evalVorticityAbs_cpu at random points for both signs;
  - monte_carlo_bs_cpu velocities (sign +1 then -1 per owner, one
    std::mt19937 as src/cpu/simulation_step.cpp:28-44 uses it) together with
    the draws that generator produced, for aux = 0 and aux = 1;
  - the five monte_carlo_bs_cpu estimates (owner and the four auxiliary
    targets at dst -/+ dx) per owner and sign, with their draws.
Run: python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import oracle as O  # noqa: E402


def main():
    R = O.ref_lib()
    if R is None:
        raise SystemExit("oracle/_ref not built (needs the reference sources)")
    out = {}
    # ---- hash
    for tag, n, rsh, seed in (("a", 2000, 32, 123), ("b", 777, 8, 5), ("c", 300, 3, 9)):
        rng = np.random.default_rng(seed)
        px = rng.random(n, dtype=np.float32)
        py = rng.random(n, dtype=np.float32)
        px[::50] = 1.0
        sub = np.arange(n, dtype=np.int32)
        num = np.zeros(rsh * rsh, np.int32)
        acc = np.zeros(rsh * rsh, np.int32)
        vpc = np.zeros(n, np.int32)
        vpm = np.zeros(n, np.int32)
        R.ref_fill_cells(rsh, rsh, 1.0, 1.0, n, sub, n, px, py, num, acc, vpc)
        hist = num.copy()
        R.ref_counting_sort(n, rsh * rsh, num, acc, vpc, vpm)
        out.update({"hash_%s_px" % tag: px, "hash_%s_py" % tag: py,
                    "hash_%s_rsh" % tag: np.int32(rsh), "hash_%s_num" % tag: hist,
                    "hash_%s_acc" % tag: acc, "hash_%s_cell" % tag: vpc,
                    "hash_%s_map" % tag: vpm})
    # ---- particles: TG strengths on a jittered 32x32 lattice, hwr 4
    Nx, hwr = 32, 4
    h = 1.0 / Nx
    x, y = O.lattice(Nx)
    rng = np.random.default_rng(0)
    x = ((x + 0.3 * h * rng.standard_normal(x.size)) % 1.0).astype(np.float32)
    y = ((y + 0.3 * h * rng.standard_normal(y.size)) % 1.0).astype(np.float32)
    w = (O.tg_vorticity(x, y) / O.tg_lambda11(Nx, hwr)).astype(np.float32)
    hw = np.float32(hwr * h)
    rsh = O.rsh_of(Nx, hwr)
    out.update(px=x, py=y, w=w, h_w=hw, rsh=np.int32(rsh))
    q = (rng.random((400, 2)) * 1.4 - 0.2).astype(np.float32)
    out["q"] = q
    for sign in (1, -1):
        v = np.zeros(len(q), np.float32)
        R.ref_eval_vorticity(x.size, x, y, w, rsh, rsh, 1.0, 1.0, len(q),
                             np.ascontiguousarray(q[:, 0]), np.ascontiguousarray(q[:, 1]), sign,
                             0, 0, hw, 1, v)
        out["query_%+d" % sign] = v
    cnt = np.zeros(2, np.int32)
    cp = np.zeros(x.size, np.float32)
    cn = np.zeros(x.size, np.float32)
    cum = np.zeros(2, np.float32)
    R.ref_subsets_f32(x.size, w, cnt, cp, cn, cum)
    out.update(cdf_pos_f32=cp[:cnt[0]], cdf_neg_f32=cn[:cnt[1]], cum_f32=cum)
    # ---- MC velocities with the generator's draws
    N1 = N2 = 16
    n_own = 40
    seed = 1234
    for aux in (0, 1):
        vel = np.zeros(2 * n_own, np.float32)
        R.ref_mc_velocities(x.size, x, y, w, rsh, rsh, 1.0, 1.0, hw, 1, N1, N2, aux, seed, 0,
                            n_own, vel)
        count = n_own * 2 * (N1 * 3 + N2 * (10 if aux else 2))
        U = np.zeros(count, np.float32)
        R.ref_uniform_stream(seed, count, U)
        out["mc_aux%d_vel" % aux] = vel.reshape(-1, 2)
        out["mc_aux%d_U" % aux] = U
    out.update(mc_N1=np.int32(N1), mc_N2=np.int32(N2), mc_owners=np.int32(n_own))
    # ---- the five estimates with real auxiliary targets (dst -/+ dx, the CUDA
    # step's placement), per owner and sign, and the draws
    dx = np.float32(0.5 * hw)
    est = np.zeros(n_own * 2 * 10, np.float32)
    R.ref_mc_aux(x.size, x, y, w, rsh, rsh, 1.0, 1.0, hw, 1, N1, N2, dx, seed, 0, n_own, est)
    count = n_own * 2 * (N1 * 3 + N2 * 10)
    U = np.zeros(count, np.float32)
    R.ref_uniform_stream(seed, count, U)
    out.update(mcaux_est=est.reshape(n_own, 2, 5, 2), mcaux_U=U, mcaux_dx=dx)
    path = os.path.join(HERE, "ref_cpu_vectors.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main()
