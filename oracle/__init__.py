"""Python access to the oracle (the CPU restatement) and to oracle/_ref.

TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg may import this package; the product package
vortexparticlemc_cuda_amd never does.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_lib = None
_ref = None

f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")

MODE_REFERENCE = 0
MODE_CONSISTENT = 1
KPER_TERMS = 24


class View(C.Structure):
    _fields_ = [("px", C.c_void_p), ("py", C.c_void_p), ("w", C.c_void_p), ("N", C.c_int),
                ("vp_sort", C.c_void_p), ("cell_acc", C.c_void_p), ("rshx", C.c_int),
                ("rshy", C.c_int), ("Lx", C.c_float), ("Ly", C.c_float),
                ("cnt_view", C.c_int), ("sub", C.c_void_p), ("F", C.c_void_p)]


class McParams(C.Structure):
    _fields_ = [("mode", C.c_int), ("N1", C.c_int), ("N2", C.c_int), ("h_w", C.c_float),
                ("Lx", C.c_float), ("Ly", C.c_float), ("periodic", C.c_int),
                ("seed", C.c_uint64), ("step", C.c_uint32), ("sim", C.c_uint32),
                ("U", C.c_void_p), ("nterms", C.c_int), ("kper_c", C.c_void_p),
                ("aux_draws", C.c_int)]


class SignSet(C.Structure):
    _fields_ = [("n", C.c_int), ("sub", C.c_void_p), ("cdf", C.c_void_p),
                ("omega", C.c_double), ("view", View)]


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            from . import build as _b
            _b.build_oracle()
        L = C.CDLL(path)
        L.orc_spline_f.restype = C.c_float
        L.orc_spline_f.argtypes = [C.c_float, C.c_float]
        L.orc_spline_d.restype = C.c_double
        L.orc_spline_d.argtypes = [C.c_double, C.c_double]
        L.orc_mass_fraction.restype = C.c_double
        L.orc_mass_fraction.argtypes = [C.c_double]
        L.orc_mass_inverse.restype = C.c_double
        L.orc_mass_inverse.argtypes = [C.c_double]
        L.orc_fill_cells.argtypes = [C.c_int, C.c_int, C.c_float, C.c_float, C.c_int, i32p,
                                     f32p, f32p, i32p, i32p, i32p]
        L.orc_counting_sort.argtypes = [C.c_int, i32p, i32p, i32p, i32p]
        L.orc_eval_vorticity.restype = C.c_float
        L.orc_eval_vorticity.argtypes = [C.POINTER(View), C.c_float, C.c_float, C.c_int,
                                         C.c_int, C.c_int, C.c_float, C.c_int]
        L.orc_build_subsets.argtypes = [C.c_int, f32p, i32p, i32p, f64p, f64p, i32p, f64p]
        L.orc_sample_idx.restype = C.c_int
        L.orc_sample_idx.argtypes = [C.c_float, f64p, C.c_int]
        L.orc_philox4x32_10.argtypes = [u32p, u32p, u32p]
        L.orc_u01.restype = C.c_float
        L.orc_u01.argtypes = [C.c_uint32]
        L.orc_kper_coeffs.argtypes = [C.c_int, f64p]
        L.orc_kernel_smooth_d.argtypes = [C.c_double, C.c_double, C.c_double, C.c_int, f64p,
                                          C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.orc_mc_velocity.argtypes = [C.POINTER(McParams), C.POINTER(SignSet), C.c_int,
                                      C.c_int, C.c_float, C.c_float, f32p, f32p, f32p]
        L.orc_step.argtypes = [C.POINTER(McParams), C.POINTER(SignSet), C.POINTER(SignSet),
                               C.c_int, f32p, f32p, C.c_float, f32p, C.c_int]
        L.orc_direct_eps.argtypes = [C.c_int, f64p, f64p, C.c_int, C.c_int, f64p, f64p, f64p,
                                     C.c_double, C.c_int, C.c_double, f64p, C.c_int]
        L.orc_direct_blob.argtypes = [C.c_int, f64p, f64p, C.c_int, f64p, f64p, f64p,
                                      C.c_double, C.c_double, C.c_int, f64p, f64p, C.c_int]
        L.orc_rbf_matvec.argtypes = [C.c_int, f64p, f64p, C.c_double, C.c_double, C.c_int,
                                     f64p, f64p, C.c_int]
        L.orc_cg_rbf.restype = C.c_int
        L.orc_cg_rbf.argtypes = [C.c_int, f64p, f64p, C.c_double, C.c_double, C.c_int, f64p,
                                 f64p, C.c_double, C.c_int, C.POINTER(C.c_double), C.c_int]
        L.orc_tg_vorticity.restype = C.c_double
        L.orc_tg_vorticity.argtypes = [C.c_double, C.c_double]
        L.orc_shear_vorticity.restype = C.c_double
        L.orc_shear_vorticity.argtypes = [C.c_double, C.c_double, C.c_double, C.c_double]
        _lib = L
    return _lib


def ref_lib():
    """oracle/_ref/libvpmc_ref.so (the reference's CPU sources), or None."""
    global _ref
    if _ref is None:
        path = os.path.join(HERE, "_ref", "libvpmc_ref.so")
        if not os.path.exists(path):
            return None
        R = C.CDLL(path)
        R.ref_spline.restype = C.c_float
        R.ref_spline.argtypes = [C.c_float, C.c_float]
        R.ref_fill_cells.argtypes = [C.c_int, C.c_int, C.c_float, C.c_float, C.c_int, i32p,
                                     C.c_int, f32p, f32p, i32p, i32p, i32p]
        R.ref_fill_cells_omp.argtypes = R.ref_fill_cells.argtypes
        R.ref_counting_sort.argtypes = [C.c_int, C.c_int, i32p, i32p, i32p, i32p]
        R.ref_eval_vorticity.argtypes = [C.c_int, f32p, f32p, f32p, C.c_int, C.c_int,
                                         C.c_float, C.c_float, C.c_int, f32p, f32p, C.c_int,
                                         C.c_int, C.c_int, C.c_float, C.c_int, f32p]
        R.ref_subsets_f32.argtypes = [C.c_int, f32p, i32p, f32p, f32p, f32p]
        R.ref_mc_velocities.argtypes = [C.c_int, f32p, f32p, f32p, C.c_int, C.c_int, C.c_float,
                                        C.c_float, C.c_float, C.c_int, C.c_int, C.c_int,
                                        C.c_int, C.c_uint32, C.c_int, C.c_int, f32p]
        R.ref_mc_velocities_mt.argtypes = [C.c_int, f32p, f32p, f32p, C.c_int, C.c_int,
                                           C.c_float, C.c_float, C.c_float, C.c_int, C.c_int,
                                           C.c_int, C.c_uint32, C.c_int, C.c_int, C.c_int, f32p]
        R.ref_mc_aux.argtypes = [C.c_int, f32p, f32p, f32p, C.c_int, C.c_int, C.c_float,
                                 C.c_float, C.c_float, C.c_int, C.c_int, C.c_int, C.c_float,
                                 C.c_uint32, C.c_int, C.c_int, f32p]
        R.ref_uniform_stream.argtypes = [C.c_uint32, C.c_int64, f32p]
        _ref = R
    return _ref


# --------------------------------------------------------------------------
# helpers

def rsh_of(Nx, hwr):
    """src/vortex_particle_mc.cu:125: rsh = ResolutionX / int(hwr) / 2"""
    return int(Nx) // int(hwr) // 2


def lattice(Nx, Ny=None, L=1.0):
    """cell-centred lattice, particle index i*Ny + j (x-major)"""
    Ny = Nx if Ny is None else Ny
    h = L / Nx
    i, j = np.meshgrid(np.arange(Nx), np.arange(Ny), indexing="ij")
    x = ((i + 0.5) * h).ravel()
    y = ((j + 0.5) * (L / Ny)).ravel()
    return x, y


def tg_vorticity(x, y):
    return 4.0 * np.pi * np.sin(2 * np.pi * x) * np.sin(2 * np.pi * y)


def tg_velocity(x, y):
    return np.stack([np.sin(2 * np.pi * x) * np.cos(2 * np.pi * y),
                     -np.cos(2 * np.pi * x) * np.sin(2 * np.pi * y)], axis=-1)


def shear_vorticity(x, y, rho=1.0 / 30.0, delta=0.05):
    base = 2 * np.pi * delta * np.cos(2 * np.pi * x)
    lo = base - 1.0 / (rho * np.cosh((y - 0.25) / rho) ** 2)
    hi = base + 1.0 / (rho * np.cosh((0.75 - y) / rho) ** 2)
    return np.where(y <= 0.5, lo, hi)


def gaussian_vortices(x, y, centres, peaks, sigmas, L=1.0):
    """sum of Gaussian vortices peak*exp(-d^2/(2 sigma^2)), d min-image"""
    out = np.zeros_like(x, dtype=np.float64)
    for (cx, cy), a, s in zip(centres, peaks, sigmas):
        dx = x - cx
        dy = y - cy
        dx -= L * np.rint(dx / L)
        dy -= L * np.rint(dy / L)
        out += a * np.exp(-(dx * dx + dy * dy) / (2 * s * s))
    return out


def spline_f(r, h_w):
    L = lib()
    return np.array([L.orc_spline_f(float(v), float(h_w)) for v in np.ravel(r)], np.float32)


def hash_build(rshx, rshy, Lx, Ly, sub, px, py):
    """FillCells_cpu + CountingSort_cpu: returns cell_num_before, cell_acc, vp_cell, vp_sort"""
    L = lib()
    sub = np.ascontiguousarray(sub, np.int32)
    n = sub.size
    cells = rshx * rshy
    num = np.zeros(cells, np.int32)
    acc = np.zeros(cells, np.int32)
    vpc = np.zeros(max(n, 1), np.int32)
    vps = np.zeros(max(n, 1), np.int32)
    px = np.ascontiguousarray(px, np.float32)
    py = np.ascontiguousarray(py, np.float32)
    L.orc_fill_cells(rshx, rshy, Lx, Ly, n, sub if n else np.zeros(1, np.int32), px, py, num,
                     acc, vpc)
    before = num.copy()
    L.orc_counting_sort(n, num, acc, vpc, vps)
    return before, acc, vpc[:n], vps[:n], num


def build_subsets(omega):
    L = lib()
    w = np.ascontiguousarray(omega, np.float32)
    n = w.size
    sp = np.zeros(max(n, 1), np.int32)
    sn = np.zeros(max(n, 1), np.int32)
    cp = np.zeros(max(n, 1), np.float64)
    cn = np.zeros(max(n, 1), np.float64)
    cnt = np.zeros(2, np.int32)
    cum = np.zeros(2, np.float64)
    L.orc_build_subsets(n, w, sp, sn, cp, cn, cnt, cum)
    a, b = int(cnt[0]), int(cnt[1])
    return dict(sub_pos=sp[:a].copy(), sub_neg=sn[:b].copy(), cdf_pos=cp[:a].copy(),
                cdf_neg=cn[:b].copy(), cum_pos=float(cum[0]), cum_neg=float(cum[1]))


class Problem:
    """A particle set with its sign subsets and reference-grid hashes, ready
    for the oracle's query and estimator.  Arrays are kept alive here."""

    def __init__(self, px, py, w, rshx, rshy, Lx=1.0, Ly=1.0, cdf_override=None, F=None):
        self.F = None if F is None else np.ascontiguousarray(F, np.float32).reshape(-1, 4)
        self.px = np.ascontiguousarray(px, np.float32)
        self.py = np.ascontiguousarray(py, np.float32)
        self.w = np.ascontiguousarray(w, np.float32)
        self.n = self.px.size
        self.rshx, self.rshy, self.Lx, self.Ly = rshx, rshy, float(Lx), float(Ly)
        s = build_subsets(self.w)
        if cdf_override is not None:
            s.update(cdf_override)
        self.sub = s
        self.sets = []
        self._keep = []
        for key in ("pos", "neg"):
            sub = np.ascontiguousarray(s["sub_" + key], np.int32)
            cdf = np.ascontiguousarray(s["cdf_" + key], np.float64)
            _, acc, _, vps, _ = hash_build(rshx, rshy, self.Lx, self.Ly, sub, self.px,
                                           self.py)
            sub_ = sub if sub.size else np.zeros(1, np.int32)
            vps_ = vps if vps.size else np.zeros(1, np.int32)
            cdf_ = cdf if cdf.size else np.zeros(1, np.float64)
            self._keep += [sub_, vps_, acc, cdf_]
            v = View(self.px.ctypes.data, self.py.ctypes.data, self.w.ctypes.data, self.n,
                     vps_.ctypes.data, acc.ctypes.data, rshx, rshy, self.Lx, self.Ly,
                     int(sub.size), sub_.ctypes.data,
                     None if self.F is None else self.F.ctypes.data)
            self.sets.append(SignSet(int(sub.size), sub_.ctypes.data, cdf_.ctypes.data,
                                     float(s["cum_" + key]), v))

    def view(self, sign):
        return self.sets[0 if sign > 0 else 1].view

    def eval(self, qx, qy, sign, h_w, periodic=True, use_signed=False, exclude_center=False):
        L = lib()
        v = self.view(sign)
        return np.array([L.orc_eval_vorticity(C.byref(v), float(a), float(b), int(sign),
                                              int(use_signed), int(exclude_center),
                                              float(h_w), int(periodic))
                         for a, b in zip(np.ravel(qx), np.ravel(qy))], np.float32)

    def params(self, mode, N1, N2, h_w, periodic=True, seed=0, step=0, sim=0, U=None,
               aux_draws=0):
        kc = kper_coeffs(KPER_TERMS)
        self._kc = kc
        self._U = None if U is None else np.ascontiguousarray(U, np.float32)
        return McParams(mode, N1, N2, h_w, self.Lx, self.Ly, int(periodic), seed, step, sim,
                        None if self._U is None else self._U.ctypes.data, KPER_TERMS,
                        kc.ctypes.data, int(aux_draws))

    def velocities(self, P, owners=None, nthreads=0):
        """per-owner velocity v+ - v- (no advection)"""
        L = lib()
        n_owner = self.n if owners is None else owners
        vel = np.zeros(2 * max(self.n, 1), np.float32)
        px = self.px.copy()
        py = self.py.copy()
        L.orc_step(C.byref(P), C.byref(self.sets[0]), C.byref(self.sets[1]), n_owner, px, py,
                   0.0, vel, nthreads)
        return vel[: 2 * n_owner].reshape(-1, 2)

    def velocity_sign(self, P, sign_idx, owner):
        L = lib()
        out = np.zeros(2, np.float32)
        L.orc_mc_velocity(C.byref(P), C.byref(self.sets[sign_idx]), sign_idx, owner,
                          float(self.px[owner]), float(self.py[owner]), self.px, self.py, out)
        return out


def kper_coeffs(nterms=KPER_TERMS):
    c = np.zeros(nterms, np.float64)
    lib().orc_kper_coeffs(nterms, c)
    return c


def kernel_smooth(rx, ry, L=1.0, nterms=KPER_TERMS):
    c = kper_coeffs(nterms)
    kx = C.c_double()
    ky = C.c_double()
    lib().orc_kernel_smooth_d(float(rx), float(ry), float(L), nterms, c, C.byref(kx),
                              C.byref(ky))
    return kx.value, ky.value


def philox(ctr, key):
    out = np.zeros(4, np.uint32)
    lib().orc_philox4x32_10(np.ascontiguousarray(ctr, np.uint32),
                            np.ascontiguousarray(key, np.uint32), out)
    return out


def direct_eps(qx, qy, px, py, g, L=1.0, periodic=True, eps=0.0, skip_self=False,
               nthreads=0):
    qx = np.ascontiguousarray(qx, np.float64)
    qy = np.ascontiguousarray(qy, np.float64)
    u = np.zeros(2 * qx.size, np.float64)
    lib().orc_direct_eps(qx.size, qx, qy, int(skip_self), len(px),
                         np.ascontiguousarray(px, np.float64),
                         np.ascontiguousarray(py, np.float64),
                         np.ascontiguousarray(g, np.float64), L, int(periodic), eps, u,
                         nthreads)
    return u.reshape(-1, 2)


def direct_blob(qx, qy, px, py, g, h_w, L=1.0, nterms=KPER_TERMS, nthreads=0):
    qx = np.ascontiguousarray(qx, np.float64)
    qy = np.ascontiguousarray(qy, np.float64)
    u = np.zeros(2 * qx.size, np.float64)
    c = kper_coeffs(max(nterms, 1))
    lib().orc_direct_blob(qx.size, qx, qy, len(px), np.ascontiguousarray(px, np.float64),
                          np.ascontiguousarray(py, np.float64),
                          np.ascontiguousarray(g, np.float64), h_w, L, nterms, c, u,
                          nthreads)
    return u.reshape(-1, 2)


def rbf_matvec(px, py, x, h_w, L=1.0, periodic=True, nthreads=0):
    y = np.zeros(len(px), np.float64)
    lib().orc_rbf_matvec(len(px), np.ascontiguousarray(px, np.float64),
                         np.ascontiguousarray(py, np.float64), h_w, L, int(periodic),
                         np.ascontiguousarray(x, np.float64), y, nthreads)
    return y


def cg_rbf(px, py, b, h_w, L=1.0, periodic=True, tol=1e-8, max_iter=1000, nthreads=0):
    w = np.zeros(len(px), np.float64)
    rr = C.c_double()
    it = lib().orc_cg_rbf(len(px), np.ascontiguousarray(px, np.float64),
                          np.ascontiguousarray(py, np.float64), h_w, L, int(periodic),
                          np.ascontiguousarray(b, np.float64), w, tol, max_iter, C.byref(rr),
                          nthreads)
    return w, it, rr.value


def tg_lambda11(Nx, hwr, L=1.0):
    """eigenvalue of the periodic lattice Gram matrix on the (1,1) Fourier mode:
    sum over lattice offsets o of W(|o|) cos(2 pi o_x) cos(2 pi o_y)"""
    h = L / Nx
    h_w = hwr * h
    R = int(np.ceil(hwr)) + 1
    a = np.arange(-R, R + 1)
    ox, oy = np.meshgrid(a * h, a * h, indexing="ij")
    r = np.hypot(ox, oy)
    L_ = lib()
    Wv = np.vectorize(lambda v: L_.orc_spline_d(float(v), float(h_w)))(r)
    return float(np.sum(Wv * np.cos(2 * np.pi * ox / L) * np.cos(2 * np.pi * oy / L)))
