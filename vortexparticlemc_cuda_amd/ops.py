"""Torch-tensor front end over the C ABI (include/vpmc.h).

PyTorch only supplies device memory and the current stream; every
computation is a libvpmc.so kernel.  Workspaces are sized with the C ABI's
*_workspace_bytes queries and allocated here, by the caller, as the ABI
requires.
"""
import ctypes as C
import math

import torch

from . import _capi
from ._capi import Field, McDesc, SignSet, View, check, lib


def _stream():
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def _p(t):
    if t is None:
        return None
    if not t.is_cuda:
        raise ValueError("device tensor required")
    if not t.is_contiguous():
        raise ValueError("contiguous tensor required")
    return C.c_void_p(t.data_ptr())


def workspace(nbytes, device="cuda"):
    return torch.empty(max(int(nbytes), 256), dtype=torch.uint8, device=device)


def _xy(pos):
    pos = pos.to(torch.float32).contiguous()
    if pos.dim() != 2 or pos.shape[1] != 2:
        raise ValueError("positions must be [N, 2]")
    return pos


# ---------------------------------------------------------------------------
# reference-grid hash (FillCells_cuda / CountingSort_cuda)

def fill_cells(rshx, rshy, Lx, Ly, pos, sub=None):
    pos = _xy(pos)
    n = pos.shape[0] if sub is None else sub.numel()
    dev = pos.device
    cells = rshx * rshy
    cell_num = torch.empty(cells, dtype=torch.int32, device=dev)
    cell_acc = torch.empty(cells, dtype=torch.int32, device=dev)
    vp_cell = torch.empty(max(n, 1), dtype=torch.int32, device=dev)
    ws = workspace(lib().vpmc_hash_workspace_bytes(n, cells), dev)
    check(lib().vpmc_fill_cells(rshx, rshy, Lx, Ly, n, _p(sub), _p(pos), _p(cell_num),
                                _p(cell_acc), _p(vp_cell), _p(ws), ws.numel(), _stream()),
          "vpmc_fill_cells")
    return cell_num, cell_acc, vp_cell[:n]


def counting_sort(cell_num, cell_acc, vp_cell):
    n = vp_cell.numel()
    cells = cell_num.numel()
    dev = cell_num.device
    vp_map = torch.empty(max(n, 1), dtype=torch.int32, device=dev)
    ws = workspace(lib().vpmc_hash_workspace_bytes(n, cells), dev)
    check(lib().vpmc_counting_sort(n, cells, _p(cell_num), _p(cell_acc), _p(vp_cell.contiguous()),
                                   _p(vp_map), _p(ws), ws.numel(), _stream()),
          "vpmc_counting_sort")
    return vp_map[:n]


# ---------------------------------------------------------------------------
# subsets, CDF, guide

def build_subsets(omega):
    omega = omega.to(torch.float32).contiguous()
    n = omega.numel()
    dev = omega.device
    m = max(n, 1)
    out = dict(sub_pos=torch.empty(m, dtype=torch.int32, device=dev),
               sub_neg=torch.empty(m, dtype=torch.int32, device=dev),
               cdf_pos=torch.empty(m, dtype=torch.float64, device=dev),
               cdf_neg=torch.empty(m, dtype=torch.float64, device=dev),
               counts=torch.zeros(2, dtype=torch.int32, device=dev),
               totals=torch.zeros(2, dtype=torch.float64, device=dev))
    ws = workspace(lib().vpmc_subsets_workspace_bytes(n), dev)
    check(lib().vpmc_build_subsets(n, _p(omega), _p(out["sub_pos"]), _p(out["sub_neg"]),
                                   _p(out["cdf_pos"]), _p(out["cdf_neg"]), _p(out["counts"]),
                                   _p(out["totals"]), _p(ws), ws.numel(), _stream()),
          "vpmc_build_subsets")
    return out


def guide_log2_for(n):
    return max(0, int(math.ceil(math.log2(max(n, 1)))))


def build_guide(cdf, count_dev, m_log2):
    guide = torch.empty((1 << m_log2) + 1, dtype=torch.int32, device=cdf.device)
    check(lib().vpmc_build_guide(_p(cdf), _p(count_dev), m_log2, _p(guide), _stream()),
          "vpmc_build_guide")
    return guide


def sample_idx(u, cdf, count_dev, guide, m_log2):
    u = u.to(torch.float32).contiguous()
    out = torch.empty(max(u.numel(), 1), dtype=torch.int32, device=u.device)
    check(lib().vpmc_sample_idx(u.numel(), _p(u), _p(cdf), _p(count_dev), _p(guide), m_log2,
                                _p(out), _stream()), "vpmc_sample_idx")
    return out[:u.numel()]


# ---------------------------------------------------------------------------
# queries

def query_view(pos, w, vp_sort, cell_acc, sub, rshx, rshy, Lx, Ly, cnt_view, pts, sign, h_w,
               periodic=True, use_signed=False, exclude_center=False):
    pos = _xy(pos)
    pts = _xy(pts)
    out = torch.empty(max(pts.shape[0], 1), dtype=torch.float32, device=pos.device)
    v = View(_p(pos), _p(w.contiguous()), pos.shape[0], _p(vp_sort), _p(cell_acc), rshx, rshy,
             Lx, Ly, cnt_view, _p(sub))
    check(lib().vpmc_query_view(C.byref(v), pts.shape[0], _p(pts), sign, int(use_signed),
                                int(exclude_center), h_w, int(periodic), _p(out), _stream()),
          "vpmc_query_view")
    return out[:pts.shape[0]]


def field_grid(h_w, Lx=1.0, Ly=1.0, cells_per_hw=8):
    nx, ny = C.c_int(), C.c_int()
    check(lib().vpmc_field_grid(h_w, Lx, Ly, cells_per_hw, C.byref(nx), C.byref(ny)),
          "vpmc_field_grid")
    return nx.value, ny.value


class FieldBuffers:
    """device buffers of one fine-grid field"""

    def __init__(self, nx, ny, Lx, Ly, n, device, deformed=False, gq_bound=0.0):
        self.nx, self.ny, self.Lx, self.Ly, self.n = nx, ny, float(Lx), float(Ly), int(n)
        m = max(n, 1)
        self.pts = torch.zeros((m, 4), dtype=torch.float32, device=device)
        self.cell_start = torch.zeros(nx * ny + 1, dtype=torch.int32, device=device)
        self.order = torch.zeros(m, dtype=torch.int32, device=device)
        self.set_xy = torch.zeros((m, 2), dtype=torch.float32, device=device)
        # deformed kernels: per-record quadratic forms, per-local-index F
        self.gq = self.set_jac = None
        self.gq_bound = float(gq_bound)
        if deformed:
            self.gq = torch.zeros((m, 4), dtype=torch.float32, device=device)
            self.set_jac = torch.zeros((m, 4), dtype=torch.float32, device=device)

    def struct(self):
        return Field(self.nx, self.ny, self.Lx, self.Ly, self.n, _p(self.pts),
                     _p(self.cell_start), _p(self.order), _p(self.gq), self.gq_bound)


def field_build(fb, pos, w, sub=None, abs_w=True, ws=None, jac=None):
    """vpmc_field_build; with jac (n x 4 F per particle) the deformed
    build vpmc_field_build_jac (fb must have been made deformed=True)"""
    pos = _xy(pos)
    if ws is None:
        ws = workspace(lib().vpmc_field_workspace_bytes(fb.n, fb.nx, fb.ny), pos.device)
    wp = _p(None if w is None else w.contiguous())
    if jac is not None:
        if fb.gq is None:
            raise ValueError("field_build: jac needs FieldBuffers(..., deformed=True)")
        jac = jac.to(torch.float32).contiguous()
        check(lib().vpmc_field_build_jac(fb.nx, fb.ny, fb.Lx, fb.Ly, fb.n, _p(sub), _p(pos), wp,
                                         int(abs_w), _p(jac), _p(fb.pts), _p(fb.gq),
                                         _p(fb.cell_start), _p(fb.order), _p(fb.set_xy),
                                         _p(fb.set_jac), _p(ws), ws.numel(), _stream()),
              "vpmc_field_build_jac")
        return fb
    if fb.gq is not None:
        raise ValueError("field_build: a deformed FieldBuffers needs jac")
    check(lib().vpmc_field_build(fb.nx, fb.ny, fb.Lx, fb.Ly, fb.n, _p(sub), _p(pos), wp,
                                 int(abs_w), _p(fb.pts), _p(fb.cell_start), _p(fb.order),
                                 _p(fb.set_xy), _p(ws), ws.numel(), _stream()),
          "vpmc_field_build")
    return fb


def field_query(fb, h_w, pts):
    pts = _xy(pts)
    out = torch.empty(max(pts.shape[0], 1), dtype=torch.float32, device=pts.device)
    f = fb.struct()
    check(lib().vpmc_field_query(C.byref(f), h_w, pts.shape[0], _p(pts), _p(out), _stream()),
          "vpmc_field_query")
    return out[:pts.shape[0]]


def reinit(pos, w, h_w, nx, ny, jac=None, Lx=1.0, Ly=1.0, tol=1e-6, max_iter=1000,
           cells_per_hw=8):
    """vpmc_reinit: the field of (pos, w[, F]) at the nx x ny cell-centred
    lattice, strengths fitted there.  Returns (lattice positions, strengths,
    iterations, relative FP64 residual, status)."""
    pos = _xy(pos)
    n = pos.shape[0]
    ws = workspace(lib().vpmc_reinit_workspace_bytes(n, nx, ny, h_w, Lx, Ly, cells_per_hw,
                                                     int(jac is not None)), pos.device)
    pos_out = torch.empty((nx * ny, 2), dtype=torch.float32, device=pos.device)
    w_out = torch.empty(nx * ny, dtype=torch.float32, device=pos.device)
    if jac is not None:
        jac = jac.to(torch.float32).contiguous()
    it, rr = C.c_int(), C.c_double()
    st = lib().vpmc_reinit(n, _p(pos), _p(w.to(torch.float32).contiguous()), _p(jac), h_w, Lx,
                           Ly, cells_per_hw, nx, ny, tol, max_iter, _p(pos_out), _p(w_out),
                           C.byref(it), C.byref(rr), _p(ws), ws.numel(), _stream())
    if st not in (0, 5, 6):   # ok, CG breakdown, not converged: reported, not raised
        check(st, "vpmc_reinit")
    return pos_out, w_out, it.value, rr.value, st


def advect(pos, vel_sign, dt, Lx=1.0, Ly=1.0, periodic=True):
    n = pos.shape[0]
    vel = torch.empty((max(n, 1), 2), dtype=torch.float32, device=pos.device)
    check(lib().vpmc_advect(n, _p(pos), _p(vel_sign), _p(vel), dt, Lx, Ly, int(periodic),
                            _stream()), "vpmc_advect")
    return vel[:n]


# ---------------------------------------------------------------------------
# RBF / CG

def cg_rbf(pos, b, h_w, Lx=1.0, Ly=1.0, tol=1e-6, max_iter=1000, refine=True, cells_per_hw=8,
           lattice=None):
    """strengths w with A w = b (vpmc_cg_rbf); a float64 b selects the FP64
    right-hand side (vpmc_cg_rbf_f64, MINRES); lattice=(nx, ny) for particles
    on that cell-centred lattice adds its circulant preconditioner
    (vpmc_cg_rbf_lattice, FP64 b).  Returns (w, iterations, relative FP64
    residual, status)."""
    pos = _xy(pos)
    n = pos.shape[0]
    if lattice is not None:
        nxl, nyl = (int(v) for v in lattice)
        if nxl * nyl != n:
            raise ValueError("cg_rbf: lattice %dx%d for %d particles" % (nxl, nyl, n))
        ws = workspace(lib().vpmc_cg_lattice_workspace_bytes(nxl, nyl, h_w, Lx, Ly,
                                                             cells_per_hw), pos.device)
        w = torch.empty(max(n, 1), dtype=torch.float32, device=pos.device)
        it, rr = C.c_int(), C.c_double()
        st = lib().vpmc_cg_rbf_lattice(nxl, nyl, _p(pos), _p(b.to(torch.float64).contiguous()),
                                       _p(w), None, h_w, Lx, Ly, cells_per_hw, tol, max_iter,
                                       C.byref(it), C.byref(rr), _p(ws), ws.numel(), _stream())
        if st not in (0, 5, 6):
            check(st, "vpmc_cg_rbf_lattice")
        return w[:n], it.value, rr.value, st
    nx, ny = field_grid(h_w, Lx, Ly, cells_per_hw)
    ws = workspace(lib().vpmc_cg_workspace_bytes(n, nx, ny), pos.device)
    w = torch.empty(max(n, 1), dtype=torch.float32, device=pos.device)
    it, rr = C.c_int(), C.c_double()
    if b.dtype == torch.float64:
        st = lib().vpmc_cg_rbf_f64(n, _p(pos), _p(b.contiguous()), _p(w), None, h_w, Lx, Ly,
                                   cells_per_hw, tol, max_iter, C.byref(it), C.byref(rr), _p(ws),
                                   ws.numel(), _stream())
    else:
        st = lib().vpmc_cg_rbf(n, _p(pos), _p(b.to(torch.float32).contiguous()), _p(w), h_w, Lx,
                               Ly, cells_per_hw, tol, max_iter, int(refine), C.byref(it),
                               C.byref(rr), _p(ws), ws.numel(), _stream())
    return w[:n], it.value, rr.value, st


def rbf_matvec(fb, h_w, x_sorted, f64=False):
    f = fb.struct()
    if f64:
        y = torch.empty(max(fb.n, 1), dtype=torch.float64, device=x_sorted.device)
        check(lib().vpmc_rbf_matvec_f64(C.byref(f), h_w, _p(x_sorted), _p(y), _stream()),
              "vpmc_rbf_matvec_f64")
    else:
        y = torch.empty(max(fb.n, 1), dtype=torch.float32, device=x_sorted.device)
        check(lib().vpmc_rbf_matvec(C.byref(f), h_w, _p(x_sorted), _p(y), _stream()),
              "vpmc_rbf_matvec")
    return y[:fb.n]


# ---------------------------------------------------------------------------
# FP64 direct sums

def direct_sum_eps(probes, pos, g, L=1.0, periodic=True, eps=0.0, skip_self=False):
    probes = probes.to(torch.float64).contiguous()
    pos = pos.to(torch.float64).contiguous()
    g = g.to(torch.float64).contiguous()
    u = torch.empty((max(probes.shape[0], 1), 2), dtype=torch.float64, device=probes.device)
    check(lib().vpmc_direct_sum_eps(probes.shape[0], _p(probes), int(skip_self), pos.shape[0],
                                    _p(pos), _p(g), L, int(periodic), eps, _p(u), _stream()),
          "vpmc_direct_sum_eps")
    return u[:probes.shape[0]]


def direct_sum_blob(probes, pos, g, h_w, L=1.0):
    probes = probes.to(torch.float64).contiguous()
    pos = pos.to(torch.float64).contiguous()
    g = g.to(torch.float64).contiguous()
    u = torch.empty((max(probes.shape[0], 1), 2), dtype=torch.float64, device=probes.device)
    check(lib().vpmc_direct_sum_blob(probes.shape[0], _p(probes), pos.shape[0], _p(pos), _p(g),
                                     h_w, L, _p(u), _stream()), "vpmc_direct_sum_blob")
    return u[:probes.shape[0]]
