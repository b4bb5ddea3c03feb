"""Device-resident matrix-free transport operator.

``TransportOperator(mesh, kappa, sigma, Phi)`` flattens the mesh
(``plan.MeshTables``), creates the device plan, evaluates the coefficient
callables at the nodes and uploads them.  ``apply`` is the matrix-free
counterpart of PETSc ``MatMult`` with the reference's assembled operator
(``src/rt/solve.py:49-52``); ``RESTRICT`` gives the interior block A_II in
full-length vectors (zeros at inflow-boundary DOFs).
"""
import ctypes as C

import numpy as np
import torch

from . import _lib
from ._lib import (OP_MASS, OP_SCAT, OP_ICONV, OP_BCONV, OP_RESTRICT, OP_ACCUM,
                   OP_QUADW, OP_SYSTEM, PREC_F64, PREC_F32, PREC_MIXED, check, ptr)
from .plan import MeshTables

__all__ = ["TransportOperator", "OP_MASS", "OP_SCAT", "OP_ICONV", "OP_BCONV",
           "OP_RESTRICT", "OP_ACCUM", "OP_SYSTEM", "PREC_F64", "PREC_F32",
           "PREC_MIXED"]


def stream_handle(stream=None):
    s = torch.cuda.current_stream() if stream is None else stream
    return C.c_void_p(s.cuda_stream)


def build_desc(t, G_off):
    """rtx_mesh_desc for MeshTables ``t``; returns (desc, arrays to keep alive)."""
    keep = []
    d = _lib.MeshDesc()
    d.ncol, d.ncell, d.nlink, d.ndof = t.ncol, t.ncell, t.nlink, t.ndof

    def P(arr, ct):
        a = np.ascontiguousarray(arr)
        keep.append(a)
        return ptr(a, ct)

    i32, i64, f64 = C.c_int32, C.c_int64, C.c_double
    d.col_cell_start = P(t.col_cell_start, i32)
    d.col_nx, d.col_ny = P(t.col_nx, i32), P(t.col_ny, i32)
    d.col_sp_off, d.col_nang = P(t.col_sp_off, i32), P(t.col_nang, i32)
    d.col_G_off = P(np.asarray(G_off, np.int64), i64)
    d.col_geom = P(t.col_geom.ravel(), f64)
    d.cell_dof_off, d.cell_col = P(t.cell_dof_off, i32), P(t.cell_col, i32)
    d.cell_nth, d.cell_quad = P(t.cell_nth, i32), P(t.cell_quad, i32)
    d.cell_ang_off, d.cell_bdry = P(t.cell_ang_off, i32), P(t.cell_bdry, i32)
    d.cell_geom = P(t.cell_geom.ravel(), f64)
    d.link_start, d.link_src = P(t.link_start, i32), P(t.link_src, i32)
    d.link_face = P(t.link_face, i32)
    d.link_Es_off, d.link_Eth_off = P(t.link_Es_off, i32), P(t.link_Eth_off, i32)
    d.link_target = P(t.link_target, i32)
    d.nEs, d.nEth = t.Es_pool.size, t.Eth_pool.size
    d.Es_pool, d.Eth_pool = P(t.Es_pool, f64), P(t.Eth_pool, f64)
    d.lgl_x, d.lgl_w = P(t.lgl_x, f64), P(t.lgl_w, f64)
    d.lg_x, d.lg_w, d.lgl_D = P(t.lg_x, f64), P(t.lg_w, f64), P(t.lgl_D, f64)
    d.col_ang_start, d.ang_pack = P(t.col_ang_start, i32), P(t.ang_pack, i32)
    d.ang_trig, d.link_meta = P(t.ang_trig, f64), P(t.link_meta, i32)
    d.col_ntrace, d.col_nitems = P(t.col_ntrace, i32), P(t.col_nitems, i32)
    d.nang = t.nang
    return d, keep


class TransportOperator:
    def __init__(self, mesh, kappa, sigma, Phi, device=None, tables=None):
        if not torch.cuda.is_available():
            raise _lib.RtxError("TransportOperator needs a GPU (torch.cuda is not available)")
        self.lib = _lib.load()
        self.device = torch.device("cuda" if device is None else device)
        self.t = MeshTables(mesh) if tables is None else tables
        t = self.t
        self.ndof = t.ndof
        f64 = C.c_double
        G_pool, G_off = t.G_pool(Phi)
        self._G_off = G_off
        d, self._keep = build_desc(t, G_off)
        with torch.cuda.device(self.device):
            h = C.c_void_p()
            check(self.lib.rtx_plan_create(C.byref(d), C.byref(h)), "rtx_plan_create")
            self.plan = h
        self._keep = []
        self.G_pool = np.ascontiguousarray(G_pool)
        self.set_coeffs(kappa, sigma)

    def set_coeffs(self, kappa, sigma, Phi=None):
        """Re-evaluate kappa(x, y), sigma(x, y) (and Phi, if given) at the
        nodes and upload them; the mesh tables are kept."""
        t = self.t
        f64 = C.c_double
        if Phi is not None:
            G_pool, G_off = t.G_pool(Phi)
            if not np.array_equal(G_off, self._G_off):
                raise ValueError("a new Phi must keep the angular signatures of the plan")
            self.G_pool = np.ascontiguousarray(G_pool)
        self.kappa_nodes = np.ascontiguousarray(t.eval_spatial(kappa))
        self.sigma_nodes = np.ascontiguousarray(t.eval_spatial(sigma))
        with torch.cuda.device(self.device):
            check(self.lib.rtx_plan_set_coeffs(
                self.plan, t.nsp, ptr(self.kappa_nodes, f64), ptr(self.sigma_nodes, f64),
                self.G_pool.size, ptr(self.G_pool, f64)), "rtx_plan_set_coeffs")

    def __del__(self):
        try:
            if getattr(self, "plan", None) is not None and self.plan.value:
                self.lib.rtx_plan_destroy(self.plan)
                self.plan = None
        except Exception:
            pass

    # ------------------------------------------------------------------
    def _check_vec(self, v, dtype):
        if not (v.is_cuda and v.dtype == dtype and v.is_contiguous() and v.numel() == self.ndof):
            raise ValueError("expected a contiguous {} CUDA tensor of {} entries, got {} {} {}"
                             .format(dtype, self.ndof, v.dtype, tuple(v.shape), v.device))

    def apply(self, x, y=None, ops=OP_SYSTEM, precision=PREC_F64, stream=None):
        """y = op(x) on the device (y += op(x) with OP_ACCUM)."""
        dtype = torch.float64 if precision == PREC_F64 else torch.float32
        self._check_vec(x, dtype)
        if y is None:
            y = torch.empty_like(x)
        self._check_vec(y, dtype)
        check(self.lib.rtx_apply(self.plan, ops, precision, C.c_void_p(x.data_ptr()),
                                 C.c_void_p(y.data_ptr()), stream_handle(stream)), "rtx_apply")
        return y

    def set_path(self, mode):
        """-1 automatic, 0 generic per-cell/per-column kernels, 1 fused per-part
        kernel, 2 uniform-bin kernel, 3 hp column kernel."""
        check(self.lib.rtx_set_path(self.plan, int(mode)), "rtx_set_path")

    def autotune(self, precision=PREC_F64, reps=10, stream=None):
        """Time every valid number of column parts (1, 2, 4) of the uniform
        kernel and of each fused order bin on a scratch vector pair and keep
        the fastest; returns plan_info() afterwards."""
        dtype = torch.float64 if precision == PREC_F64 else torch.float32
        x = torch.rand(self.ndof, dtype=dtype, device=self.device)
        y = torch.empty_like(x)
        check(self.lib.rtx_autotune(self.plan, int(precision), C.c_void_p(x.data_ptr()),
                                    C.c_void_p(y.data_ptr()), int(reps), stream_handle(stream)),
              "rtx_autotune")
        return self.plan_info()

    def plan_info(self):
        """Launch shapes of the apply: uniform bin (usable, parts per column),
        the fused order bins (parts, columns; by LDS class) and the column and cell
        counts left to the generic kernels."""
        out = (C.c_int64 * 25)()
        n = self.lib.rtx_plan_info(self.plan, out, 25)
        v = list(out)[:n]
        bins = [tuple(v[2 + 3 * b: 5 + 3 * b]) for b in range(6)]
        return {"uniform": bool(v[0]), "uniform_parts": v[1],
                "fused_bins": [(H, nc) for H, nc, g in bins if nc and not g],
                "bins_on_generic": [nc for H, nc, g in bins if nc and g],
                "generic_cols": v[20], "generic_cells": v[21],
                "column_kernel": bool(v[22]), "column_kernel_cols": v[23],
                "uniform_output_layout": bool(v[24])}

    def set_bin_generic(self, b, generic=True):
        """Run fused order bin ``b`` on the generic kernels (or back)."""
        check(self.lib.rtx_set_bin_generic(self.plan, int(b), int(bool(generic))),
              "rtx_set_bin_generic")

    def set_uniform_parts(self, H):
        """Force the uniform kernel's parts per column (0: default)."""
        check(self.lib.rtx_set_uniform_parts(self.plan, int(H)), "rtx_set_uniform_parts")

    def mask(self, x, keep_interior=True, stream=None):
        """Zero the inflow-boundary DOFs (keep_interior) or the interior ones."""
        prec = PREC_F64 if x.dtype == torch.float64 else PREC_F32
        self._check_vec(x, x.dtype)
        check(self.lib.rtx_mask(self.plan, 0 if keep_interior else 1, prec,
                                C.c_void_p(x.data_ptr()), stream_handle(stream)), "rtx_mask")
        return x

    def nodal(self, func):
        """func(x, y, th) at every DOF node in the global DOF order (host)."""
        return self.t.eval_nodal_dofs(func)
