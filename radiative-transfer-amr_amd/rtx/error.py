"""Device evaluation of the analytic L2 error of a nodal solution.

Counterpart of ``src/amr/Error_Indicator/error/error_analytic.py:22-149``:
every cell's nodal values are interpolated to n + 3 nodes per direction (LGL
in space, Gauss-Legendre in angle) and compared with the analytic solution
there, weighted by the n + 3 point rules (``rtx_error_l2``).  The analytic
solution is a host callable, evaluated here on the high-resolution nodes of
every cell; the interpolation, the squared differences and the per-cell
sums run on the GPU.
"""
import ctypes as C

import numpy as np
import torch

from . import _lib
from ._lib import check, NMAX
from .plan import MeshTables
from .operator import build_desc, stream_handle
from dg.quadrature import quad_xyth, lagrange_matrix

MM = NMAX + 3
TAB_STRIDE = 2 * NMAX * MM + 2 * MM      # RTX_ERR_TAB_STRIDE of rtx.h


def error_tables():
    """Interpolation matrices n -> n + 3 (LGL, LG) and the n + 3 point weights
    for n = 1..NMAX, in the layout ``rtx_error_l2`` reads."""
    tab = np.zeros((NMAX + 1) * TAB_STRIDE)
    for n in range(1, NMAX + 1):
        m = n + 3
        xl, _, _, _, xg, _ = quad_xyth(nnodes_x=n, nnodes_th=n)
        yl, wl, _, _, yg, wg = quad_xyth(nnodes_x=m, nnodes_th=m)
        b = n * TAB_STRIDE
        tab[b:b + n * m] = lagrange_matrix(xl, yl).ravel()
        tab[b + NMAX * MM:b + NMAX * MM + n * m] = lagrange_matrix(xg, yg).ravel()
        tab[b + 2 * NMAX * MM:b + 2 * NMAX * MM + m] = wl
        tab[b + 2 * NMAX * MM + MM:b + 2 * NMAX * MM + MM + m] = wg
    return tab


class ErrorEvaluator:
    """Per-cell squared L2 errors of nodal solutions on one mesh."""

    def __init__(self, mesh, device=None, tables=None):
        if not torch.cuda.is_available():
            raise _lib.RtxError("ErrorEvaluator needs a GPU (torch.cuda is not available)")
        self.lib = _lib.load()
        self.device = torch.device("cuda" if device is None else device)
        self.t = MeshTables(mesh) if tables is None else tables
        t = self.t
        # no coefficients are needed: the G offsets of the descriptor stay 0
        d, keep = build_desc(t, np.zeros(t.ncol, np.int64))
        with torch.cuda.device(self.device):
            h = C.c_void_p()
            check(self.lib.rtx_plan_create(C.byref(d), C.byref(h)), "rtx_plan_create")
            self.plan = h
        del keep
        nx = t.col_nx[t.cell_col].astype(np.int64)
        ny = t.col_ny[t.cell_col].astype(np.int64)
        nt = t.cell_nth.astype(np.int64)
        self.hr_sizes = (nx + 3) * (ny + 3) * (nt + 3)
        self.hr_off = np.zeros(t.ncell + 1, np.int64)
        self.hr_off[1:] = np.cumsum(self.hr_sizes)
        self.d_hr_off = torch.from_numpy(self.hr_off[:-1].copy()).to(self.device)
        self.d_tab = torch.from_numpy(error_tables()).to(self.device)

    def __del__(self):
        try:
            if getattr(self, "plan", None) is not None and self.plan.value:
                self.lib.rtx_plan_destroy(self.plan)
                self.plan = None
        except Exception:
            pass

    def hr_values(self, anl_sol):
        """The analytic solution at every cell's (nx + 3)(ny + 3)(nth + 3)
        nodes, cell after cell ((p, q, r) row-major), as the reference calls
        it: on broadcast node arrays of one cell (``error_analytic.py:85-117``)."""
        t = self.t
        out = np.empty(int(self.hr_off[-1]))
        rules = {}

        def rule(n, kind):
            key = (n, kind)
            if key not in rules:
                xl, _, _, _, xg, _ = quad_xyth(nnodes_x=n, nnodes_th=n)
                rules[key] = xl if kind == 0 else xg
            return rules[key]
        for k in range(t.ncell):
            c = int(t.cell_col[k])
            x0, y0, x1, y1 = t.col_geom[c]
            mx, my = int(t.col_nx[c]) + 3, int(t.col_ny[c]) + 3
            mt = int(t.cell_nth[k]) + 3
            th0, th1 = t.cell_geom[k]
            xf = (x0 + (x1 - x0) / 2.0 * (rule(mx, 0) + 1.0)).reshape(mx, 1, 1)
            yf = (y0 + (y1 - y0) / 2.0 * (rule(my, 0) + 1.0)).reshape(1, my, 1)
            tf = (th0 + (th1 - th0) / 2.0 * (rule(mt, 1) + 1.0)).reshape(1, 1, mt)
            v = np.broadcast_to(np.asarray(anl_sol(xf, yf, tf), dtype=np.float64), (mx, my, mt))
            out[self.hr_off[k]:self.hr_off[k + 1]] = v.ravel()
        return out

    def cell_errors(self, uh, u_hr):
        """(dx dy dth / 8) sum w (u - I uh)^2 per cell (global cell order) for
        the nodal vector ``uh`` (FP64, host or device) and the analytic
        values ``u_hr`` (from ``hr_values``)."""
        uh_d = torch.as_tensor(np.ascontiguousarray(uh) if isinstance(uh, np.ndarray) else uh,
                               dtype=torch.float64).to(self.device).contiguous()
        if uh_d.numel() != self.t.ndof:
            raise ValueError("expected {} nodal values, got {}".format(self.t.ndof, uh_d.numel()))
        u_d = torch.as_tensor(np.ascontiguousarray(u_hr) if isinstance(u_hr, np.ndarray)
                              else u_hr, dtype=torch.float64).to(self.device).contiguous()
        if u_d.numel() != int(self.hr_off[-1]):
            raise ValueError("expected {} high-resolution values, got {}"
                             .format(int(self.hr_off[-1]), u_d.numel()))
        err = torch.empty(self.t.ncell, dtype=torch.float64, device=self.device)
        vp = lambda x: C.c_void_p(x.data_ptr())
        check(self.lib.rtx_error_l2(self.plan, vp(uh_d), vp(u_d), vp(self.d_hr_off),
                                    vp(self.d_tab), vp(err), stream_handle()), "rtx_error_l2")
        return err.cpu().numpy()
