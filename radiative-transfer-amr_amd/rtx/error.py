"""Device evaluation of the error measures of a nodal solution.

* Analytic L2 error (``src/amr/Error_Indicator/error/error_analytic.py:22-149``):
  every cell's nodal values are interpolated to n + 3 nodes per direction
  (LGL in space, Gauss-Legendre in angle) and compared with the analytic
  solution there, weighted by the n + 3 point rules (``rtx_error_l2``).  The
  analytic solution is a host callable, evaluated on the high-resolution
  nodes of every cell; the interpolation, the squared differences and the
  per-cell sums run on the GPU.
* High-resolution error (``error_high_resolution.py:65-203``): the same
  against the nodal solution of a p-refined mesh (``rtx_error_l2_nodal``).
* Angular and spatial jump indicators (``error_angular_jump.py:14-99``,
  ``error_spatial_jump.py:17-236``) and the projection integrals hp steering
  reads (``src/dg/projection/cell_intg_xy.py``, ``col_intg_th.py``):
  ``rtx_error_angular_jump``, ``rtx_error_spatial_jump``,
  ``rtx_indicator_moments``.  The face-pair tables of the spatial jumps
  (own cell, neighbour cell, projections onto the shared interval) are built
  here once per mesh.
"""
import ctypes as C

import numpy as np
import torch

from . import _lib
from ._lib import check, NMAX
from .plan import MeshTables
from .operator import build_desc, stream_handle
from dg.quadrature import quad_xyth, lagrange_matrix, lag_eval
from dg.projection import push_forward, pull_back

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


def nodal_error_tables():
    """Interpolation blocks n -> m (LGL, LG) with the m-point weights for
    n, m = 1..NMAX, in the layout ``rtx_error_l2_nodal`` reads:
    block (n, m) at (n (MM + 1) + m) TAB_STRIDE, laid out as a block of
    ``error_tables`` with m in place of n + 3."""
    tab = np.zeros((NMAX + 1) * (MM + 1) * TAB_STRIDE)
    for n in range(1, NMAX + 1):
        xl, _, _, _, xg, _ = quad_xyth(nnodes_x=n, nnodes_th=n)
        for m in range(1, NMAX + 1):
            yl, wl, _, _, yg, wg = quad_xyth(nnodes_x=m, nnodes_th=m)
            b = (n * (MM + 1) + m) * TAB_STRIDE
            tab[b:b + n * m] = lagrange_matrix(xl, yl).ravel()
            tab[b + NMAX * MM:b + NMAX * MM + n * m] = lagrange_matrix(xg, yg).ravel()
            tab[b + 2 * NMAX * MM:b + 2 * NMAX * MM + m] = wl
            tab[b + 2 * NMAX * MM + MM:b + 2 * NMAX * MM + MM + m] = wg
    return tab


def trace_tables():
    """lend[(n 2 + s) NMAX + a] = l_a(2 s - 1) of the n-point Gauss-Legendre
    basis: the bottom (s = 0) and top (s = 1) angular traces
    (``error_angular_jump.py:46-51``)."""
    lend = np.zeros((NMAX + 1) * 2 * NMAX)
    for n in range(1, NMAX + 1):
        tb = quad_xyth(nnodes_th=n)[4]
        for s, x in ((0, -1.0), (1, 1.0)):
            for a in range(n):
                lend[(n * 2 + s) * NMAX + a] = lag_eval(tb, a, x)
    return lend


def build_jump_pairs(mesh, t):
    """Face pairs of the spatial-jump indicator of ``mesh`` (host tables
    ``t``) in the reference's visiting
    order (``error_spatial_jump.py:35-225``): per column (ascending key),
    face, neighbour column (``list(set(nhbr_keys[F]))``), own cell
    (ascending key) and neighbour cell (``nhbr_cells_in_nhbr_col``).
    Returns (col_pair_start, pair ints, scale, pool) as host arrays."""
    from .plan import _lgl, _lg
    index = {key: k for k, key in enumerate(t.cell_keys)}
    pool, pool_key = [], {}
    npool = [0]

    def put(key, fn):
        if key not in pool_key:
            M = np.ascontiguousarray(fn(), dtype=np.float64)
            pool_key[key] = npool[0]
            pool.append(M.ravel())
            npool[0] += M.size
        return pool_key[key]

    def proj(nodes_w, n_own, lo_own, hi_own, nB, lo_A, hi_A):
        # l_i of the own rule at the nB-point rule of the shared interval,
        # pulled back into the own interval (:88-98, 116-126, 177-187)
        xb = nodes_w(n_own)[0]
        yB = nodes_w(nB)[0]
        pts = pull_back(lo_own, hi_own, push_forward(lo_A, hi_A, yB))
        return put((id(nodes_w), n_own, nB, pts.tobytes()),
                   lambda: lagrange_matrix(xb, pts))

    start = np.zeros(t.ncol + 1, np.int32)
    ints, scale = [], []
    for ci, ck0 in enumerate(t.col_keys):
        col0 = mesh.cols[ck0]
        x0_0, y0_0, x1_0, y1_0 = col0.pos
        nx0, ny0 = col0.ndofs
        for F in range(4):
            for ck1 in list(set(col0.nhbr_keys[F])):
                if ck1 is None:
                    continue
                col1 = mesh.cols[ck1]
                x0_1, y0_1, x1_1, y1_1 = col1.pos
                nx1, ny1 = col1.ndofs
                if F % 2 == 0:
                    lo0, hi0, lo1, hi1, n0, n1 = y0_0, y1_0, y0_1, y1_1, ny0, ny1
                else:
                    lo0, hi0, lo1, hi1, n0, n1 = x0_0, x1_0, x0_1, x1_1, nx0, nx1
                loA, hiA = (lo0, hi0) if (hi0 - lo0) <= (hi1 - lo1) else (lo1, hi1)
                nB = max(n0, n1)
                p0 = proj(_lgl, n0, lo0, hi0, nB, loA, hiA)
                p1 = proj(_lgl, n1, lo1, hi1, nB, loA, hiA)
                for kk0 in sorted(col0.cells.keys()):
                    cell0 = col0.cells[kk0]
                    th0_0, th1_0 = cell0.pos
                    nt0 = cell0.ndofs[0]
                    for kk1 in mesh.nhbr_cells_in_nhbr_col(ck0, kk0, ck1):
                        if kk1 is None:
                            continue
                        cell1 = col1.cells[kk1]
                        th0_1, th1_1 = cell1.pos
                        nt1 = cell1.ndofs[0]
                        if (th1_0 - th0_0) <= (th1_1 - th0_1):
                            gA, gB = th0_0, th1_0
                        else:
                            gA, gB = th0_1, th1_1
                        nH = max(nt0, nt1)
                        q0 = proj(_lg, nt0, th0_0, th1_0, nH, gA, gB)
                        q1 = proj(_lg, nt1, th0_1, th1_1, nH, gA, gB)
                        ints.append((index[(ck0, kk0)], index[(ck1, kk1)], F, nB, nH,
                                     p0, p1, q0, q1))
                        scale.append((hiA - loA) * (gB - gA) / 4.0)
        start[ci + 1] = len(ints)
    pair = np.asarray(ints, np.int32).reshape(-1, 9)
    if npool[0] >= 2 ** 31:
        raise ValueError("projection pool exceeds the int32 offset range")
    return (start, np.ascontiguousarray(pair.ravel()), np.asarray(scale, np.float64),
            np.concatenate(pool) if pool else np.zeros(1))


class ErrorEvaluator:
    """Per-cell squared L2 errors and error indicators of nodal solutions on
    one mesh."""

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
        self.mesh = mesh
        self._jump = None
        self._d_jump = None
        self._ang_nbr = None
        self._d_lend = None
        self._d_tab_nodal = None

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

    def _vec(self, x, n, what):
        d = torch.as_tensor(np.ascontiguousarray(x) if isinstance(x, np.ndarray) else x,
                            dtype=torch.float64).to(self.device).contiguous()
        if d.numel() != n:
            raise ValueError("expected {} {}, got {}".format(n, what, d.numel()))
        return d

    # ------------------------------------------------------------ moments
    def moments(self, uh):
        """(intg_xy, intg_th, norm2) of the nodal vector ``uh``:
        intg_xy[col_ang_start + cell_ang_off + a] = ``cell_intg_xy`` of every
        cell, intg_th[col_sp_off + i ny + j] = ``col_intg_th`` of every
        column, norm2[k] = (dx dy dth / 8) sum w w w uh^2 per cell."""
        t = self.t
        u = self._vec(uh, t.ndof, "nodal values")
        xy = torch.empty(max(t.nang, 1), dtype=torch.float64, device=self.device)
        th = torch.empty(max(t.nsp, 1), dtype=torch.float64, device=self.device)
        n2 = torch.empty(max(t.ncell, 1), dtype=torch.float64, device=self.device)
        vp = lambda x: C.c_void_p(x.data_ptr())
        check(self.lib.rtx_indicator_moments(self.plan, vp(u), vp(xy), vp(th), vp(n2),
                                             stream_handle()), "rtx_indicator_moments")
        return (xy.cpu().numpy()[:t.nang], th.cpu().numpy()[:t.nsp],
                n2.cpu().numpy()[:t.ncell])

    # ------------------------------------------------------- angular jumps
    def angular_jump(self, uh):
        """Per-cell angular-jump sums (1/4) sum w w (1/2)(jump_lo^2 +
        jump_up^2) (``error_angular_jump.py:53-75``), global cell order."""
        t = self.t
        if self._ang_nbr is None:
            index = {key: k for k, key in enumerate(t.cell_keys)}
            nbr = np.zeros(2 * t.ncell, np.int32)
            for k, (ck, kk) in enumerate(t.cell_keys):
                lo, up = self.mesh.cols[ck].cells[kk].nhbr_keys[:2]
                nbr[2 * k], nbr[2 * k + 1] = index[(ck, lo)], index[(ck, up)]
            self._ang_nbr = torch.from_numpy(nbr).to(self.device)
            self._d_lend = torch.from_numpy(trace_tables()).to(self.device)
        u = self._vec(uh, t.ndof, "nodal values")
        err = torch.empty(max(t.ncell, 1), dtype=torch.float64, device=self.device)
        vp = lambda x: C.c_void_p(x.data_ptr())
        check(self.lib.rtx_error_angular_jump(self.plan, vp(u), vp(self._ang_nbr),
                                              vp(self._d_lend), vp(err), stream_handle()),
              "rtx_error_angular_jump")
        return err.cpu().numpy()[:t.ncell]

    # ------------------------------------------------------- spatial jumps
    def jump_pairs(self):
        """The face-pair tables of ``build_jump_pairs`` for this mesh."""
        if self._jump is None:
            self._jump = build_jump_pairs(self.mesh, self.t)
        return self._jump

    def spatial_jump(self, uh):
        """Per-column spatial-jump errors col_jump / (perimeter 2 pi)
        (``error_spatial_jump.py:227``), column order."""
        t = self.t
        if self._d_jump is None:
            start, pair, scale, pool = self.jump_pairs()
            to = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(self.device)
            self._d_jump = (to(start), to(pair if pair.size else np.zeros(9, np.int32)),
                            to(scale if scale.size else np.zeros(1)), to(pool))
        u = self._vec(uh, t.ndof, "nodal values")
        err = torch.empty(max(t.ncol, 1), dtype=torch.float64, device=self.device)
        vp = lambda x: C.c_void_p(x.data_ptr())
        ds, dp, dsc, dpl = self._d_jump
        check(self.lib.rtx_error_spatial_jump(self.plan, vp(u), vp(ds), vp(dp), vp(dsc), vp(dpl),
                                              vp(err), stream_handle()),
              "rtx_error_spatial_jump")
        return err.cpu().numpy()[:t.ncol]

    # ----------------------------------------------- high-resolution error
    def nodal_errors(self, uh, uh_hr, tables_hr):
        """(cell_err, hr_norm2) of ``uh`` against the nodal solution ``uh_hr``
        of a p-refined mesh with the same column and cell keys
        (``tables_hr``: its MeshTables): cell_err[k] = (dx dy dth / 8) sum w w
        w (uh_hr - I uh)^2 on the high-resolution nodes of cell k and
        hr_norm2[k] its share of int uh_hr^2 (``error_high_resolution.py:
        65-97, 180-190``)."""
        t, th = self.t, tables_hr
        if th.cell_keys != t.cell_keys:
            raise ValueError("the high-resolution mesh must have the same column and cell keys")
        mx = th.col_nx[th.cell_col].astype(np.int32)
        my = th.col_ny[th.cell_col].astype(np.int32)
        mt = th.cell_nth.astype(np.int32)
        if min(mx.min(initial=1), my.min(initial=1), mt.min(initial=1)) < 1 or \
                max(mx.max(initial=1), my.max(initial=1), mt.max(initial=1)) > NMAX:
            raise ValueError("high-resolution node counts must lie in 1..{}".format(NMAX))
        if self._d_tab_nodal is None:
            self._d_tab_nodal = torch.from_numpy(nodal_error_tables()).to(self.device)
        dims = torch.from_numpy(np.ascontiguousarray(np.stack([mx, my, mt], 1).ravel())).to(
            self.device)
        off = torch.from_numpy(th.cell_dof_off.astype(np.int64)).to(self.device)
        u = self._vec(uh, t.ndof, "nodal values")
        v = self._vec(uh_hr, th.ndof, "high-resolution nodal values")
        err = torch.empty(max(t.ncell, 1), dtype=torch.float64, device=self.device)
        n2 = torch.empty(max(t.ncell, 1), dtype=torch.float64, device=self.device)
        vp = lambda x: C.c_void_p(x.data_ptr())
        check(self.lib.rtx_error_l2_nodal(self.plan, vp(u), vp(v), vp(off), vp(dims),
                                          vp(self._d_tab_nodal), vp(err), vp(n2),
                                          stream_handle()), "rtx_error_l2_nodal")
        return err.cpu().numpy()[:t.ncell], n2.cpu().numpy()[:t.ncell]

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
