"""Flatten an hp mesh into the device tables of ``rtx_mesh_desc``.

Integer tables (all int32, bit-exact with the reference's ordering):
  * columns in ascending key order, cells in ascending key order within a
    column; DOF offsets follow (``src/rt/mass_matrix.py:54-74``);
  * per cell the 4-bit set of masked faces: a domain face (no neighbour)
    whose directions enter the domain for the cell's quadrant
    (``src/dg/matrix/get_masks.py:80-110``);
  * upwind "links": for every cell, every face F on which the cell is not
    outflow (``boundary_convection_matrix.py:270-273``), every distinct
    neighbour column on F and every cell returned by
    ``nhbr_cells_in_nhbr_col`` there (``boundary_convection_matrix.py:348-524``).

Floating-point tables: column/cell geometry, the 1-D rules, and deduplicated
pools of the face-coupling matrices E^x/E^y (keyed like the reference's cache
``get_Ex.py:42-44``: orders and level relation) and E^theta (keyed by both
cells' angular intervals, orders and face, ``get_Eth.py:47-49``).

The four E branches are evaluated here in matrix form
(``get_Ex.py:51-124``, ``get_Eth.py:51-146``):
  dlv in {0,+1}, n0 >= n1 : E = L1(T01 xi0) diag(w0 t0)
  dlv in {0,+1}, n0 <  n1 : E = L1(T01 xi1) diag(w1 t1) L0(xi1)^T
  dlv = -1,      n0 >  n1 : E = 1/2 L1(xi0) diag(w0 t0) L0(T10 xi0)^T
  dlv = -1,      n0 <= n1 : E = 1/2 diag(w1 t1) L0(T10 xi1)^T
where L(p)[i, k] = l_i(p_k), T01/T10 map between the two intervals and t is
Theta_F at the physical angle of each quadrature node (1 in space).
"""
import numpy as np

from dg.quadrature import quad_xyth, lagrange_matrix, lagrange_deriv_matrix
from dg.projection import push_forward, pull_back, eval_nodal
from ._lib import NMAX

INFLOW_Q = {0: (1, 2), 1: (2, 3), 2: (3, 0), 3: (0, 1)}


def _rules(n):
    xb, wx, _, _, tb, wt = quad_xyth(nnodes_x=n, nnodes_y=1, nnodes_th=n)
    return xb, wx, tb, wt


def face_matrix(n0, a0, b0, lv0, n1, a1, b1, lv1, nodes_w, F=None):
    """E in R^{n1 x n0}; ``nodes_w(n)`` returns (nodes, weights); when F is
    given the weights carry Theta_F (angular coupling)."""
    x0, w0 = nodes_w(n0)
    x1, w1 = nodes_w(n1)
    dlv = lv0 - lv1
    if dlv not in (-1, 0, 1):
        raise ValueError("neighbour level difference {} unsupported".format(dlv))

    def tf(iv_lo, iv_hi, xi):
        if F is None:
            return np.ones(len(xi))
        return np.cos(push_forward(iv_lo, iv_hi, xi) - F * np.pi / 2.0)

    if dlv >= 0:
        if n0 >= n1:
            t = pull_back(a1, b1, push_forward(a0, b0, x0))
            E = lagrange_matrix(x1, t) * (w0 * tf(a0, b0, x0))[None, :]
        else:
            t = pull_back(a1, b1, push_forward(a0, b0, x1))
            E = (lagrange_matrix(x1, t) * (w1 * tf(a0, b0, x1))[None, :]) @ \
                lagrange_matrix(x0, x1).T
    else:
        if n0 > n1:
            t = pull_back(a0, b0, push_forward(a1, b1, x0))
            E = 0.5 * (lagrange_matrix(x1, x0) * (w0 * tf(a1, b1, x0))[None, :]) @ \
                lagrange_matrix(x0, t).T
        else:
            t = pull_back(a0, b0, push_forward(a1, b1, x1))
            E = 0.5 * (w1 * tf(a1, b1, x1))[:, None] * lagrange_matrix(x0, t).T
    if F is not None:
        E = np.where(np.abs(E) < 1.0e-15, 0.0, E)
    return np.ascontiguousarray(E, dtype=np.float64)


def _lgl(n):
    r = quad_xyth(nnodes_x=n)
    return r[0], r[1]


def _lg(n):
    r = quad_xyth(nnodes_th=n)
    return r[4], r[5]


def quad_tables():
    """1-D rules for n = 1..NMAX in the layout of rtx_mesh_desc."""
    T = (NMAX + 1) * NMAX
    lx, lw, gx, gw = (np.zeros(T) for _ in range(4))
    D = np.zeros(T * NMAX)
    for n in range(1, NMAX + 1):
        xb, wx = _lgl(n)
        tb, wt = _lg(n)
        lx[n * NMAX:n * NMAX + n] = xb
        lw[n * NMAX:n * NMAX + n] = wx
        gx[n * NMAX:n * NMAX + n] = tb
        gw[n * NMAX:n * NMAX + n] = wt
        Dn = lagrange_deriv_matrix(xb)        # D[q, j] = l_q'(xi_j)
        for q in range(n):
            D[(n * NMAX + q) * NMAX:(n * NMAX + q) * NMAX + n] = Dn[q]
    return lx, lw, gx, gw, D


class MeshTables:
    """Host copies of every table; ``ang_sig`` groups columns by angular mesh."""

    def __init__(self, mesh):
        if not mesh.has_th:
            raise ValueError("the transport operator needs a mesh with an angular "
                             "dimension (has_th=True)")
        col_keys = sorted(k for k, c in mesh.cols.items() if c.is_lf)
        self.col_keys = col_keys
        ncol = len(col_keys)
        col_index = {k: i for i, k in enumerate(col_keys)}
        cell_index = {}
        col_cell_start = np.zeros(ncol + 1, np.int32)
        col_nx = np.zeros(ncol, np.int32)
        col_ny = np.zeros(ncol, np.int32)
        col_sp_off = np.zeros(ncol, np.int32)
        col_nang = np.zeros(ncol, np.int32)
        col_geom = np.zeros((ncol, 4))
        cd, cc, cn, cq, ca, cb, cg = [], [], [], [], [], [], []
        cell_keys = []
        dof = 0
        sp = 0
        for ci, ck in enumerate(col_keys):
            col = mesh.cols[ck]
            nx, ny = col.ndofs
            if not (1 <= nx <= NMAX and 1 <= ny <= NMAX):
                raise ValueError("column {}: spatial orders {} exceed the device "
                                 "limit of {} nodes".format(ck, col.ndofs, NMAX))
            col_nx[ci], col_ny[ci] = nx, ny
            col_sp_off[ci] = sp
            sp += nx * ny
            col_geom[ci] = col.pos
            col_cell_start[ci] = len(cd)
            open_f = [col.nhbr_keys[F][0] is None and col.nhbr_keys[F][1] is None
                      for F in range(4)]
            ang = 0
            for kk in sorted(col.cells.keys()):
                cell = col.cells[kk]
                if not cell.is_lf:
                    continue
                nth = cell.ndofs[0]
                if not 1 <= nth <= NMAX:
                    raise ValueError("cell ({}, {}): angular order {} exceeds the "
                                     "device limit of {}".format(ck, kk, nth, NMAX))
                cell_index[(ck, kk)] = len(cd)
                cell_keys.append((ck, kk))
                cd.append(dof)
                cc.append(ci)
                cn.append(nth)
                q = cell.quadrant
                cq.append(-1 if q is None else q)
                ca.append(ang)
                bits = 0
                for F in range(4):
                    if open_f[F] and q in INFLOW_Q[F]:
                        bits |= 1 << F
                cb.append(bits)
                cg.append(cell.pos)
                ang += nth
                dof += nx * ny * nth
            col_nang[ci] = ang
        col_cell_start[ncol] = len(cd)
        if dof >= 2 ** 31:
            raise ValueError("{} DOFs exceed the int32 DOF index range".format(dof))
        self.ndof = dof
        self.nsp = sp
        self.ncol, self.ncell = ncol, len(cd)
        self.col_cell_start, self.col_nx, self.col_ny = col_cell_start, col_nx, col_ny
        self.col_sp_off, self.col_nang = col_sp_off, col_nang
        self.col_geom = np.ascontiguousarray(col_geom)
        self.cell_dof_off = np.asarray(cd, np.int32)
        self.cell_col = np.asarray(cc, np.int32)
        self.cell_nth = np.asarray(cn, np.int32)
        self.cell_quad = np.asarray(cq, np.int32)
        self.cell_ang_off = np.asarray(ca, np.int32)
        self.cell_bdry = np.asarray(cb, np.int32)
        self.cell_geom = np.ascontiguousarray(np.asarray(cg, np.float64).reshape(-1, 2))
        self.cell_keys = cell_keys
        self._links(mesh, col_keys, cell_index)
        self._angular_signatures(mesh, col_keys)
        (self.lgl_x, self.lgl_w, self.lg_x, self.lg_w, self.lgl_D) = quad_tables()
        self._packed_column_tables()

    # ------------------------------------------ packed tables (fused kernel)
    def _packed_column_tables(self):
        """Per-angle and per-link metadata packed for one coalesced load per
        column (layout documented in rtx_mesh_desc)."""
        nth = self.cell_nth.astype(np.int64)
        self.nang = int(nth.sum())
        self.col_ang_start = np.zeros(self.ncol + 1, np.int32)
        self.col_ang_start[1:] = np.cumsum(self.col_nang)
        first_dof = self.cell_dof_off[self.col_cell_start[self.cell_col]]
        cell_base = (self.cell_dof_off - first_dof).astype(np.int64)
        cell_of_ang = np.repeat(np.arange(self.ncell), nth)
        ang_start = np.concatenate([[0], np.cumsum(nth)[:-1]])
        loc = np.arange(self.nang) - ang_start[cell_of_ang]
        self.ang_pack = (cell_base[cell_of_ang] | (nth[cell_of_ang] << 20) | (loc << 25)).astype(np.int32)
        th0 = self.cell_geom[cell_of_ang, 0]
        th1 = self.cell_geom[cell_of_ang, 1]
        xi = np.concatenate([_lg(int(n))[0] for n in self.cell_nth]) if self.ncell else np.zeros(0)
        wa = np.concatenate([_lg(int(n))[1] for n in self.cell_nth]) if self.ncell else np.zeros(0)
        th = th0 + (th1 - th0) * 0.5 * (xi + 1.0)
        self.ang_trig = np.ascontiguousarray(np.stack([np.cos(th), np.sin(th), wa], axis=1).ravel())
        L = self.nlink
        meta = np.zeros((max(L, 1), 4), np.int64)
        ntr = np.zeros(self.ncol, np.int64)
        nit = np.zeros(self.ncol, np.int64)
        if L:
            s, F, k = self.link_src.astype(np.int64), self.link_face.astype(np.int64), \
                self.link_target.astype(np.int64)
            c0, c1 = self.cell_col[k].astype(np.int64), self.cell_col[s].astype(np.int64)
            nx1, ny1 = self.col_nx[c1].astype(np.int64), self.col_ny[c1].astype(np.int64)
            nt1 = self.cell_nth[s].astype(np.int64)
            kk = k - self.col_cell_start[c0]
            meta[:L, 0] = self.cell_dof_off[s]
            meta[:L, 1] = F | (nx1 << 2) | (ny1 << 6) | (nt1 << 10) | \
                (self.cell_bdry[s].astype(np.int64) << 14) | (kk << 18)
            tr_sz = np.where(F % 2 == 1, nx1, ny1) * nt1
            it_sz = np.where(F % 2 == 1, self.col_nx[c0], self.col_ny[c0]) * self.cell_nth[k]
            # column-local exclusive prefix sums (links are grouped by column)
            for sz, col in ((tr_sz, 2), (it_sz, 3)):
                cs = np.cumsum(sz) - sz
                first = np.searchsorted(c0, c0, side="left")
                meta[:L, col] = cs - cs[first]
            ntr = np.bincount(c0, weights=tr_sz, minlength=self.ncol).astype(np.int64)
            nit = np.bincount(c0, weights=it_sz, minlength=self.ncol).astype(np.int64)
        self.link_meta = np.ascontiguousarray(meta.astype(np.int32).ravel())
        self.col_ntrace = ntr.astype(np.int32)
        self.col_nitems = nit.astype(np.int32)

    # --------------------------------------------------------------- links
    def _links(self, mesh, col_keys, cell_index):
        es_pool, es_key = [], {}
        et_pool, et_key = [], {}
        es_n, et_n = [0], [0]
        links = []     # (target cell, src cell, F, es_off, et_off)

        def es_get(ck0, ck1, F):
            c0, c1 = mesh.cols[ck0], mesh.cols[ck1]
            if F % 2 == 0:
                n0, a0, b0, n1, a1, b1 = c0.ndofs[1], c0.pos[1], c0.pos[3], \
                    c1.ndofs[1], c1.pos[1], c1.pos[3]
            else:
                n0, a0, b0, n1, a1, b1 = c0.ndofs[0], c0.pos[0], c0.pos[2], \
                    c1.ndofs[0], c1.pos[0], c1.pos[2]
            dlv = c0.lv - c1.lv
            side = 0 if dlv == 0 else (1 if (a1 + b1) > (a0 + b0) else -1)
            key = (n0, n1, dlv, side)
            if key not in es_key:
                E = face_matrix(n0, a0, b0, c0.lv, n1, a1, b1, c1.lv, _lgl)
                es_key[key] = es_n[0]
                es_pool.append(E.ravel())
                es_n[0] += E.size
            return es_key[key]

        def et_get(cell0, cell1, F):
            key = (cell0.ndofs[0], cell0.pos[0], cell0.pos[1], cell0.lv,
                   cell1.ndofs[0], cell1.pos[0], cell1.pos[1], cell1.lv, F)
            if key not in et_key:
                E = face_matrix(cell0.ndofs[0], cell0.pos[0], cell0.pos[1], cell0.lv,
                                cell1.ndofs[0], cell1.pos[0], cell1.pos[1], cell1.lv,
                                _lg, F=F)
                et_key[key] = et_n[0]
                et_pool.append(E.ravel())
                et_n[0] += E.size
            return et_key[key]

        for ck0 in col_keys:
            col0 = mesh.cols[ck0]
            for F in range(4):
                k1, k2 = col0.nhbr_keys[F]
                nbrs = [k for k in ([k1] + ([k2] if k2 != k1 else [])) if k is not None]
                if not nbrs:
                    continue
                for kk0 in sorted(col0.cells.keys()):
                    cell0 = col0.cells[kk0]
                    q = cell0.quadrant
                    if q is not None and F in (q, (q + 1) % 4):
                        continue       # outflow face: own-trace term, no coupling
                    t = cell_index[(ck0, kk0)]
                    for nk in nbrs:
                        es = es_get(ck0, nk, F)
                        for kk1 in mesh.nhbr_cells_in_nhbr_col(ck0, kk0, nk):
                            if kk1 is None:
                                continue
                            cell1 = mesh.cols[nk].cells[kk1]
                            links.append((t, cell_index[(nk, kk1)], F, es,
                                          et_get(cell0, cell1, F)))
        links.sort(key=lambda r: r[0])       # stable: keeps face/neighbour order
        L = np.asarray(links, dtype=np.int64).reshape(-1, 5)
        self.nlink = L.shape[0]
        counts = np.bincount(L[:, 0], minlength=self.ncell) if self.nlink else \
            np.zeros(self.ncell, np.int64)
        self.link_start = np.concatenate([[0], np.cumsum(counts)]).astype(np.int32)
        self.link_src = np.ascontiguousarray(L[:, 1], dtype=np.int32)
        self.link_face = np.ascontiguousarray(L[:, 2], dtype=np.int32)
        self.link_Es_off = np.ascontiguousarray(L[:, 3], dtype=np.int32)
        self.link_Eth_off = np.ascontiguousarray(L[:, 4], dtype=np.int32)
        self.link_target = np.ascontiguousarray(L[:, 0], dtype=np.int32)
        self.Es_pool = np.concatenate(es_pool) if es_pool else np.zeros(1)
        self.Eth_pool = np.concatenate(et_pool) if et_pool else np.zeros(1)

    # ------------------------------------------------------- angular meshes
    def _angular_signatures(self, mesh, col_keys):
        sigs, sig_of_col = {}, np.zeros(len(col_keys), np.int64)
        for ci, ck in enumerate(col_keys):
            col = mesh.cols[ck]
            sig = tuple((tuple(col.cells[kk].pos), col.cells[kk].ndofs[0])
                        for kk in sorted(col.cells.keys()) if col.cells[kk].is_lf)
            if sig not in sigs:
                sigs[sig] = len(sigs)
            sig_of_col[ci] = sigs[sig]
        self.signatures = list(sigs.keys())
        self.col_sig = sig_of_col

    def angular_nodes(self, sig):
        """Physical angles, weights w_a * dth_K / 2 and dth_K of a signature."""
        th, w, dth = [], [], []
        for (lo, hi), nth in sig:
            tb, wt = _lg(nth)
            th.append(push_forward(lo, hi, tb))
            w.append(wt)
            dth.append(np.full(nth, hi - lo))
        return np.concatenate(th), np.concatenate(w), np.concatenate(dth)

    # -------------------------------------------------------- coefficients
    def node_coords(self):
        """Physical (x, y) of every spatial node, column order."""
        xs, ys = np.zeros(self.nsp), np.zeros(self.nsp)
        for ci in range(self.ncol):
            x0, y0, x1, y1 = self.col_geom[ci]
            nx, ny = int(self.col_nx[ci]), int(self.col_ny[ci])
            xb, yb = _lgl(nx)[0], _lgl(ny)[0]
            X, Y = np.meshgrid(push_forward(x0, x1, xb), push_forward(y0, y1, yb),
                               indexing="ij")
            o = self.col_sp_off[ci]
            xs[o:o + nx * ny] = X.ravel()
            ys[o:o + nx * ny] = Y.ravel()
        return xs, ys

    def eval_spatial(self, func):
        """func(x, y) at every spatial node (per column, broadcast call with a
        scalar fallback like the reference's array calls)."""
        out = np.zeros(self.nsp)
        for ci in range(self.ncol):
            x0, y0, x1, y1 = self.col_geom[ci]
            nx, ny = int(self.col_nx[ci]), int(self.col_ny[ci])
            xf = push_forward(x0, x1, _lgl(nx)[0])
            yf = push_forward(y0, y1, _lgl(ny)[0])
            o = self.col_sp_off[ci]
            out[o:o + nx * ny] = eval_nodal(func, xf, yf).ravel()
        return out

    def eval_nodal_dofs(self, func):
        """func(x, y, th) at every DOF in the global order; one broadcast call
        per column over (x nodes, y nodes, all angular nodes of the column)."""
        out = np.zeros(self.ndof)
        for ci in range(self.ncol):
            x0, y0, x1, y1 = self.col_geom[ci]
            nx, ny = int(self.col_nx[ci]), int(self.col_ny[ci])
            xf = push_forward(x0, x1, _lgl(nx)[0])
            yf = push_forward(y0, y1, _lgl(ny)[0])
            k0, k1 = self.col_cell_start[ci], self.col_cell_start[ci + 1]
            ths = [push_forward(self.cell_geom[k, 0], self.cell_geom[k, 1],
                                _lg(int(self.cell_nth[k]))[0]) for k in range(k0, k1)]
            vals = eval_nodal(func, xf, yf, np.concatenate(ths))
            a = 0
            for k, th in zip(range(k0, k1), ths):
                nt = th.size
                o = self.cell_dof_off[k]
                out[o:o + nx * ny * nt] = vals[:, :, a:a + nt].ravel()
                a += nt
        return out

    def G_pool(self, Phi):
        """G^T[A, R] = (dth_K / 8)(dth_K' / 2) w_r w_a Phi(th_R, th_A) per
        angular signature; returns (pool, per-column offsets)."""
        offs_sig, blocks, n = [], [], 0
        for sig in self.signatures:
            th, w, dth = self.angular_nodes(sig)
            Ph = eval_nodal(Phi, th, th)                  # Ph[R, A]
            Gp = (dth[:, None] / 8.0) * (dth[None, :] / 2.0) * w[:, None] * w[None, :] * Ph
            blocks.append(np.ascontiguousarray(Gp.T).ravel())
            offs_sig.append(n)
            n += Gp.size
        offs_sig = np.asarray(offs_sig, np.int64)
        return np.concatenate(blocks), offs_sig[self.col_sig]
