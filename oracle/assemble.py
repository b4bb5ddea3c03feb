"""Vectorised assembly of the reference's operator matrices.

Same matrices and drop thresholds as ``loops.py`` (which follows the
reference's loop nests line by line), vectorised over nodes so meshes of
10^5-10^6 DOFs assemble in seconds.  ``cols=`` restricts the assembly to the
rows owned by a subset of columns (all of their couplings included), which
is how large meshes are spot-checked against the device operator.

Reference: ``src/rt/mass_matrix.py:140-173``,
``interior_convection_matrix.py:161-243``,
``boundary_convection_matrix.py:258-537``, ``scattering_matrix.py:142-237``,
``preconditioner_matrix.py:225-248``, ``forcing_vector.py:141-153``,
``boundary_conditions_vector.py:198-205``.

Pinned against ``loops.py`` on tiny meshes and against closed-form
invariants (tests/test_oracle_pins.py).
"""
import numpy as np
import scipy.sparse as sp

from .quad import lgl, lg, dmat, push
from .emat import E_space, E_theta, theta_F
from .meshq import leaf_order, dof_offsets, nhbr_cells

EPS = 1.0e-15
SMALL = 10.0 * EPS


def eval_grid(func, *coords):
    shape = tuple(len(c) for c in coords)
    grids = []
    for d, c in enumerate(coords):
        s = [1] * len(coords)
        s[d] = len(c)
        grids.append(np.asarray(c, dtype=float).reshape(s))
    try:
        with np.errstate(all="ignore"):
            v = np.broadcast_to(np.asarray(func(*grids), dtype=float), shape)
        return np.array(v)
    except Exception:
        out = np.empty(shape)
        for idx in np.ndindex(*shape):
            out[idx] = func(*(float(coords[d][idx[d]]) for d in range(len(coords))))
        return out


class _Acc:
    def __init__(self):
        self.R, self.C, self.V = [], [], []

    def add(self, r, c, v):
        self.R.append(np.asarray(r, dtype=np.int64).ravel())
        self.C.append(np.asarray(c, dtype=np.int64).ravel())
        self.V.append(np.asarray(v, dtype=float).ravel())

    def csr(self, n):
        if not self.R:
            return sp.csr_matrix((n, n))
        A = sp.coo_matrix((np.concatenate(self.V),
                           (np.concatenate(self.R), np.concatenate(self.C))),
                          shape=(n, n)).tocsr()
        A.eliminate_zeros()
        return A


def _geom(col):
    x0, y0, x1, y1 = col.pos
    nx, ny = col.ndofs
    xb, wx = lgl(nx)
    yb, wy = lgl(ny)
    return x0, y0, x1, y1, nx, ny, xb, wx, yb, wy


def assemble(mesh, kappa, sigma, Phi, cols=None, parts=("M", "S", "IC", "BC", "P"), drop=True):
    """Dict of CSR matrices (full n x n shape).  ``S`` is the reference's
    returned scattering matrix (i.e. -S); ``P`` = M + S_ret computed as the
    reference's preconditioner_matrix does.  ``drop=False`` keeps the entries
    the reference's thresholds drop (|S| <= 1e-14, ``scattering_matrix.py:114``;
    |BC| <= 1e-15, ``boundary_convection_matrix.py:310,335``): the exact
    discrete operator, which the matrix-free device path applies."""
    small = SMALL if drop else -1.0
    offs, n = dof_offsets(mesh)
    sel = None if cols is None else set(cols)
    acc = {p: _Acc() for p in parts}
    order = leaf_order(mesh)
    for ck, kks in order:
        if sel is not None and ck not in sel:
            continue
        col = mesh.cols[ck]
        x0, y0, x1, y1, nx, ny, xb, wx, yb, wy = _geom(col)
        dx, dy = x1 - x0, y1 - y0
        xf, yf = push(x0, x1, xb), push(y0, y1, yb)
        ww = np.outer(wx, wy)
        kap = eval_grid(kappa, xf, yf) if ("M" in parts or "P" in parts) else None
        sig = eval_grid(sigma, xf, yf) if ("S" in parts or "P" in parts) else None
        cinfo = []
        for kk in kks:
            c = col.cells[kk]
            nth = c.ndofs[0]
            tb, wth = lg(nth)
            cinfo.append((kk, c, nth, wth, push(c.pos[0], c.pos[1], tb),
                          c.pos[1] - c.pos[0], offs[(ck, kk)]))
        I, J = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")
        for kk, c, nth, wth, thf, dth, o in cinfo:
            dc = dx * dy * dth / 8.0
            idx = o + np.arange(nx * ny * nth)
            if "M" in parts:
                acc["M"].add(idx, idx, dc * (ww * kap)[:, :, None] * wth[None, None, :])
            if "IC" in parts:
                a = np.arange(nth)
                # y-derivative part: rows (i, q, a), cols (i, j, a)
                Dy = dmat(yb)
                ii, jj, aa, qq = np.meshgrid(np.arange(nx), np.arange(ny), a,
                                             np.arange(ny), indexing="ij")
                v = (dx * dth / 4.0) * ww[ii, jj] * (wth * np.sin(thf))[aa] * Dy[qq, jj]
                acc["IC"].add(o + nth * ny * ii + nth * qq + aa,
                              o + nth * ny * ii + nth * jj + aa, v)
                Dx = dmat(xb)
                ii, jj, aa, pp = np.meshgrid(np.arange(nx), np.arange(ny), a,
                                             np.arange(nx), indexing="ij")
                v = (dy * dth / 4.0) * ww[ii, jj] * (wth * np.cos(thf))[aa] * Dx[pp, ii]
                acc["IC"].add(o + nth * ny * pp + nth * jj + aa,
                              o + nth * ny * ii + nth * jj + aa, v)
        if "S" in parts or "P" in parts:
            wws = ww * sig
            s_ok = np.abs(wws) > small
            sidx = (I * ny + J)
            for k0, c0, n0, w0, th0, dth0, o0 in cinfo:
                dc = dx * dy * dth0 / 8.0
                for k1, c1, n1, w1, th1, dth1, o1 in cinfo:
                    Ph = eval_grid(Phi, th0, th1)
                    base = -dc * (dth1 / 2.0) * (w0[:, None] * w1[None, :] * Ph)
                    vals = wws[:, :, None, None] * base[None, None, :, :]
                    rows = o0 + sidx[:, :, None, None] * n0 + np.arange(n0)[None, None, :, None]
                    colsi = o1 + sidx[:, :, None, None] * n1 + np.arange(n1)[None, None, None, :]
                    rows = np.broadcast_to(rows, vals.shape)
                    colsi = np.broadcast_to(colsi, vals.shape)
                    keep = np.abs(vals) > small
                    if k0 != k1:
                        keep &= s_ok[:, :, None, None]
                    if "S" in parts:
                        acc["S"].add(rows[keep], colsi[keep], vals[keep])
                    if "P" in parts:
                        pv = vals.copy()
                        if k0 == k1:
                            wwk = (ww * kap)
                            r_eq = np.arange(n0)
                            pv[:, :, r_eq, r_eq] = dc * (w1[None, None, :] * wwk[:, :, None]
                                                         - (dth1 / 2.0) * wws[:, :, None]
                                                         * (w0 * w1 * np.diag(Ph))[None, None, :])
                        pkeep = np.abs(pv) > small
                        if k0 != k1:
                            pkeep &= s_ok[:, :, None, None]
                        acc["P"].add(rows[pkeep], colsi[pkeep], pv[pkeep])
        if "BC" in parts:
            _bc_column(mesh, ck, col, cinfo, offs, acc["BC"], drop)
    return {p: acc[p].csr(n) for p in parts}


def _bc_column(mesh, ck, col, cinfo, offs, acc, drop=True):
    tol = 1.0e-15 if drop else -1.0
    x0, y0, x1, y1, nx, ny, xb, wx, yb, wy = _geom(col)
    dx, dy = x1 - x0, y1 - y0
    for F in range(4):
        k1, k2 = col.nhbr_keys[F]
        nbrs = [k for k in ([k1] + ([k2] if k2 != k1 else [])) if k is not None]
        Es = {}
        for nk in nbrs:
            c1 = mesh.cols[nk]
            if F % 2 == 0:
                Es[nk] = E_space(ny, y0, y1, col.lv, c1.ndofs[1], c1.pos[1], c1.pos[3], c1.lv)
            else:
                Es[nk] = E_space(nx, x0, x1, col.lv, c1.ndofs[0], c1.pos[0], c1.pos[2], c1.lv)
        for kk, c, nth, wth, thf, dth, o in cinfo:
            dc = (dy if F % 2 == 0 else dx) * dth / 4.0
            q = c.quadrant
            if q is not None and F in (q, (q + 1) % 4):
                TF = theta_F(thf, F)
                if F % 2 == 0:
                    ix = nx - 1 if F == 0 else 0
                    v = dc * wy[:, None] * (wth * TF)[None, :]
                    b = o + nth * ny * ix + nth * np.arange(ny)[:, None] + np.arange(nth)[None, :]
                else:
                    jy = ny - 1 if F == 1 else 0
                    v = dc * wx[:, None] * (wth * TF)[None, :]
                    b = o + nth * ny * np.arange(nx)[:, None] + nth * jy + np.arange(nth)[None, :]
                keep = np.abs(v) > tol
                acc.add(b[keep], b[keep], v[keep])
                continue
            for nk in nbrs:
                col1 = mesh.cols[nk]
                nx1, ny1 = col1.ndofs
                E = Es[nk]
                for kk1 in nhbr_cells(mesh, ck, kk, nk):
                    c1 = col1.cells[kk1]
                    nth1 = c1.ndofs[0]
                    o1 = offs[(nk, kk1)]
                    Eth = E_theta(nth, c.pos[0], c.pos[1], c.lv, nth1,
                                  c1.pos[0], c1.pos[1], c1.lv, F)
                    # vals[m1, m0, a, r] = dc E[m1, m0] Eth[a, r]
                    v = dc * E[:, :, None, None] * Eth[None, None, :, :]
                    m1 = np.arange(E.shape[0])[:, None, None, None]
                    m0 = np.arange(E.shape[1])[None, :, None, None]
                    a = np.arange(nth1)[None, None, :, None]
                    r = np.arange(nth)[None, None, None, :]
                    if F % 2 == 0:
                        i0 = nx - 1 if F == 0 else 0
                        i1 = 0 if F == 0 else nx1 - 1
                        rows = o + nth * ny * i0 + nth * m0 + r
                        cols = o1 + nth1 * ny1 * i1 + nth1 * m1 + a
                    else:
                        j0 = ny - 1 if F == 1 else 0
                        j1 = 0 if F == 1 else ny1 - 1
                        rows = o + nth * ny * m0 + nth * j0 + r
                        cols = o1 + nth1 * ny1 * m1 + nth1 * j1 + a
                    rows = np.broadcast_to(rows, v.shape)
                    cols = np.broadcast_to(cols, v.shape)
                    keep = np.abs(v) > tol
                    acc.add(rows[keep], cols[keep], v[keep])


def system_matrix(mats):
    """A = (BC - IC) + P (``src/rt/solve.py:49-52``)."""
    return ((mats["BC"] - mats["IC"]) + mats["P"]).tocsr()


def forcing(mesh, f):
    offs, n = dof_offsets(mesh)
    out = np.zeros(n)
    for ck, kks in leaf_order(mesh):
        col = mesh.cols[ck]
        x0, y0, x1, y1, nx, ny, xb, wx, yb, wy = _geom(col)
        xf, yf = push(x0, x1, xb), push(y0, y1, yb)
        for kk in kks:
            c = col.cells[kk]
            nth = c.ndofs[0]
            tb, wth = lg(nth)
            thf = push(c.pos[0], c.pos[1], tb)
            dc = (x1 - x0) * (y1 - y0) * (c.pos[1] - c.pos[0]) / 8.0
            fv = eval_grid(f, xf, yf, thf)
            o = offs[(ck, kk)]
            out[o:o + nx * ny * nth] = (dc * wx[:, None, None] * wy[None, :, None]
                                        * wth[None, None, :] * fv).ravel()
    return out


def bcs(mesh, bcs_fn):
    offs, n = dof_offsets(mesh)
    out = np.zeros(n)
    for ck, kks in leaf_order(mesh):
        col = mesh.cols[ck]
        x0, y0, x1, y1, nx, ny, xb, wx, yb, wy = _geom(col)
        xf, yf = push(x0, x1, xb), push(y0, y1, yb)
        for kk in kks:
            c = col.cells[kk]
            nth = c.ndofs[0]
            thf = push(c.pos[0], c.pos[1], lg(nth)[0])
            o = offs[(ck, kk)]
            out[o:o + nx * ny * nth] = eval_grid(bcs_fn, xf, yf, thf).ravel()
    return out


def intr_mask(mesh):
    offs, n = dof_offsets(mesh)
    m = np.ones(n, dtype=bool)
    for ck, kks in leaf_order(mesh):
        col = mesh.cols[ck]
        nx, ny = col.ndofs
        nb = [col.nhbr_keys[F][0] is None and col.nhbr_keys[F][1] is None for F in range(4)]
        for kk in kks:
            c = col.cells[kk]
            nth = c.ndofs[0]
            q = c.quadrant
            blk = np.ones((nx, ny, nth), dtype=bool)
            if nb[0] and q in (1, 2):
                blk[nx - 1] = False
            if nb[1] and q in (2, 3):
                blk[:, ny - 1] = False
            if nb[2] and q in (3, 0):
                blk[0] = False
            if nb[3] and q in (0, 1):
                blk[:, 0] = False
            o = offs[(ck, kk)]
            m[o:o + nx * ny * nth] = blk.ravel()
    return m
