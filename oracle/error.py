"""Analytic L2 error of a DG solution: CPU restatement (test infrastructure).

Follows ``src/amr/Error_Indicator/error/error_analytic.py:22-149``: per
cell, the nodal solution (nx, ny, nth) is interpolated to nx + 3, ny + 3,
nth + 3 nodes (LGL in space, Gauss-Legendre in angle; ``:53-120``), the
analytic solution is evaluated at those nodes (``:116-117``), and

    cell_err = (dx dy dth / 8) sum_pqr wx_p wy_q wth_r (u - u_hr)^2   (:136-137)

Column and mesh errors sum the cell errors; reported errors are
sqrt(err / int u^2) with int u^2 over [0, Lx] x [0, Ly] x [0, 2 pi]
(``:28-32``, ``:139-147``).  The interpolation is the literal quintuple loop
of ``:120-132`` for small cells here, written as three 1-D contractions
(the same sums in a different order).
"""
import numpy as np

from .quad import lgl, lg, lag, push
from .meshq import leaf_order, dof_offsets


def interp_matrix(rule, n, m):
    """M[i, p] = l_i(xi_p): basis on the n-point rule at the m-point rule's
    nodes (``error_analytic.py:57-66``)."""
    x, _ = rule(n)
    y, _ = rule(m)
    return np.array([[lag(x, i, y[p]) for p in range(m)] for i in range(n)])


def cell_errors(mesh, u_vec, anl_sol):
    """Squared (unnormalised) error of every leaf cell, in global order:
    list of (col_key, cell_key, err)."""
    offs, _ = dof_offsets(mesh)
    out = []
    for ck, kks in leaf_order(mesh):
        col = mesh.cols[ck]
        x0, y0, x1, y1 = col.pos
        nx, ny = col.ndofs
        mx, my = nx + 3, ny + 3
        Px, Py = interp_matrix(lgl, nx, mx), interp_matrix(lgl, ny, my)
        xb, wx = lgl(mx)
        yb, wy = lgl(my)
        xf = push(x0, x1, xb).reshape(mx, 1, 1)
        yf = push(y0, y1, yb).reshape(1, my, 1)
        for kk in kks:
            cell = col.cells[kk]
            th0, th1 = cell.pos
            nth = cell.ndofs[0]
            mt = nth + 3
            Pt = interp_matrix(lg, nth, mt)
            tb, wt = lg(mt)
            thf = push(th0, th1, tb).reshape(1, 1, mt)
            o = offs[(ck, kk)]
            uh = np.asarray(u_vec[o:o + nx * ny * nth]).reshape(nx, ny, nth)
            uh_hr = np.einsum("ijk,ip,jq,kr->pqr", uh, Px, Py, Pt)
            u = np.broadcast_to(anl_sol(xf, yf, thf), (mx, my, mt))
            w = wx.reshape(mx, 1, 1) * wy.reshape(1, my, 1) * wt.reshape(1, 1, mt)
            err = (x1 - x0) * (y1 - y0) * (th1 - th0) / 8.0 * np.sum(w * (u - uh_hr) ** 2)
            out.append((ck, kk, float(err)))
    return out


def error_analytic(mesh, u_vec, anl_sol, intg_u2):
    """(mesh error, {col_key: column error}, {(col_key, cell_key): cell
    error}), each sqrt(err / intg_u2)."""
    cells = cell_errors(mesh, u_vec, anl_sol)
    cols = {}
    tot = 0.0
    for ck, kk, e in cells:
        cols[ck] = cols.get(ck, 0.0) + e
        tot += e
    return (np.sqrt(tot / intg_u2),
            {ck: np.sqrt(e / intg_u2) for ck, e in cols.items()},
            {(ck, kk): np.sqrt(e / intg_u2) for ck, kk, e in cells})
