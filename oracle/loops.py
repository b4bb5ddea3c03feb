"""Literal scalar-loop assembly of the reference's matrices (tiny meshes).

Each function walks the same loop nests as the reference and applies the
same drop thresholds, emitting COO triplets:

* ``mass``           - ``src/rt/mass_matrix.py:108-180``
* ``scattering``     - ``src/rt/scattering_matrix.py:106-247`` (returns -S,
                       drops |val| <= 10 EPS = 1e-14, skips nodes with
                       |w w sigma| <= 1e-14 when K != K')
* ``interior_conv``  - ``src/rt/interior_convection_matrix.py:110-250``
* ``boundary_conv``  - ``src/rt/boundary_convection_matrix.py:27-158,160-540``
                       (outflow own-trace terms, upwind neighbour terms,
                       drops |val| <= 1e-15)
* ``preconditioner`` - ``src/rt/preconditioner_matrix.py:116-267`` (M + (-S))
* ``forcing`` / ``bcs`` / ``intr_mask`` - ``forcing_vector.py:103-156``,
  ``boundary_conditions_vector.py:101-207``, ``get_masks.py:22-129``
* ``dirac_bcs``      - the Dirac-delta branch of
                       ``boundary_conditions_vector.py:142-196`` (per-column
                       write index, cells skipped without advancing it)

Only usable on meshes of a few hundred DOFs.  ``assemble.py`` is checked
against these on such meshes.
"""
import numpy as np
import scipy.sparse as sp

from .quad import lgl, lg, dlag, push, pull, lag
from .emat import E_space, E_theta, theta_F
from .meshq import leaf_order, dof_offsets, nhbr_cells

EPS = 1.0e-15
SMALL = 10.0 * EPS


def _coo(rows, cols, vals, n):
    A = sp.coo_matrix((np.asarray(vals, dtype=float),
                       (np.asarray(rows, dtype=np.int64), np.asarray(cols, dtype=np.int64))),
                      shape=(n, n)).tocsr()
    A.eliminate_zeros()
    return A


def _col_nodes(col, kappa):
    x0, y0, x1, y1 = col.pos
    nx, ny = col.ndofs
    xb, wx = lgl(nx)
    yb, wy = lgl(ny)
    xf = push(x0, x1, xb).reshape(nx, 1)
    yf = push(y0, y1, yb).reshape(1, ny)
    k = np.broadcast_to(np.asarray(kappa(xf, yf), dtype=float), (nx, ny))
    return x1 - x0, y1 - y0, nx, ny, wx, wy, k


def mass(mesh, kappa):
    offs, n = dof_offsets(mesh)
    R, C, V = [], [], []
    for ck, kks in leaf_order(mesh):
        col = mesh.cols[ck]
        dx, dy, nx, ny, wx, wy, kap = _col_nodes(col, kappa)
        for kk in kks:
            cell = col.cells[kk]
            nth = cell.ndofs[0]
            _, wth = lg(nth)
            dth = cell.pos[1] - cell.pos[0]
            dc = dx * dy * dth / 8.0
            o = offs[(ck, kk)]
            for i in range(nx):
                for j in range(ny):
                    for a in range(nth):
                        b = o + nth * ny * i + nth * j + a
                        R.append(b); C.append(b)
                        V.append(dc * wth[a] * (wx[i] * wy[j] * kap[i, j]))
    return _coo(R, C, V, n)


def _scat_like(mesh, sigma, Phi, kappa=None):
    offs, n = dof_offsets(mesh)
    R, C, V = [], [], []
    for ck, kks in leaf_order(mesh):
        col = mesh.cols[ck]
        dx, dy, nx, ny, wx, wy, sig = _col_nodes(col, sigma)
        wws = wx.reshape(nx, 1) * wy.reshape(1, ny) * sig
        if kappa is not None:
            _, _, _, _, _, _, kap = _col_nodes(col, kappa)
            wwk = wx.reshape(nx, 1) * wy.reshape(1, ny) * kap
        for k0 in kks:
            c0 = col.cells[k0]
            n0 = c0.ndofs[0]
            t0b, w0 = lg(n0)
            th0 = push(c0.pos[0], c0.pos[1], t0b)
            dth0 = c0.pos[1] - c0.pos[0]
            dc = dx * dy * dth0 / 8.0
            o0 = offs[(ck, k0)]
            for k1 in kks:
                c1 = col.cells[k1]
                n1 = c1.ndofs[0]
                t1b, w1 = lg(n1)
                th1 = push(c1.pos[0], c1.pos[1], t1b)
                dth1 = c1.pos[1] - c1.pos[0]
                o1 = offs[(ck, k1)]
                Ph = np.array([[Phi(th0[r], th1[a]) for a in range(n1)]
                               for r in range(n0)], dtype=float)
                for i in range(nx):
                    for j in range(ny):
                        s = wws[i, j]
                        if k0 != k1 and not abs(s) > SMALL:
                            continue
                        for r in range(n0):
                            for a in range(n1):
                                if kappa is not None and k0 == k1 and r == a:
                                    val = dc * (w1[a] * wwk[i, j]
                                                - (dth1 / 2.0) * s * w0[r] * w1[a] * Ph[r, a])
                                else:
                                    val = -dc * (dth1 / 2.0) * s * w0[r] * w1[a] * Ph[r, a]
                                if abs(val) > SMALL:
                                    R.append(o0 + n0 * ny * i + n0 * j + r)
                                    C.append(o1 + n1 * ny * i + n1 * j + a)
                                    V.append(val)
    return _coo(R, C, V, n)


def scattering(mesh, sigma, Phi):
    return _scat_like(mesh, sigma, Phi)


def preconditioner(mesh, kappa, sigma, Phi):
    return _scat_like(mesh, sigma, Phi, kappa=kappa)


def interior_conv(mesh):
    offs, n = dof_offsets(mesh)
    R, C, V = [], [], []
    for ck, kks in leaf_order(mesh):
        col = mesh.cols[ck]
        x0, y0, x1, y1 = col.pos
        dx, dy = x1 - x0, y1 - y0
        nx, ny = col.ndofs
        xb, wx = lgl(nx)
        yb, wy = lgl(ny)
        Dy = np.array([[dlag(yb, q, yb[j]) for j in range(ny)] for q in range(ny)])
        Dx = np.array([[dlag(xb, p, xb[i]) for i in range(nx)] for p in range(nx)])
        for kk in kks:
            cell = col.cells[kk]
            nth = cell.ndofs[0]
            tb, wth = lg(nth)
            thf = push(cell.pos[0], cell.pos[1], tb)
            dth = cell.pos[1] - cell.pos[0]
            ws = wth * np.sin(thf)
            wc = wth * np.cos(thf)
            o = offs[(ck, kk)]
            dcy = dx * dth / 4.0
            for i in range(nx):
                for j in range(ny):
                    for a in range(nth):
                        for q in range(ny):
                            R.append(o + nth * ny * i + nth * q + a)
                            C.append(o + nth * ny * i + nth * j + a)
                            V.append(dcy * wx[i] * wy[j] * ws[a] * Dy[q, j])
            dcx = dy * dth / 4.0
            for i in range(nx):
                for j in range(ny):
                    for a in range(nth):
                        for p in range(nx):
                            R.append(o + nth * ny * p + nth * j + a)
                            C.append(o + nth * ny * i + nth * j + a)
                            V.append(dcx * wx[i] * wy[j] * wc[a] * Dx[p, i])
    return _coo(R, C, V, n)


def _is_outflow(quad, F):
    return quad is not None and F in (quad, (quad + 1) % 4)


def boundary_conv(mesh):
    offs, n = dof_offsets(mesh)
    R, C, V = [], [], []
    tol = 1.0e-15
    for ck0, kks0 in leaf_order(mesh):
        col0 = mesh.cols[ck0]
        x00, y00, x10, y10 = col0.pos
        dx0, dy0 = x10 - x00, y10 - y00
        nx0, ny0 = col0.ndofs
        _, wx0 = lgl(nx0)
        _, wy0 = lgl(ny0)
        for F in range(4):
            k1, k2 = col0.nhbr_keys[F]
            nbrs = []
            if k1 is not None:
                nbrs.append(k1)
            if k2 is not None and k2 != k1:
                nbrs.append(k2)
            Es = {}
            for nk in nbrs:
                col1 = mesh.cols[nk]
                if F % 2 == 0:   # x-face: couple along y
                    Es[nk] = E_space(ny0, y00, y10, col0.lv, col1.ndofs[1],
                                     col1.pos[1], col1.pos[3], col1.lv)
                else:
                    Es[nk] = E_space(nx0, x00, x10, col0.lv, col1.ndofs[0],
                                     col1.pos[0], col1.pos[2], col1.lv)
            for kk0 in kks0:
                c0 = col0.cells[kk0]
                nth0 = c0.ndofs[0]
                th00, th10 = c0.pos
                dth0 = th10 - th00
                o0 = offs[(ck0, kk0)]
                dc = (dy0 if F % 2 == 0 else dx0) * dth0 / 4.0
                if _is_outflow(c0.quadrant, F):
                    tb, wth = lg(nth0)
                    TF = theta_F(push(th00, th10, tb), F)
                    if F % 2 == 0:
                        ix = nx0 - 1 if F == 0 else 0
                        for j in range(ny0):
                            for a in range(nth0):
                                val = dc * wy0[j] * wth[a] * TF[a]
                                if abs(val) > tol:
                                    b = o0 + nth0 * ny0 * ix + nth0 * j + a
                                    R.append(b); C.append(b); V.append(val)
                    else:
                        jy = ny0 - 1 if F == 1 else 0
                        for i in range(nx0):
                            for a in range(nth0):
                                val = dc * wx0[i] * wth[a] * TF[a]
                                if abs(val) > tol:
                                    b = o0 + nth0 * ny0 * i + nth0 * jy + a
                                    R.append(b); C.append(b); V.append(val)
                    continue
                for nk in nbrs:
                    col1 = mesh.cols[nk]
                    nx1, ny1 = col1.ndofs
                    for kk1 in nhbr_cells(mesh, ck0, kk0, nk):
                        c1 = col1.cells[kk1]
                        nth1 = c1.ndofs[0]
                        o1 = offs[(nk, kk1)]
                        Eth = E_theta(nth0, th00, th10, c0.lv, nth1, c1.pos[0],
                                      c1.pos[1], c1.lv, F)
                        E = Es[nk]
                        if F % 2 == 0:
                            i0 = nx0 - 1 if F == 0 else 0
                            i1 = 0 if F == 0 else nx1 - 1
                            for j in range(ny1):
                                for q in range(ny0):
                                    for a in range(nth1):
                                        for r in range(nth0):
                                            val = dc * E[j, q] * Eth[a, r]
                                            if abs(val) > tol:
                                                R.append(o0 + nth0 * ny0 * i0 + nth0 * q + r)
                                                C.append(o1 + nth1 * ny1 * i1 + nth1 * j + a)
                                                V.append(val)
                        else:
                            j0 = ny0 - 1 if F == 1 else 0
                            j1 = 0 if F == 1 else ny1 - 1
                            for i in range(nx1):
                                for p in range(nx0):
                                    for a in range(nth1):
                                        for r in range(nth0):
                                            val = dc * E[i, p] * Eth[a, r]
                                            if abs(val) > tol:
                                                R.append(o0 + nth0 * ny0 * p + nth0 * j0 + r)
                                                C.append(o1 + nth1 * ny1 * i + nth1 * j1 + a)
                                                V.append(val)
    return _coo(R, C, V, n)


def forcing(mesh, f):
    offs, n = dof_offsets(mesh)
    out = np.zeros(n)
    for ck, kks in leaf_order(mesh):
        col = mesh.cols[ck]
        x0, y0, x1, y1 = col.pos
        nx, ny = col.ndofs
        xb, wx = lgl(nx)
        yb, wy = lgl(ny)
        xf, yf = push(x0, x1, xb), push(y0, y1, yb)
        for kk in kks:
            cell = col.cells[kk]
            nth = cell.ndofs[0]
            tb, wth = lg(nth)
            thf = push(cell.pos[0], cell.pos[1], tb)
            dc = (x1 - x0) * (y1 - y0) * (cell.pos[1] - cell.pos[0]) / 8.0
            o = offs[(ck, kk)]
            for i in range(nx):
                for j in range(ny):
                    for a in range(nth):
                        out[o + nth * ny * i + nth * j + a] = \
                            dc * wx[i] * wy[j] * wth[a] * f(xf[i], yf[j], thf[a])
    return out


def bcs(mesh, bcs_fn):
    offs, n = dof_offsets(mesh)
    out = np.zeros(n)
    for ck, kks in leaf_order(mesh):
        col = mesh.cols[ck]
        x0, y0, x1, y1 = col.pos
        nx, ny = col.ndofs
        xf, yf = push(x0, x1, lgl(nx)[0]), push(y0, y1, lgl(ny)[0])
        for kk in kks:
            cell = col.cells[kk]
            nth = cell.ndofs[0]
            thf = push(cell.pos[0], cell.pos[1], lg(nth)[0])
            o = offs[(ck, kk)]
            for i in range(nx):
                for j in range(ny):
                    for a in range(nth):
                        out[o + nth * ny * i + nth * j + a] = bcs_fn(xf[i], yf[j], thf[a])
    return out


def dirac_bcs(mesh, bcs_fn, dirac):
    """``boundary_conditions_vector.py:101-207`` as written: per column a zero
    vector filled at a running index ``idx`` (``:127-196``), the delta branch
    taken when ``any(dirac)`` (``:142``), each cell written only when one of
    its three "in cell" flags holds (``:164``), scalar bcs calls; the column
    vectors are placed at the columns' first DOFs (``:55-84``)."""
    offs, n = dof_offsets(mesh)
    out = np.zeros(n)
    for ck, kks in leaf_order(mesh):
        col = mesh.cols[ck]
        x0, y0, x1, y1 = col.pos
        nx, ny = col.ndofs
        xxb, yyb = lgl(nx)[0], lgl(ny)[0]
        xxf, yyf = push(x0, x1, xxb), push(y0, y1, yyb)
        col_ndof = sum(nx * ny * col.cells[kk].ndofs[0] for kk in kks)
        col_vec = np.zeros(col_ndof)
        idx = 0
        for kk in kks:
            cell = col.cells[kk]
            th0, th1 = cell.pos
            nth = cell.ndofs[0]
            thb = lg(nth)[0]
            thf = push(th0, th1, thb)
            if any(dirac):
                in_cell = [True, True, True]
                if dirac[0] is not None:
                    if dirac[0] < x0 or x1 < dirac[0]:
                        in_cell[0] = False
                    else:
                        xs_b = pull(x0, x1, dirac[0])
                if dirac[1] is not None:
                    if dirac[1] < y0 or y1 < dirac[1]:
                        in_cell[1] = False
                    else:
                        ys_b = pull(y0, y1, dirac[1])
                if dirac[2] is not None:
                    if dirac[2] < th0 or th1 < dirac[2]:
                        in_cell[2] = False
                    else:
                        ths_b = pull(th0, th1, dirac[2])
                if any(in_cell):
                    for ii in range(nx):
                        if dirac[0] is not None and in_cell[0]:
                            phi_i, x_i = max(0.0, lag(xxb, ii, xs_b)), dirac[0]
                        else:
                            phi_i, x_i = 1.0, xxf[ii]
                        for jj in range(ny):
                            if dirac[1] is not None and in_cell[1]:
                                psi_j, y_j = max(0.0, lag(yyb, jj, ys_b)), dirac[1]
                            else:
                                psi_j, y_j = 1.0, yyf[jj]
                            for aa in range(nth):
                                if dirac[2] is not None and in_cell[2]:
                                    xsi_a, th_a = max(0.0, lag(thb, aa, ths_b)), dirac[2]
                                else:
                                    xsi_a, th_a = 1.0, thf[aa]
                                col_vec[idx] = bcs_fn(x_i, y_j, th_a) * phi_i * psi_j * xsi_a
                                idx += 1
            else:
                for ii in range(nx):
                    for jj in range(ny):
                        for aa in range(nth):
                            col_vec[idx] = bcs_fn(xxf[ii], yyf[jj], thf[aa])
                            idx += 1
        o = offs[(ck, kks[0])]
        out[o:o + col_ndof] = col_vec
    return out


def intr_mask(mesh):
    offs, n = dof_offsets(mesh)
    m = np.ones(n, dtype=bool)
    for ck, kks in leaf_order(mesh):
        col = mesh.cols[ck]
        nx, ny = col.ndofs
        nb = [col.nhbr_keys[F][0] is None and col.nhbr_keys[F][1] is None
              for F in range(4)]
        for kk in kks:
            cell = col.cells[kk]
            nth = cell.ndofs[0]
            q = cell.quadrant
            o = offs[(ck, kk)]
            for i in range(nx):
                for j in range(ny):
                    for a in range(nth):
                        b = o + nth * ny * i + nth * j + a
                        if nb[0] and q in (1, 2) and i == nx - 1:
                            m[b] = False
                        if nb[1] and q in (2, 3) and j == ny - 1:
                            m[b] = False
                        if nb[2] and q in (3, 0) and i == 0:
                            m[b] = False
                        if nb[3] and q in (0, 1) and j == 0:
                            m[b] = False
    return m
