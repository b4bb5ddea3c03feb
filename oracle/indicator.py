"""Error indicators, projection integrals and hp steering: CPU restatement
(test infrastructure).

Literal loops over the reference's data model, for small meshes:

* ``angular_jump``: ``src/amr/Error_Indicator/error/error_angular_jump.py:
  26-81``.  Traces at the bottom (theta pulled back to -1) and top (+1) of
  every cell, squared jumps against the lower / upper neighbour's opposite
  trace, cell error (1/4) sum wx wy (1/2)(jump_lo + jump_up).
* ``spatial_jump``: ``error_spatial_jump.py:30-227``.  For every face, every
  neighbour column, every own cell and every neighbour cell: both face traces
  projected onto the larger node count of the smaller (shared) interval in
  space and in angle (the quadruple loops of ``:193-222``), squared
  difference weighted by those rules times d_A dth_G / 4; column error the
  sum over (2 dx + 2 dy) 2 pi.
* ``cell_intg_xy`` / ``col_intg_th``: ``src/dg/projection/cell_intg_xy.py:
  13-36`` and ``col_intg_th.py:13-37``.
* ``cell_hp_steer`` / ``col_hp_steer``: ``cell_hp_steer.py:15-48`` and
  ``col_hp_steer.py:14-64``, per-coefficient loops.
* ``high_resolution``: ``error_high_resolution.py:65-190``; the sextuple loop
  of ``:180-188`` for the interpolation to the p-refined cell.

Pins (tests/test_oracle_indicator.py): fields the jumps must not see
(continuous polynomials of low degree, theta-independent fields), piecewise
constants with closed-form jumps, a linear angle field whose only jump is the
2 pi wrap, closed-form integrals, Legendre-mode fields with closed-form
steering coefficients, and exact interpolation with a planted one-node
difference for the high-resolution error.
"""
import numpy as np
from scipy.special import eval_legendre

from .quad import lgl, lg, lag, push, pull
from .meshq import leaf_order, dof_offsets, nhbr_cells


def _vals(mesh, u_vec, offs, ck, kk):
    nx, ny = mesh.cols[ck].ndofs
    nth = mesh.cols[ck].cells[kk].ndofs[0]
    o = offs[(ck, kk)]
    return np.asarray(u_vec[o:o + nx * ny * nth], dtype=np.float64).reshape(nx, ny, nth)


def angular_jump(mesh, u_vec):
    """{(col_key, cell_key): sum} with sum = (1/4) sum_ij wx wy (1/2)
    ((bot - top_lo)^2 + (top - bot_up)^2); the reported errors are its
    square root (cells) and the square roots of its sums (columns, mesh)."""
    offs, _ = dof_offsets(mesh)
    tops, bots = {}, {}
    for ck, kks in leaf_order(mesh):
        nx, ny = mesh.cols[ck].ndofs
        for kk in kks:
            nth = mesh.cols[ck].cells[kk].ndofs[0]
            tb, _ = lg(nth)
            v = _vals(mesh, u_vec, offs, ck, kk)
            top = np.zeros((nx, ny))
            bot = np.zeros((nx, ny))
            for a in range(nth):
                top += v[:, :, a] * lag(tb, a, 1.0)
                bot += v[:, :, a] * lag(tb, a, -1.0)
            tops[(ck, kk)], bots[(ck, kk)] = top, bot
    out = {}
    for ck, kks in leaf_order(mesh):
        nx, ny = mesh.cols[ck].ndofs
        _, wx = lgl(nx)
        _, wy = lgl(ny)
        for kk in kks:
            lo, up = mesh.cols[ck].cells[kk].nhbr_keys[:2]
            jl = (bots[(ck, kk)] - tops[(ck, lo)]) ** 2
            ju = (tops[(ck, kk)] - bots[(ck, up)]) ** 2
            s = 0.0
            for i in range(nx):
                for j in range(ny):
                    s += wx[i] * wy[j] * 0.5 * (jl[i, j] + ju[i, j])
            out[(ck, kk)] = 0.25 * s
    return out


def _face_trace(v, F, own):
    """Trace on face F of the own cell (own=True) or on the opposite face of
    the neighbour (``error_spatial_jump.py:144-151, 170-177``)."""
    if own:
        return {0: v[-1, :, :], 1: v[:, -1, :], 2: v[0, :, :], 3: v[:, 0, :]}[F]
    return {0: v[0, :, :], 1: v[:, 0, :], 2: v[-1, :, :], 3: v[:, -1, :]}[F]


def spatial_jump(mesh, u_vec):
    """{col_key: col_jump / ((2 dx + 2 dy) 2 pi)}."""
    offs, _ = dof_offsets(mesh)
    out = {}
    for ck0, kks0 in leaf_order(mesh):
        col0 = mesh.cols[ck0]
        x0_0, y0_0, x1_0, y1_0 = col0.pos
        dx0, dy0 = x1_0 - x0_0, y1_0 - y0_0
        nx0, ny0 = col0.ndofs
        jump = 0.0
        for F in range(4):
            for ck1 in list(set(col0.nhbr_keys[F])):
                if ck1 is None:
                    continue
                col1 = mesh.cols[ck1]
                x0_1, y0_1, x1_1, y1_1 = col1.pos
                nx1, ny1 = col1.ndofs
                if F % 2 == 0:
                    a0, b0, a1, b1, n0, n1 = y0_0, y1_0, y0_1, y1_1, ny0, ny1
                else:
                    a0, b0, a1, b1, n0, n1 = x0_0, x1_0, x0_1, x1_1, nx0, nx1
                aA, bA = (a0, b0) if (b0 - a0) <= (b1 - a1) else (a1, b1)
                nB = max(n0, n1)
                xB, wB = lgl(nB)
                xf = push(aA, bA, xB)
                p0 = pull(a0, b0, xf)
                p1 = pull(a1, b1, xf)
                x0n, _ = lgl(n0)
                x1n, _ = lgl(n1)
                psi0 = np.array([[lag(x0n, j, p0[jp]) for jp in range(nB)] for j in range(n0)])
                psi1 = np.array([[lag(x1n, q, p1[jp]) for jp in range(nB)] for q in range(n1)])
                for kk0 in kks0:
                    c0 = col0.cells[kk0]
                    th0_0, th1_0 = c0.pos
                    nt0 = c0.ndofs[0]
                    u0 = _face_trace(_vals(mesh, u_vec, offs, ck0, kk0), F, True)
                    for kk1 in nhbr_cells(mesh, ck0, kk0, ck1):
                        c1 = col1.cells[kk1]
                        th0_1, th1_1 = c1.pos
                        nt1 = c1.ndofs[0]
                        u1 = _face_trace(_vals(mesh, u_vec, offs, ck1, kk1), F, False)
                        if (th1_0 - th0_0) <= (th1_1 - th0_1):
                            gA, gB = th0_0, th1_0
                        else:
                            gA, gB = th0_1, th1_1
                        nH = max(nt0, nt1)
                        tH, wH = lg(nH)
                        tf = push(gA, gB, tH)
                        q0 = pull(th0_0, th1_0, tf)
                        q1 = pull(th0_1, th1_1, tf)
                        t0n, _ = lg(nt0)
                        t1n, _ = lg(nt1)
                        xsi0 = np.array([[lag(t0n, a, q0[ap]) for ap in range(nH)]
                                         for a in range(nt0)])
                        xsi1 = np.array([[lag(t1n, r, q1[ap]) for ap in range(nH)]
                                         for r in range(nt1)])
                        U0 = np.zeros((nB, nH))
                        for j in range(n0):
                            for a in range(nt0):
                                for jp in range(nB):
                                    for ap in range(nH):
                                        U0[jp, ap] += u0[j, a] * psi0[j, jp] * xsi0[a, ap]
                        U1 = np.zeros((nB, nH))
                        for q in range(n1):
                            for r in range(nt1):
                                for jp in range(nB):
                                    for ap in range(nH):
                                        U1[jp, ap] += u1[q, r] * psi1[q, jp] * xsi1[r, ap]
                        jump += (bA - aA) * (gB - gA) / 4.0 * np.sum(
                            wB.reshape(nB, 1) * wH.reshape(1, nH) * (U0 - U1) ** 2)
        out[ck0] = jump / ((2.0 * dx0 + 2.0 * dy0) * 2.0 * np.pi)
    return out


def cell_intg_xy(mesh, u_vec, ck, kk):
    offs, _ = dof_offsets(mesh)
    x0, y0, x1, y1 = mesh.cols[ck].pos
    nx, ny = mesh.cols[ck].ndofs
    _, wx = lgl(nx)
    _, wy = lgl(ny)
    v = _vals(mesh, u_vec, offs, ck, kk)
    out = np.zeros(v.shape[2])
    for a in range(v.shape[2]):
        s = 0.0
        for i in range(nx):
            for j in range(ny):
                s += wx[i] * wy[j] * v[i, j, a]
        out[a] = (x1 - x0) * (y1 - y0) / 4.0 * s
    return out


def col_intg_th(mesh, u_vec, ck):
    offs, _ = dof_offsets(mesh)
    nx, ny = mesh.cols[ck].ndofs
    out = np.zeros((nx, ny))
    for kk in sorted(mesh.cols[ck].cells.keys()):
        cell = mesh.cols[ck].cells[kk]
        th0, th1 = cell.pos
        _, wt = lg(cell.ndofs[0])
        v = _vals(mesh, u_vec, offs, ck, kk)
        for i in range(nx):
            for j in range(ny):
                out[i, j] += (th1 - th0) / 2.0 * np.sum(wt * v[i, j, :])
    return out


def cell_steer_lp(nth, uh_hat):
    """(a_nth, lp) of ``cell_hp_steer.py:34-41`` for the space-integrated
    values uh_hat at the nth Gauss-Legendre nodes; the cell is p-refined
    when lp - 1/2 >= nth."""
    tb, wt = lg(nth)
    a = 0.0
    for r in range(nth):
        a += wt[r] * uh_hat[r] * eval_legendre(nth - 1, tb[r])
    a *= (2.0 * nth - 1.0) / 2.0
    with np.errstate(all="ignore"):
        lp = np.log((2.0 * nth - 1.0) / (2 * (a * a * 2.0 / (2.0 * nth - 1.0)))) \
            / (2.0 * np.log(nth - 1))
    return a, lp


def col_steer_lp(nx, ny, uh_hat):
    """(ax_sq, ay_sq, lp) of ``col_hp_steer.py:31-56`` for the
    angle-integrated values uh_hat[i, j] at the LGL nodes; the column is
    p-refined when lp - 1/2 >= (nx + ny) / 2."""
    xb, wx = lgl(nx)
    yb, wy = lgl(ny)
    ax_sq = 0.0
    for j in range(ny):
        s = 0.0
        for i in range(nx):
            for q in range(ny):
                s += eval_legendre(nx - 1, xb[i]) * wx[i] * wy[q] * uh_hat[i, q] \
                    * eval_legendre(j, yb[q])
        a = (2.0 * nx - 1.0) * (2.0 * j + 1.0) / 4.0 * s
        ax_sq += a * a * 2.0 / (2.0 * j + 1.0)
    ay_sq = 0.0
    for i in range(nx):
        s = 0.0
        for p in range(nx):
            for q in range(ny):
                s += eval_legendre(i, xb[p]) * wx[p] * wy[q] * uh_hat[p, q] \
                    * eval_legendre(ny - 1, yb[q])
        a = (2.0 * i + 1.0) * (2.0 * ny - 1.0) / 4.0 * s
        ay_sq += a * a * 2.0 / (2.0 * i + 1.0)
    with np.errstate(all="ignore"):
        t0 = np.log((2.0 * nx - 1.0) / (2 * ax_sq)) / (2.0 * np.log(nx - 1.0))
        t1 = np.log((2.0 * ny - 1.0) / (2 * ay_sq)) / (2.0 * np.log(ny - 1.0))
    return ax_sq, ay_sq, 0.5 * (t0 + t1)


def high_resolution(mesh, u_vec, mesh_hr, u_hr_vec):
    """({(col_key, cell_key): cell_err}, intg_uh_hr2): cell_err = (dx dy dth
    / 8) sum w w w (uh_hr - I uh)^2 on the p-refined cell's nodes and int
    uh_hr^2 over the p-refined mesh; reported errors are sqrt(err / intg)."""
    offs, _ = dof_offsets(mesh)
    offs_hr, _ = dof_offsets(mesh_hr)
    intg = 0.0
    for ck, kks in leaf_order(mesh_hr):
        x0, y0, x1, y1 = mesh_hr.cols[ck].pos
        nx, ny = mesh_hr.cols[ck].ndofs
        _, wx = lgl(nx)
        _, wy = lgl(ny)
        for kk in kks:
            th0, th1 = mesh_hr.cols[ck].cells[kk].pos
            nt = mesh_hr.cols[ck].cells[kk].ndofs[0]
            _, wt = lg(nt)
            v = _vals(mesh_hr, u_hr_vec, offs_hr, ck, kk)
            intg += (x1 - x0) * (y1 - y0) / 4.0 * (th1 - th0) / 2.0 * np.sum(
                wx.reshape(nx, 1, 1) * wy.reshape(1, ny, 1) * wt.reshape(1, 1, nt) * v ** 2)
    out = {}
    for ck, kks in leaf_order(mesh):
        x0, y0, x1, y1 = mesh.cols[ck].pos
        nx, ny = mesh.cols[ck].ndofs
        mx, my = mesh_hr.cols[ck].ndofs
        xb, _ = lgl(nx)
        yb, _ = lgl(ny)
        xh, wxh = lgl(mx)
        yh, wyh = lgl(my)
        phi = np.array([[lag(xb, i, xh[p]) for p in range(mx)] for i in range(nx)])
        psi = np.array([[lag(yb, j, yh[q]) for q in range(my)] for j in range(ny)])
        for kk in kks:
            th0, th1 = mesh.cols[ck].cells[kk].pos
            nt = mesh.cols[ck].cells[kk].ndofs[0]
            mt = mesh_hr.cols[ck].cells[kk].ndofs[0]
            tb, _ = lg(nt)
            th_, wth = lg(mt)
            xsi = np.array([[lag(tb, a, th_[r]) for r in range(mt)] for a in range(nt)])
            v = _vals(mesh, u_vec, offs, ck, kk)
            vi = np.zeros((mx, my, mt))
            for p in range(mx):
                for q in range(my):
                    for r in range(mt):
                        s = 0.0
                        for i in range(nx):
                            for j in range(ny):
                                for a in range(nt):
                                    s += v[i, j, a] * phi[i, p] * psi[j, q] * xsi[a, r]
                        vi[p, q, r] = s
            vh = _vals(mesh_hr, u_hr_vec, offs_hr, ck, kk)
            out[(ck, kk)] = (x1 - x0) * (y1 - y0) * (th1 - th0) / 8.0 * np.sum(
                wxh.reshape(mx, 1, 1) * wyh.reshape(1, my, 1) * wth.reshape(1, 1, mt)
                * (vh - vi) ** 2)
    return out, intg
