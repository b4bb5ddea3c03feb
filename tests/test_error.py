"""Pins of the oracle's analytic-error restatement
(``src/amr/Error_Indicator/error/error_analytic.py:22-149``) and of the
device kernel against it.

The checks use what the mathematics fixes: a solution inside the DG space
has error 0; the zero solution has error exactly 1 against int u^2 when the
n + 3 point rules integrate u^2 exactly; the interpolation error of x^n on
the n nodes of a column is its monic nodal polynomial prod (x - X_i), whose
squared integral a high-order Gauss rule computes exactly (and likewise for
theta^n on a cell's Gauss nodes).
"""
import math

import numpy as np
import pytest

from oracle import error as oerr, meshq, quad
from rtx import problems


def _mesh():
    m = problems.uniform_mesh([1.0, 1.0], [3, 3, 3], 1, 1)
    keys = sorted(m.cols)
    m.ref_col(keys[0], "spt", "p")
    m.ref_cell(keys[-1], sorted(m.cols[keys[-1]].cells)[0], "p")
    return m


def _nodal(mesh, f):
    """Nodal values of f(x, y, th) in global order (the oracle's nodes)."""
    offs, N = meshq.dof_offsets(mesh)
    u = np.zeros(N)
    for ck, kks in meshq.leaf_order(mesh):
        col = mesh.cols[ck]
        x0, y0, x1, y1 = col.pos
        nx, ny = col.ndofs
        xs = quad.push(x0, x1, quad.lgl(nx)[0])
        ys = quad.push(y0, y1, quad.lgl(ny)[0])
        for kk in kks:
            c = col.cells[kk]
            nth = c.ndofs[0]
            ts = quad.push(c.pos[0], c.pos[1], quad.lg(nth)[0])
            o = offs[(ck, kk)]
            v = f(xs[:, None, None], ys[None, :, None], ts[None, None, :])
            u[o:o + nx * ny * nth] = np.broadcast_to(v, (nx, ny, nth)).ravel()
    return u


def test_solution_in_space_has_zero_error():
    mesh = _mesh()
    f = lambda x, y, th: (1.0 + x) * (1.0 - 2.0 * y) * (0.5 + th)
    err, cols, cells = oerr.error_analytic(mesh, _nodal(mesh, f), f, 1.0)
    assert err < 1e-13 and max(cells.values()) < 1e-13


def test_zero_solution_has_unit_error():
    mesh = _mesh()
    f = lambda x, y, th: 2.0 * (1.0 + x * y) + 0.0 * th
    # int over [0,1]^2 x [0, 2 pi] of (2 (1 + x y))^2 = 4 * 2 pi * (1 + 1/2 + 1/9)
    intg = 4.0 * 2.0 * math.pi * (1.0 + 0.5 + 1.0 / 9.0)
    N = meshq.dof_offsets(mesh)[1]
    err, cols, cells = oerr.error_analytic(mesh, np.zeros(N), f, intg)
    assert abs(err - 1.0) < 1e-13
    assert abs(sum(e * e for e in cols.values()) - 1.0) < 1e-13


def _monic_nodal_sq_integral(nodes, a, b):
    """int_a^b prod (x - X_i)^2 dx by a 12-point Gauss rule (exact to degree
    23), the product evaluated in factored form to avoid cancellation."""
    g, w = np.polynomial.legendre.leggauss(12)
    x = a + (b - a) * 0.5 * (g + 1.0)
    om = np.prod([x - xi for xi in nodes], axis=0)
    return float(np.sum(w * 0.5 * (b - a) * om * om))


def test_interpolation_error_is_the_nodal_polynomial():
    """u = x^n (n = the column's node count): per column the error is
    (y1 - y0) * 2 pi * int (prod (x - X_i))^2; u = th^m likewise per cell."""
    # x^n with n varying by column is not one function: one column order
    mesh = problems.uniform_mesh([1.0, 1.0], [4, 3, 3], 1, 1)
    f = lambda x, y, th: x ** 4 + 0.0 * y + 0.0 * th
    N = meshq.dof_offsets(mesh)[1]
    err, cols, cells = oerr.error_analytic(mesh, _nodal(mesh, f), f, 1.0)
    want = 0.0
    for ck, kks in meshq.leaf_order(mesh):
        col = mesh.cols[ck]
        x0, y0, x1, y1 = col.pos
        X = quad.push(x0, x1, quad.lgl(col.ndofs[0])[0])
        c = (y1 - y0) * 2.0 * math.pi * _monic_nodal_sq_integral(X, x0, x1)
        assert abs(cols[ck] ** 2 - c) < 1e-12 * max(c, 1e-30)
        want += c
    assert abs(err ** 2 - want) < 1e-12 * want
    g = lambda x, y, th: th ** 3 + 0.0 * x + 0.0 * y
    err, cols, cells = oerr.error_analytic(mesh, _nodal(mesh, g), g, 1.0)
    for (ck, kk), e in cells.items():
        col = mesh.cols[ck]
        cell = col.cells[kk]
        x0, y0, x1, y1 = col.pos
        T = quad.push(cell.pos[0], cell.pos[1], quad.lg(cell.ndofs[0])[0])
        c = (x1 - x0) * (y1 - y0) * _monic_nodal_sq_integral(T, *cell.pos)
        assert abs(e ** 2 - c) < 1e-11 * max(c, 1e-30), (ck, kk, e ** 2, c)


def test_high_resolution_rule_has_n_plus_3_nodes():
    """u = x^6 against uh = 0 on one column with 2 nodes: u^2 = x^12 is above
    the exactness of the 5-point rule, so the error is that rule's sum
    (closed-form 5-point LGL nodes and weights), not the exact integral;
    likewise y^6.  A rule with another node count gives another value."""
    from oracle import meshq
    mesh = problems.uniform_mesh([1.0, 1.0], [2, 2, 2], 0, 0)
    N = meshq.dof_offsets(mesh)[1]
    xi = np.array([-1.0, -math.sqrt(3 / 7), 0.0, math.sqrt(3 / 7), 1.0])
    wi = np.array([1 / 10, 49 / 90, 32 / 45, 49 / 90, 1 / 10])
    x = 0.5 * (xi + 1.0)
    want = 2.0 * math.pi * float(np.sum(0.5 * wi * x ** 12))
    assert abs(want - 2.0 * math.pi / 13.0) > 1e-4      # the rule is not exact here
    for f in (lambda x, y, th: x ** 6 + 0.0 * y + 0.0 * th,
              lambda x, y, th: y ** 6 + 0.0 * x + 0.0 * th):
        err, _, _ = oerr.error_analytic(mesh, np.zeros(N), f, 1.0)
        assert abs(err ** 2 - want) < 1e-13
