"""Dirac-delta boundary data (``src/rt/boundary_conditions_vector.py:142-196``).

The product's vectorised ``rt.problem.dirac_bcs_vector`` is compared with the
oracle's literal restatement (``oracle.loops.dirac_bcs``), and both are pinned
by what the reference's branch fixes independently of either restatement:

* ``any(dirac)`` is the trigger: a delta at coordinate 0.0 alone gives the
  nodal vector (``oracle.loops.bcs``);
* cells whose interval holds no delta keep their nodal values;
* a delta on a cell's node turns the Lagrange factor into a Kronecker delta;
* a delta inside a cell weights the nodal value at the delta's coordinate by
  max(0, l_i(xi)), with l_i from numpy's barycentric form;
* with all three coordinates given, a cell none of whose intervals holds its
  delta is skipped without advancing the column's write index, so the next
  written cell lands at its place and the column's tail stays zero.
"""
import numpy as np
import pytest
from scipy.interpolate import BarycentricInterpolator

from oracle import loops
from oracle.meshq import dof_offsets, leaf_order
from oracle.quad import lgl, lg, push, pull
from rtx import problems


def _bcs(x, y, th):
    return (1.0 + x) * (2.0 + y) * (3.0 + np.cos(th))


def _lag_bary(nodes, xi):
    """l_i(xi) for every i, by barycentric interpolation of unit data."""
    n = len(nodes)
    return np.array([BarycentricInterpolator(nodes, np.eye(n)[i])(xi) for i in range(n)])


@pytest.fixture(scope="module")
def mesh():
    return problems.uniform_mesh([1.0, 1.0], [3, 3, 3], 2, 2)


def _product(mesh, dirac):
    from rt.problem import dirac_bcs_vector
    return dirac_bcs_vector(mesh, _bcs, dirac)


CASES = [[None, None, 0.3], [0.5, None, None], [0.37, None, 2.0], [0.3, 0.6, 2.0],
         [None, 0.25, None], [0.3, 0.6, 1.0]]


@pytest.mark.parametrize("dirac", CASES, ids=[str(c) for c in CASES])
def test_product_matches_oracle_loops(mesh, dirac):
    got = _product(mesh, dirac)
    ref = loops.dirac_bcs(mesh, _bcs, dirac)
    assert got.shape == ref.shape
    assert np.abs(got - ref).max() <= 1e-15 * np.abs(ref).max()


def test_delta_at_zero_takes_the_nodal_branch(mesh):
    nodal = loops.bcs(mesh, _bcs)
    # the reference tests any(dirac): a lone 0.0 is falsy
    assert np.array_equal(loops.dirac_bcs(mesh, _bcs, [0.0, None, None]), nodal)
    assert np.array_equal(loops.dirac_bcs(mesh, _bcs, [None, 0.0, 0.0]), nodal)
    # a planted "is not None" trigger would take the delta branch and differ
    assert not np.allclose(loops.dirac_bcs(mesh, _bcs, [1e-300, None, None]), nodal)


def test_theta_delta_inside_cells(mesh):
    ths = 0.3
    v = loops.dirac_bcs(mesh, _bcs, [None, None, ths])
    nodal = loops.bcs(mesh, _bcs)
    offs, _ = dof_offsets(mesh)
    hit = 0
    for ck, kks in leaf_order(mesh):
        col = mesh.cols[ck]
        nx, ny = col.ndofs
        xf = push(col.pos[0], col.pos[2], lgl(nx)[0])
        yf = push(col.pos[1], col.pos[3], lgl(ny)[0])
        for kk in kks:
            cell = col.cells[kk]
            nth = cell.ndofs[0]
            o = offs[(ck, kk)]
            blk = v[o:o + nx * ny * nth].reshape(nx, ny, nth)
            th0, th1 = cell.pos
            if th0 <= ths <= th1:
                hit += 1
                lth = np.maximum(0.0, _lag_bary(lg(nth)[0], pull(th0, th1, ths)))
                X, Y = np.meshgrid(xf, yf, indexing="ij")
                exp = _bcs(X, Y, ths)[:, :, None] * lth[None, None, :]
                assert np.abs(blk - exp).max() <= 1e-14 * np.abs(exp).max()
            else:
                assert np.abs(blk.ravel() - nodal[o:o + blk.size]).max() <= 1e-15 * 30
    assert hit == 16        # one angular cell per column holds theta = 0.3


def test_x_delta_on_a_column_face_is_a_kronecker_delta(mesh):
    xs = 0.5                # the shared face of the two middle column rows
    v = loops.dirac_bcs(mesh, _bcs, [xs, None, None])
    nodal = loops.bcs(mesh, _bcs)
    offs, _ = dof_offsets(mesh)
    touched = 0
    for ck, kks in leaf_order(mesh):
        col = mesh.cols[ck]
        x0, x1 = col.pos[0], col.pos[2]
        nx, ny = col.ndofs
        for kk in kks:
            nth = col.cells[kk].ndofs[0]
            o = offs[(ck, kk)]
            blk = v[o:o + nx * ny * nth].reshape(nx, ny, nth)
            ref = nodal[o:o + nx * ny * nth].reshape(nx, ny, nth)
            if x0 <= xs <= x1:
                node = 0 if x0 == xs else nx - 1
                assert np.abs(blk[node] - ref[node]).max() <= 1e-14 * np.abs(ref).max()
                others = [i for i in range(nx) if i != node]
                assert np.abs(blk[others]).max() <= 1e-15 * np.abs(ref).max()
                touched += 1
            else:
                assert np.array_equal(blk, ref)
    assert touched == 2 * 4 * 4          # two column rows of four, four cells each


def test_skipped_cells_shift_the_column_write_index(mesh):
    # all three given: a column holding neither xs nor ys writes only the
    # angular cell holding ths (the second of four here), at the first
    # cell's place; the rest of the column stays zero
    xs, ys, ths = 0.3, 0.6, 2.0
    v = loops.dirac_bcs(mesh, _bcs, [xs, ys, ths])
    offs, _ = dof_offsets(mesh)
    checked = 0
    for ck, kks in leaf_order(mesh):
        col = mesh.cols[ck]
        x0, y0, x1, y1 = col.pos
        if x0 <= xs <= x1 or y0 <= ys <= y1:
            continue
        nx, ny = col.ndofs
        cells = [col.cells[kk] for kk in kks]
        assert not (cells[0].pos[0] <= ths <= cells[0].pos[1])
        assert cells[1].pos[0] <= ths <= cells[1].pos[1]
        nth = cells[1].ndofs[0]
        lth = np.maximum(0.0, _lag_bary(lg(nth)[0], pull(cells[1].pos[0], cells[1].pos[1], ths)))
        X, Y = np.meshgrid(push(x0, x1, lgl(nx)[0]), push(y0, y1, lgl(ny)[0]), indexing="ij")
        exp = (_bcs(X, Y, ths)[:, :, None] * lth[None, None, :]).ravel()
        o = offs[(ck, kks[0])]
        col_ndof = sum(nx * ny * c.ndofs[0] for c in cells)
        assert np.abs(v[o:o + exp.size] - exp).max() <= 1e-14 * np.abs(exp).max()
        assert not v[o + exp.size:o + col_ndof].any()
        checked += 1
    assert checked == 9                  # 3 x 3 columns miss both xs and ys
