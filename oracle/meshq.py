"""Mesh queries restated for the oracle.

Reads any mesh object with the reference's data model (``mesh.cols`` dict of
columns with ``pos, idx, lv, ndofs, cells, nhbr_keys``; cells with
``pos, idx, lv, ndofs, quadrant, nhbr_keys``).

* ``leaf_order`` / ``dof_offsets``: ascending column keys, ascending cell keys
  (``src/rt/mass_matrix.py:54-74``, ``src/dg/matrix/get_idxs.py:42-80``).
* ``nhbr_cells``: cells of a neighbour column overlapping a cell in angle:
  same key, else parent, else both children
  (``src/dg/mesh/nhbr_cells_in_nhbr_col.py:14-65``).
* ``geometric_face_neighbours``: brute-force adjacency from positions, an
  independent check of the stored neighbour keys.
"""
import numpy as np


def cell_key(idx, lv):
    return (2 ** lv - 1 + idx) if lv >= 0 else idx


def leaf_order(mesh):
    out = []
    for ck in sorted(mesh.cols.keys()):
        col = mesh.cols[ck]
        if not col.is_lf:
            continue
        out.append((ck, [kk for kk in sorted(col.cells.keys())
                         if col.cells[kk].is_lf]))
    return out


def dof_offsets(mesh):
    """{(col_key, cell_key): first global DOF}, and the total."""
    offs, n = {}, 0
    for ck, kks in leaf_order(mesh):
        col = mesh.cols[ck]
        nx, ny = col.ndofs
        for kk in kks:
            offs[(ck, kk)] = n
            n += nx * ny * col.cells[kk].ndofs[0]
    return offs, n


def nhbr_cells(mesh, ck0, kk0, ck1):
    cell = mesh.cols[ck0].cells[kk0]
    cells1 = mesh.cols[ck1].cells
    if cell.key in cells1:
        return [cell.key] if cells1[cell.key].is_lf else []
    out = []
    pk = cell_key(cell.idx // 2, cell.lv - 1)
    if pk in cells1 and cells1[pk].is_lf:
        out.append(pk)
    c0 = cell_key(2 * cell.idx, cell.lv + 1)
    c1 = cell_key(2 * cell.idx + 1, cell.lv + 1)
    if c0 in cells1:
        if cells1[c0].is_lf:
            out.append(c0)
        if c1 in cells1 and cells1[c1].is_lf:
            out.append(c1)
    return out


def geometric_face_neighbours(mesh, ck, tol=1e-12):
    """Per face F, the sorted keys of leaf columns sharing a positive-length
    stretch of that face (non-periodic domains)."""
    x0, y0, x1, y1 = mesh.cols[ck].pos
    out = [[], [], [], []]
    for k2 in sorted(mesh.cols.keys()):
        if k2 == ck:
            continue
        a0, b0, a1, b1 = mesh.cols[k2].pos
        ov_y = min(y1, b1) - max(y0, b0)
        ov_x = min(x1, a1) - max(x0, a0)
        if ov_y > tol and abs(a0 - x1) < tol:
            out[0].append(k2)
        if ov_x > tol and abs(b0 - y1) < tol:
            out[1].append(k2)
        if ov_y > tol and abs(a1 - x0) < tol:
            out[2].append(k2)
        if ov_x > tol and abs(b1 - y0) < tol:
            out[3].append(k2)
    return out
