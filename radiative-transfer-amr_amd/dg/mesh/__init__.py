"""Angular-spatial hp mesh (``src/dg/mesh``).

The mesh is a dict of spatial columns keyed by ``calc_key(idx, lv)``; each
column holds a dict of angular cells.  Refinement keeps spatial and angular
neighbours within one level of each other (2:1 balance) by refining a
coarser neighbour first, exactly as the reference does
(``ref_col_spt.py:44-62``, ``ref_cell.py:70-98``), so the set of keys after a
refinement sequence matches the reference's.

Public surface (same names and arguments as the reference):
``Mesh(Ls, pbcs, ndofs, has_th)``, ``ref_mesh``, ``ref_col``,
``ref_col_spt``, ``ref_col_ang``, ``ref_cell``, ``nhbr_cells_in_nhbr_col``,
``nhbr_cols``, ``add_col``, ``del_col``, ``get_ndof``, ``to_file`` and the
module function ``from_file``.
"""
import copy
import json

import numpy as np

from . import column
from .column import Column, calc_key as calc_col_key
from .column.cell import Cell, calc_key as calc_cell_key

__all__ = ["Mesh", "from_file", "column", "Column", "Cell"]

PI = np.pi

# Children of a column are ordered bottom-left, top-left, bottom-right,
# top-right; child ii has index [2i + di, 2j + dj] with idx[0] along y.
_CHILD_OFFSETS = ((0, 0), (1, 0), (0, 1), (1, 1))
# On face F the two neighbour slots are ordered counter-clockwise; the child
# of a refined column that touches slot n of face F (ref_col_spt.py:152-165).
_FACE_SLOT_CHILD = {(0, 0): 2, (0, 1): 3, (1, 0): 3, (1, 1): 1,
                    (2, 0): 1, (2, 1): 0, (3, 0): 0, (3, 1): 2}
# For each child, which faces are shared with a sibling (and the sibling).
_SIBLING = {0: {0: 2, 1: 1}, 1: {0: 3, 3: 0}, 2: {1: 3, 2: 0}, 3: {2: 1, 3: 2}}
# For each child and outer face: which parent slot it inherits, and which
# sibling replaces a column that was its own neighbour (periodic edge case).
_OUTER = {0: {2: (1, 2), 3: (0, 1)}, 1: {1: (1, 0), 2: (0, 3)},
          2: {0: (0, 0), 3: (1, 3)}, 3: {0: (1, 1), 1: (0, 2)}}


class Mesh:
    def __init__(self, Ls: list, pbcs: list, ndofs: list = [2, 2, 2],
                 has_th: bool = False):
        self.Ls = Ls
        self.pbcs = pbcs
        self.has_th = has_th
        if has_th:
            cell = Cell(pos=[0, 2 * PI], idx=0, lv=0, is_lf=True,
                        ndofs=[ndofs[2]], quadrant=None, nhbr_keys=[0, 0])
        else:
            cell = Cell(pos=[0, 0], idx=0, lv=0, is_lf=True, ndofs=[1],
                        quadrant=None, nhbr_keys=[None, None])
        nhbr_keys = [[None, None] for _ in range(4)]
        if pbcs[0]:
            nhbr_keys[0] = [0, 0]
            nhbr_keys[2] = [0, 0]
        if pbcs[1]:
            nhbr_keys[1] = [0, 0]
            nhbr_keys[3] = [0, 0]
        col = Column(pos=[0, 0, Ls[0], Ls[1]], idx=[0, 0], lv=0, is_lf=True,
                     ndofs=list(ndofs[0:2]), cells={0: cell},
                     nhbr_keys=nhbr_keys)
        self.cols = {col.key: col}

    # ------------------------------------------------------------------ basic
    def add_col(self, col: Column) -> None:
        self.cols[col.key] = col

    def del_col(self, col_key: int) -> None:
        del self.cols[col_key]

    def get_ndof(self) -> int:
        """Total DOFs over leaf columns and cells (``get_ndof.py:9-21``)."""
        total = 0
        for col in self.cols.values():
            if not col.is_lf:
                continue
            nx, ny = col.ndofs
            for cell in col.cells.values():
                if cell.is_lf:
                    total += int(nx * ny * cell.ndofs[0])
        return int(total)

    def nhbr_cols(self, col_key: int) -> list:
        """Unique neighbour column keys per face."""
        col = self.cols[col_key]
        return [sorted({k for k in col.nhbr_keys[F] if k is not None})
                for F in range(4)]

    # ------------------------------------------------------------- neighbours
    def nhbr_cells_in_nhbr_col(self, col_key: int, cell_key: int,
                               nhbr_col_key: int) -> list:
        """Leaf cells of ``nhbr_col_key`` that overlap ``cell_key`` in angle:
        the same cell, else its parent, else its two children
        (``nhbr_cells_in_nhbr_col.py:14-65``).  Returns ``[k0, k1]`` with
        ``None`` for missing entries."""
        cell = self.cols[col_key].cells[cell_key]
        cells_1 = self.cols[nhbr_col_key].cells
        out = [None, None]
        same = cells_1.get(cell.key)
        if same is not None:
            if same.is_lf:
                out[0] = same.key
            return out
        parent_key = calc_cell_key(cell.idx // 2, cell.lv - 1)
        parent = cells_1.get(parent_key)
        if parent is not None and parent.is_lf:
            out[0] = parent_key
        if parent is None:
            c0 = cells_1.get(calc_cell_key(2 * cell.idx, cell.lv + 1))
            if c0 is not None:
                if c0.is_lf:
                    out[0] = c0.key
                c1 = cells_1.get(calc_cell_key(2 * cell.idx + 1, cell.lv + 1))
                if c1 is not None and c1.is_lf:
                    out[1] = c1.key
        return out

    # ------------------------------------------------------------- refinement
    def ref_mesh(self, kind: str = "all", form: str = "h") -> None:
        """Refine every column present at call time (``ref_mesh.py:9-17``)."""
        col_keys = sorted(self.cols.keys())
        if kind in ("ang", "all"):
            for key in col_keys:
                self.ref_col(key, kind="ang", form=form)
        if kind in ("spt", "all"):
            for key in col_keys:
                self.ref_col(key, kind="spt", form=form)

    def ref_col(self, col_key: int, kind: str = "all", form: str = "h") -> None:
        """Refine one column in angle and/or space (``ref_col.py:11-27``)."""
        if kind not in ("ang", "spt", "all"):
            raise ValueError("unsupported refinement kind - {}".format(kind))
        col = self.cols[col_key]
        assert col.is_lf
        if kind in ("ang", "all"):
            for cell_key in sorted(col.cells.keys()):
                self.ref_cell(col_key, cell_key, form=form)
        if kind in ("spt", "all"):
            self.ref_col_spt(col_key, form=form)

    def ref_col_ang(self, col_key: int) -> None:
        col = self.cols[col_key]
        assert col.is_lf
        for cell_key in sorted(col.cells.keys()):
            self.ref_cell(col_key, cell_key)

    def ref_col_spt(self, col_key: int, form: str = "h") -> None:
        if form not in ("h", "p", "hp"):
            raise ValueError("unsupported refinement form - {}".format(form))
        if form in ("p", "hp"):
            self._ref_col_spt_p(col_key)
        if form in ("h", "hp"):
            self._ref_col_spt_h(col_key)

    def _ref_col_spt_p(self, col_key: int) -> None:
        """Raise nx, ny by one; cascade to neighbours more than one order
        behind in x (``ref_col_spt.py:26-44``)."""
        col = self.cols[col_key]
        assert col.is_lf
        col.ndofs[0] += 1
        col.ndofs[1] += 1
        for F in range(4):
            for nkey in set(col.nhbr_keys[F]):
                if nkey is None:
                    continue
                nhbr = self.cols[nkey]
                assert nhbr.is_lf
                if abs(col.ndofs[0] - nhbr.ndofs[0]) > 1:
                    self._ref_col_spt_p(nkey)

    def _ref_col_spt_h(self, col_key: int) -> None:
        """Split a column into four children (``ref_col_spt.py:46-177``)."""
        col = self.cols[col_key]
        assert col.is_lf
        lv = col.lv
        # Coarser neighbours are split first to keep the 2:1 balance.
        for F in range(4):
            for nkey in set(col.nhbr_keys[F]):
                if nkey is None:
                    continue
                nhbr = self.cols[nkey]
                assert nhbr.is_lf
                if lv - nhbr.lv == 1:
                    self.ref_col(nkey, kind="spt")
        i, j = col.idx
        x0, y0, xf, yf = col.pos
        xm, ym = (x0 + xf) / 2.0, (y0 + yf) / 2.0
        spans = {0: [x0, y0, xm, ym], 1: [x0, ym, xm, yf],
                 2: [xm, y0, xf, ym], 3: [xm, ym, xf, yf]}
        idxs = [[2 * i + di, 2 * j + dj] for di, dj in _CHILD_OFFSETS]
        keys = [calc_col_key(ix, lv + 1) for ix in idxs]
        for ii in range(4):
            nk = [[None, None] for _ in range(4)]
            for F, sib in _SIBLING[ii].items():
                nk[F] = [keys[sib], keys[sib]]
            for F, (slot, self_sib) in _OUTER[ii].items():
                parent_nkey = col.nhbr_keys[F][slot]
                k = keys[self_sib] if parent_nkey == col_key else parent_nkey
                nk[F] = [k, k]
            self.add_col(Column(pos=spans[ii], idx=idxs[ii][:], lv=lv + 1,
                                is_lf=True, ndofs=col.ndofs[:],
                                cells=copy.deepcopy(col.cells), nhbr_keys=nk))
        # Point the outside neighbours at the new children.
        for F in range(4):
            for n in range(2):
                nkey = col.nhbr_keys[F][n]
                if nkey is None:
                    continue
                nhbr = self.cols[nkey]
                assert nhbr.is_lf
                child = keys[_FACE_SLOT_CHILD[(F, n)]]
                if nhbr.lv == lv:
                    nhbr.nhbr_keys[(F + 2) % 4][(n + 1) % 2] = child
                elif nhbr.lv - lv == 1:
                    nhbr.nhbr_keys[(F + 2) % 4][0] = child
                    nhbr.nhbr_keys[(F + 2) % 4][1] = child
        self.del_col(col_key)

    def ref_cell(self, col_key: int, cell_key: int, form: str = "h") -> None:
        if form not in ("h", "p", "hp"):
            raise ValueError("unsupported refinement form - {}".format(form))
        if form in ("p", "hp"):
            self._ref_cell_p(col_key, cell_key)
        if form in ("h", "hp"):
            self._ref_cell_h(col_key, cell_key)

    def _ref_cell_p(self, col_key: int, cell_key: int) -> None:
        """Raise a cell's angular order; cascade to spatial and angular
        neighbours more than one order behind (``ref_cell.py:28-67``)."""
        col = self.cols[col_key]
        cell = col.cells[cell_key]
        assert cell.is_lf
        cell.ndofs[0] += 1
        for F in range(4):
            for nkey in set(col.nhbr_keys[F]):
                if nkey is None:
                    continue
                nhbr_col = self.cols[nkey]
                assert nhbr_col.is_lf
                for ckey in set(self.nhbr_cells_in_nhbr_col(col_key, cell_key,
                                                            nkey)):
                    if ckey is None:
                        continue
                    ncell = nhbr_col.cells[ckey]
                    assert ncell.is_lf
                    if abs(cell.ndofs[0] - ncell.ndofs[0]) > 1:
                        self._ref_cell_p(nkey, ckey)
        for loc in ("+", "-"):
            ncell = col.nhbr_cell(cell.key, loc)
            if abs(cell.ndofs[0] - ncell.ndofs[0]) > 1:
                self._ref_cell_p(col_key, ncell.key)

    def _ref_cell_h(self, col_key: int, cell_key: int) -> None:
        """Split an angular cell in two (``ref_cell.py:69-167``)."""
        col = self.cols[col_key]
        cell = col.cells[cell_key]
        assert cell.is_lf
        idx, lv = cell.idx, cell.lv
        z0, zf = cell.pos
        for F in range(2):
            nkey = cell.nhbr_keys[F]
            if nkey is None:
                continue
            ncell = col.cells[nkey]
            assert ncell.is_lf
            if lv - ncell.lv == 1:
                self.ref_cell(col_key, nkey)
        for F in range(4):
            for nkey in col.nhbr_keys[F]:
                if nkey is None:
                    continue
                nhbr_col = self.cols[nkey]
                assert nhbr_col.is_lf
                for ckey in self.nhbr_cells_in_nhbr_col(col_key, cell_key,
                                                        nkey):
                    if ckey is None:
                        continue
                    ncell = nhbr_col.cells[ckey]
                    assert ncell.is_lf
                    if lv - ncell.lv == 1:
                        self.ref_cell(nkey, ckey)
        zm = (z0 + zf) / 2.0
        spans = [[z0, zm], [zm, zf]]
        if cell.quadrant is not None:
            quads = [cell.quadrant, cell.quadrant]
        else:
            bounds = [[0, PI / 2], [PI / 2, PI], [PI, 3 * PI / 2],
                      [3 * PI / 2, 2 * PI]]
            quads = [None, None]
            for ii, span in enumerate(spans):
                for S, (b0, b1) in enumerate(bounds):
                    if span[0] >= b0 and span[1] <= b1:
                        quads[ii] = S
        keys = [calc_cell_key(2 * idx, lv + 1), calc_cell_key(2 * idx + 1, lv + 1)]
        nks = [[None, keys[1]], [keys[0], None]]
        for F in range(2):
            nkey = cell.nhbr_keys[F]
            if nkey is None:
                continue
            if nkey == cell.key:
                nks[F][F] = nks[F][(F + 1) % 2]
            elif col.cells[nkey].is_lf:
                nks[F][F] = nkey
            else:
                raise RuntimeError("2-neighbour assumption violated while "
                                   "splitting a cell")
        for ii in range(2):
            col.add_cell(Cell(pos=spans[ii], idx=2 * idx + ii, lv=lv + 1,
                              is_lf=True, ndofs=cell.ndofs[:],
                              quadrant=quads[ii], nhbr_keys=nks[ii][:]))
        for F, nkey in enumerate(cell.nhbr_keys):
            if nkey is not None and nkey != cell.key:
                ncell = col.cells[nkey]
                assert ncell.is_lf
                ncell.nhbr_keys[(F + 1) % 2] = keys[F]
        col.del_cell(cell_key)

    # -------------------------------------------------------------------- I/O
    def to_file(self, file_path: str = "mesh.json") -> None:
        """JSON mesh file in the reference's layout (``to_file.py:10-49``)."""
        d = {"Ls": self.Ls, "pbcs": self.pbcs, "has_th": self.has_th,
             "cols": {}}
        for ck, col in self.cols.items():
            cd = {"pos": col.pos, "idx": col.idx, "lv": col.lv, "key": col.key,
                  "is_lf": col.is_lf, "ndofs": col.ndofs,
                  "nhbr_keys": col.nhbr_keys, "cells": {}}
            for kk, cell in col.cells.items():
                cd["cells"][kk] = {"pos": cell.pos, "idx": cell.idx,
                                   "lv": cell.lv, "key": cell.key,
                                   "is_lf": cell.is_lf, "ndofs": cell.ndofs,
                                   "quadrant": cell.quadrant,
                                   "nhbr_keys": cell.nhbr_keys}
            d["cols"][ck] = cd
        with open(file_path, "w") as fh:
            json.dump(d, fh)

    def __eq__(self, other):
        return (self.Ls == other.Ls and self.pbcs == other.pbcs
                and self.has_th == other.has_th and self.cols == other.cols)

    def __str__(self):
        return ("Ls     :  {}\n".format(self.Ls)
                + "pbcs   :  {}\n".format(self.pbcs)
                + "has_th :  {}\n".format(self.has_th)
                + "cols   :  {}\n".format(sorted(self.cols.keys())))


def from_file(file_path: str = "mesh.json") -> Mesh:
    """Read a mesh written by ``Mesh.to_file`` (``from_file.py:13-63``)."""
    with open(file_path, "r") as fh:
        d = json.load(fh)
    mesh = Mesh(Ls=d["Ls"], pbcs=d["pbcs"], has_th=d["has_th"])
    mesh.del_col(0)
    for cd in d["cols"].values():
        cells = {}
        for kk, c in cd["cells"].items():
            cells[int(kk)] = Cell(pos=c["pos"], idx=c["idx"], lv=c["lv"],
                                  is_lf=c["is_lf"], ndofs=c["ndofs"],
                                  quadrant=c["quadrant"],
                                  nhbr_keys=c["nhbr_keys"])
        mesh.add_col(Column(pos=cd["pos"], idx=cd["idx"], lv=cd["lv"],
                            is_lf=cd["is_lf"], ndofs=cd["ndofs"], cells=cells,
                            nhbr_keys=cd["nhbr_keys"]))
    return mesh
