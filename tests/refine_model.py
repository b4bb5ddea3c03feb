"""Geometric model of the reference's hp refinement rules (test helper).

An independent restatement of the refinement semantics of
``src/dg/mesh/ref_col_spt.py:13-177`` and ``ref_cell.py:15-167`` that
keeps no neighbour-key bookkeeping at all: columns are (lv, ix, iy) squares
of a quadtree and cells are (lv, idx) arcs of a binary tree on the circle,
and every neighbour query is answered from geometry (shared edge of positive
length, overlapping angular arcs, adjacent arcs with wrap-around).

Rules modelled (the reference's):
  * spatial p: nx, ny += 1, then every face-neighbour column whose nx
    differs by more than one is p-refined (recursively);
  * spatial h: every face-neighbour column one level coarser is h-refined
    first, then the column splits into four children that inherit orders
    and cells;
  * angular p: nth += 1, then spatially overlapping cells in face-neighbour
    columns and the two angular neighbours are p-refined when their orders
    differ by more than one;
  * angular h: angular neighbours one level coarser, then overlapping cells
    one level coarser in face-neighbour columns, are h-refined first; the
    children inherit the order and get the quadrant that contains them.
Keys: column (4^lv - 1)/3 + 2^lv ix + iy (ix along x, iy along y), cell
2^lv - 1 + idx.
"""


def col_key(lv, ix, iy):
    return (4 ** lv - 1) // 3 + (2 ** lv) * ix + iy


def cell_key(lv, idx):
    return 2 ** lv - 1 + idx


class RefineModel:
    def __init__(self, nx, ny, nth):
        self.cols = {(0, 0, 0): {"nd": [nx, ny], "cells": {(0, 0): nth}}}

    # ------------------------------------------------------------ geometry
    @staticmethod
    def _box(c, L):
        lv, ix, iy = c
        s = 1 << (L - lv)
        return ix * s, iy * s, (ix + 1) * s, (iy + 1) * s

    def col_nbrs(self, c, F):
        out = []
        for d in self.cols:
            if d == c:
                continue
            L = max(c[0], d[0])
            x0, y0, x1, y1 = self._box(c, L)
            a0, b0, a1, b1 = self._box(d, L)
            oy = min(y1, b1) - max(y0, b0)
            ox = min(x1, a1) - max(x0, a0)
            if ((F == 0 and a0 == x1 and oy > 0) or (F == 1 and b0 == y1 and ox > 0)
                    or (F == 2 and a1 == x0 and oy > 0) or (F == 3 and b1 == y0 and ox > 0)):
                out.append(d)
        return sorted(out)

    @staticmethod
    def _arc(k, L):
        lv, idx = k
        s = 1 << (L - lv)
        return idx * s, (idx + 1) * s

    def overlapping(self, d, k):
        out = []
        for m in self.cols[d]["cells"]:
            L = max(k[0], m[0])
            a0, a1 = self._arc(k, L)
            b0, b1 = self._arc(m, L)
            if min(a1, b1) - max(a0, b0) > 0:
                out.append(m)
        return sorted(out)

    def ang_nbrs(self, c, k):
        """Leaf arcs adjacent at the - and + ends (wrap-around)."""
        res = []
        for side in (-1, +1):
            for m in self.cols[c]["cells"]:
                L = max(k[0], m[0])
                R = 1 << L
                a0, a1 = self._arc(k, L)
                b0, b1 = self._arc(m, L)
                if (side < 0 and b1 % R == a0 % R) or (side > 0 and b0 % R == a1 % R):
                    res.append(m)
                    break
        return res

    # ------------------------------------------------------------ spatial
    def col_p(self, c):
        self.cols[c]["nd"] = [n + 1 for n in self.cols[c]["nd"]]
        for F in range(4):
            for d in self.col_nbrs(c, F):
                if d in self.cols and abs(self.cols[c]["nd"][0] - self.cols[d]["nd"][0]) > 1:
                    self.col_p(d)

    def col_h(self, c):
        for F in range(4):
            for d in self.col_nbrs(c, F):
                if d in self.cols and c[0] - d[0] == 1:
                    self.col_h(d)
        lv, ix, iy = c
        rec = self.cols.pop(c)
        for di in (0, 1):
            for dj in (0, 1):
                self.cols[(lv + 1, 2 * ix + di, 2 * iy + dj)] = {
                    "nd": list(rec["nd"]), "cells": dict(rec["cells"])}

    # ------------------------------------------------------------ angular
    def cell_p(self, c, k):
        cells = self.cols[c]["cells"]
        cells[k] += 1
        for F in range(4):
            for d in self.col_nbrs(c, F):
                for m in self.overlapping(d, k):
                    if abs(cells[k] - self.cols[d]["cells"][m]) > 1:
                        self.cell_p(d, m)
        for m in self.ang_nbrs(c, k):
            if m in cells and abs(cells[k] - cells[m]) > 1:
                self.cell_p(c, m)

    def cell_h(self, c, k):
        for m in self.ang_nbrs(c, k):
            if m in self.cols[c]["cells"] and k[0] - m[0] == 1:
                self.cell_h(c, m)
        for F in range(4):
            for d in self.col_nbrs(c, F):
                for m in self.overlapping(d, k):
                    if k[0] - m[0] == 1:
                        self.cell_h(d, m)
        cells = self.cols[c]["cells"]
        n = cells.pop(k)
        cells[(k[0] + 1, 2 * k[1])] = n
        cells[(k[0] + 1, 2 * k[1] + 1)] = n

    # ------------------------------------------------------------ driver API
    def find_col(self, key):
        for c in self.cols:
            if col_key(*c) == key:
                return c
        raise KeyError(key)

    def find_cell(self, c, key):
        for k in self.cols[c]["cells"]:
            if cell_key(*k) == key:
                return k
        raise KeyError(key)

    def ref_col(self, key, kind, form):
        c = self.find_col(key)
        if kind in ("ang", "all"):
            for kk in sorted(cell_key(*k) for k in self.cols[c]["cells"]):
                self.ref_cell(key, kk, form)
        if kind in ("spt", "all"):
            c = self.find_col(key)
            if form in ("p", "hp"):
                self.col_p(c)
            if form in ("h", "hp"):
                self.col_h(c)

    def ref_cell(self, key, ckey, form):
        c = self.find_col(key)
        k = self.find_cell(c, ckey)
        if form in ("p", "hp"):
            self.cell_p(c, k)
        if form in ("h", "hp"):
            self.cell_h(c, k)

    def summary(self):
        """{col_key: (nx, ny, {cell_key: (nth, quadrant)})}."""
        out = {}
        for (lv, ix, iy), rec in self.cols.items():
            cells = {}
            for (clv, idx), n in rec["cells"].items():
                q = (idx >> (clv - 2)) if clv >= 2 else None
                cells[cell_key(clv, idx)] = (n, q)
            out[col_key(lv, ix, iy)] = (rec["nd"][0], rec["nd"][1], cells)
        return out


def mesh_summary(mesh):
    out = {}
    for ck, col in mesh.cols.items():
        cells = {kk: (c.ndofs[0], c.quadrant) for kk, c in col.cells.items()}
        out[ck] = (col.ndofs[0], col.ndofs[1], cells)
    return out
