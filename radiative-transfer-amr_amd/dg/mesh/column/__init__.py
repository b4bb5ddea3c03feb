"""Spatial column of the hp mesh (``src/dg/mesh/column``).

A column covers [x0, x1] x [y0, y1] (``pos = [x0, y0, x1, y1]``), has LGL
orders ``ndofs = [nx, ny]``, a dict of angular cells and, per face
F = 0 right, 1 top, 2 left, 3 bottom, the keys of its one or two face
neighbours ordered counter-clockwise (``[None, None]`` on the domain
boundary).
"""
from . import cell
from .cell import Cell


def calc_key(idx: list, lv: int) -> int:
    """Column key = sum_{l<lv} 4^l + 2^lv * idx[1] + idx[0]
    (``column/calc_key.py:9-17``); idx[0] counts along y, idx[1] along x."""
    return ((1 << (2 * lv)) - 1) // 3 + (1 << lv) * idx[1] + idx[0]


class Column:
    def __init__(self, pos: list, idx: list, lv: int, is_lf: bool,
                 ndofs: list, cells: dict, nhbr_keys: list):
        self.pos = pos
        self.idx = idx
        self.lv = lv
        self.key = calc_key(idx, lv)
        self.is_lf = is_lf
        self.ndofs = ndofs
        self.cells = cells
        self.nhbr_keys = nhbr_keys

    def add_cell(self, cell: Cell) -> None:
        self.cells[cell.key] = cell

    def del_cell(self, cell_key: int) -> None:
        del self.cells[cell_key]

    def nhbr_cell(self, cell_key: int, nhbr_loc: str = "+"):
        """Leaf angular neighbour of a cell inside this column, looking at the
        same level, then the parent, then the adjacent child
        (``column/nhbr_cell.py:12-62``).  Angles wrap around 2 pi."""
        cell = self.cells[cell_key]
        step = 1 if nhbr_loc == "+" else -1
        lv = cell.lv
        nidx = (cell.idx + step) % (1 << lv)
        cand = self.cells.get(calc_cell_key(nidx, lv))
        if cand is not None and cand.is_lf:
            return cand
        cand = self.cells.get(calc_cell_key(nidx // 2, lv - 1))
        if cand is not None and cand.is_lf:
            return cand
        cidx = 2 * nidx if nhbr_loc == "+" else 2 * nidx + 1
        cand = self.cells.get(calc_cell_key(cidx, lv + 1))
        if cand is not None and cand.is_lf:
            return cand
        return None

    def __eq__(self, other):
        return (self.pos == other.pos and self.idx == other.idx
                and self.lv == other.lv and self.key == other.key
                and self.is_lf == other.is_lf and self.ndofs == other.ndofs
                and self.cells == other.cells
                and self.nhbr_keys == other.nhbr_keys)

    def __str__(self):
        return ("    Column:  {}, {}\n".format(self.idx, self.lv)
                + "       key:  {}\n".format(self.key)
                + "       pos:  {}\n".format(self.pos)
                + "     is_lf:  {}\n".format(self.is_lf)
                + "     ndofs:  {}\n".format(self.ndofs)
                + "     cells:  {}\n".format(list(self.cells.keys()))
                + " nhbr_keys:  {}\n".format(self.nhbr_keys))


from .cell import calc_key as calc_cell_key  # noqa: E402
