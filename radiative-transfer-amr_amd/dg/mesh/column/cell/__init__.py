"""Angular cell of a spatial column (``src/dg/mesh/column/cell``).

A cell covers [th0, th1] of [0, 2 pi), carries its number of angular nodes,
its angular quadrant (set once the cell lies inside one quarter of the
circle, ``ref_cell.py:108-120``; ``None`` above that) and the keys of its
lower/upper angular neighbours in the same column.
"""
import numpy as np


def calc_key(idx: int, lv: int) -> int:
    """Cell key = 2^lv - 1 + idx (``column/cell/calc_key.py:9-17``)."""
    return (1 << lv) - 1 + idx if lv >= 0 else idx


class Cell:
    def __init__(self, pos: list, idx: int, lv: int, is_lf: bool, ndofs: list,
                 quadrant, nhbr_keys: list):
        self.pos = pos
        self.idx = idx
        self.lv = lv
        self.key = calc_key(idx, lv)
        self.is_lf = is_lf
        self.ndofs = ndofs
        self.quadrant = quadrant
        self.nhbr_keys = nhbr_keys

    def __eq__(self, other):
        return (self.pos == other.pos and self.idx == other.idx
                and self.lv == other.lv and self.key == other.key
                and self.is_lf == other.is_lf and self.ndofs == other.ndofs
                and self.quadrant == other.quadrant
                and self.nhbr_keys == other.nhbr_keys)

    def __str__(self):
        return ("   Cell  :  {}, {}\n".format(self.idx, self.lv)
                + "    key  :  {}\n".format(self.key)
                + "    pos  :  [{:3.2f} pi, {:3.2f} pi]\n".format(
                    self.pos[0] / np.pi, self.pos[1] / np.pi)
                + "    is_lf:  {}\n".format(self.is_lf)
                + "    ndofs:  {}\n".format(self.ndofs)
                + " quadrant:  {}\n".format(self.quadrant)
                + "nhbr_keys:  {}\n".format(self.nhbr_keys))
