"""Nodal fields on the hp mesh (``src/dg/projection``).

``Projection(mesh, func)`` stores, per leaf column and cell, the nodal values
``vals[nx, ny, nth]`` of ``func(x, y, th)`` at the LGL (space) x LG (angle)
nodes (``class_Projection.py:17-83``).  ``to_vector`` / ``from_vector`` use the
global DOF order of ``dg.matrix`` (``to_vector.py:11-44``,
``from_vector.py:11-40``); ``to_file`` writes the mesh as JSON and the vector
as headerless float64 (``to_file.py:10-19``), read back by ``from_file``.

``func`` is evaluated on broadcast node arrays when it accepts arrays and
returns a matching shape; otherwise (e.g. it branches on scalar
coordinates) it is called once per node like the reference does.
"""
import copy

import numpy as np

from ..mesh import Mesh, from_file as mesh_from_file
from ..quadrature import quad_xyth

__all__ = ["Projection", "from_file", "push_forward", "pull_back",
           "eval_nodal"]


def push_forward(x0: float, xf: float, nodes):
    """[-1, 1] -> [x0, xf] (``push_pull.py:10-17``)."""
    return x0 + (xf - x0) / 2.0 * (np.asarray(nodes) + 1.0)


def pull_back(x0: float, xf: float, nodes):
    """[x0, xf] -> [-1, 1] (``push_pull.py:19-26``)."""
    return -1.0 + 2.0 / (xf - x0) * (np.asarray(nodes) - x0)


def eval_nodal(func, *coords) -> np.ndarray:
    """Evaluate ``func`` on the tensor grid of the 1-D ``coords``.

    Tries one broadcast call first; falls back to per-point scalar calls when
    the callable does not accept arrays or returns a wrong shape."""
    shape = tuple(len(c) for c in coords)
    grids = []
    for d, c in enumerate(coords):
        s = [1] * len(coords)
        s[d] = len(c)
        grids.append(np.asarray(c, dtype=np.float64).reshape(s))
    try:
        with np.errstate(all="ignore"):
            v = func(*grids)
        v = np.broadcast_to(np.asarray(v, dtype=np.float64), shape)
        return np.array(v, dtype=np.float64)
    except Exception:
        out = np.empty(shape)
        for idx in np.ndindex(*shape):
            out[idx] = func(*(float(coords[d][idx[d]]) for d in range(len(coords))))
        return out


class Projection_Cell:
    def __init__(self, cell, vals: np.ndarray):
        self.pos = cell.pos[:]
        self.idx = cell.idx
        self.lv = cell.lv
        self.key = cell.key
        self.is_lf = cell.is_lf
        self.ndofs = cell.ndofs[:]
        self.quadrant = cell.quadrant
        self.nhbr_keys = cell.nhbr_keys
        self.vals = vals

    def __eq__(self, other):
        return (self.pos == other.pos and self.idx == other.idx
                and self.lv == other.lv and self.key == other.key
                and self.is_lf == other.is_lf and self.ndofs == other.ndofs
                and self.quadrant == other.quadrant
                and self.nhbr_keys == other.nhbr_keys
                and bool(np.all(self.vals == other.vals)))


class Projection_Column:
    def __init__(self, col, cells: dict):
        self.pos = col.pos[:]
        self.idx = col.idx[:]
        self.lv = col.lv
        self.key = col.key
        self.is_lf = col.is_lf
        self.ndofs = col.ndofs[:]
        self.nhbr_keys = col.nhbr_keys[:]
        self.cells = cells

    def __eq__(self, other):
        return (self.pos == other.pos and self.idx == other.idx
                and self.lv == other.lv and self.key == other.key
                and self.is_lf == other.is_lf and self.ndofs == other.ndofs
                and self.cells == other.cells
                and self.nhbr_keys == other.nhbr_keys)


class Projection:
    def __init__(self, mesh: Mesh, func=None):
        self.mesh = copy.deepcopy(mesh)
        if func is None:
            f = None
        elif not mesh.has_th:
            def f(x, y, th):
                return func(x, y)
        else:
            f = func
        self.cols = {}
        for ckey in sorted(self.mesh.cols.keys()):
            col = self.mesh.cols[ckey]
            assert col.is_lf
            x0, y0, x1, y1 = col.pos
            nx, ny = col.ndofs
            xb, _, yb, _, _, _ = quad_xyth(nnodes_x=nx, nnodes_y=ny)
            xf, yf = push_forward(x0, x1, xb), push_forward(y0, y1, yb)
            cells = {}
            for kkey in sorted(col.cells.keys()):
                cell = col.cells[kkey]
                assert cell.is_lf
                if self.mesh.has_th:
                    nth = cell.ndofs[0]
                    thb = quad_xyth(nnodes_th=nth)[4]
                    thf = push_forward(cell.pos[0], cell.pos[1], thb)
                else:
                    nth, thf = 1, np.zeros(1)
                if f is None:
                    vals = np.zeros((nx, ny, nth))
                else:
                    vals = eval_nodal(f, xf, yf, thf)
                cells[kkey] = Projection_Cell(cell, vals)
            self.cols[ckey] = Projection_Column(col, cells)

    def to_vector(self) -> np.ndarray:
        parts = []
        for ckey in sorted(self.cols.keys()):
            col = self.cols[ckey]
            for kkey in sorted(col.cells.keys()):
                parts.append(np.asarray(col.cells[kkey].vals, dtype=np.float64).ravel())
        return np.concatenate(parts) if parts else np.zeros(0)

    def from_vector(self, vec) -> "Projection":
        vec = np.asarray(vec, dtype=np.float64)
        g = 0
        for ckey in sorted(self.mesh.cols.keys()):
            col = self.mesh.cols[ckey]
            nx, ny = col.ndofs
            for kkey in sorted(col.cells.keys()):
                nth = col.cells[kkey].ndofs[0]
                n = nx * ny * nth
                self.cols[ckey].cells[kkey].vals = vec[g:g + n].reshape(nx, ny, nth).copy()
                g += n
        return self

    def to_file(self, file_path: str = "projection.npy", **kwargs) -> None:
        kw = {"write_mesh": True, "mesh_file_path": "mesh.json", **kwargs}
        if kw["write_mesh"]:
            self.mesh.to_file(kw["mesh_file_path"])
        self.to_vector().tofile(file_path)

    def col_intg_th(self, col_key: int) -> np.ndarray:
        """int over theta of the column's field at each spatial node
        (``col_intg_th.py:13-37``)."""
        col = self.cols[col_key]
        nx, ny = col.ndofs
        out = np.zeros((nx, ny))
        for kkey in sorted(col.cells.keys()):
            cell = col.cells[kkey]
            wth = quad_xyth(nnodes_th=cell.ndofs[0])[5]
            out += (cell.pos[1] - cell.pos[0]) / 2.0 * (cell.vals @ wth)
        return out

    def cell_intg_xy(self, col_key: int, cell_key: int) -> np.ndarray:
        """int over the column's area at each angular node
        (``cell_intg_xy.py:13-36``)."""
        col = self.cols[col_key]
        x0, y0, x1, y1 = col.pos
        nx, ny = col.ndofs
        _, wx, _, wy, _, _ = quad_xyth(nnodes_x=nx, nnodes_y=ny)
        w = np.outer(wx, wy)
        vals = col.cells[cell_key].vals
        return (x1 - x0) * (y1 - y0) / 4.0 * np.einsum("ij,ija->a", w, vals)

    def __eq__(self, other):
        return self.mesh == other.mesh and self.cols == other.cols


def from_file(mesh_file_path: str = "mesh.json",
              projection_file_path: str = "projection.npy") -> Projection:
    mesh = mesh_from_file(mesh_file_path)
    proj = Projection(mesh, None)
    proj.from_vector(np.fromfile(projection_file_path))
    return proj
