"""Global DOF indexing, the inflow-boundary mask and interior/boundary
splitting (``src/dg/matrix``).

DOF order: columns by ascending key, cells by ascending key, then within a
cell beta(i, j, a) = nth * ny * i + nth * j + a (``get_idxs.py:11-39``).
A DOF is a boundary DOF when it sits on a domain-boundary face (no neighbour)
and its cell's quadrant streams into the domain through that face
(``get_masks.py:80-110``).  The reference's PETSc index-set splitting
(``split_matrix.py:15-49``, ``split_vector.py:15-48``,
``merge_vectors.py:14-56``) becomes fancy indexing of scipy/numpy objects.
"""
import numpy as np
import scipy.sparse as sp

__all__ = ["get_idx_map", "get_idx_inv_map", "get_col_idxs", "get_cell_idxs",
           "get_intr_mask", "get_bdry_mask", "split_matrix", "split_vector",
           "merge_vectors", "get_col_dof_offsets"]

# Quadrants whose directions enter the domain through face F (F = 0 right,
# 1 top, 2 left, 3 bottom).
INFLOW_QUADRANTS = {0: (1, 2), 1: (2, 3), 2: (3, 0), 3: (0, 1)}


def get_idx_map(ndof_x: int, ndof_y: int, ndof_th: int):
    def idx_map(ii, jj, aa):
        return ndof_th * ndof_y * ii + ndof_th * jj + aa
    return idx_map


def get_idx_inv_map(ndof_x: int, ndof_y: int, ndof_th: int):
    def idx_inv_map(idx):
        a = idx % ndof_th
        j = ((idx - a) // ndof_th) % ndof_y
        i = (idx - (ndof_th * j + a)) // (ndof_th * ndof_y)
        return [i, j, a]
    return idx_inv_map


def get_col_idxs(mesh) -> list:
    """[ncols, {col_key: position in ascending-key order}] over leaf columns."""
    col_idxs = {}
    for key in sorted(mesh.cols.keys()):
        if mesh.cols[key].is_lf:
            col_idxs[key] = len(col_idxs)
    return [len(col_idxs), col_idxs]


def get_cell_idxs(mesh, col_key: int) -> list:
    cells = mesh.cols[col_key].cells
    cell_idxs = {}
    for key in sorted(cells.keys()):
        if cells[key].is_lf:
            cell_idxs[key] = len(cell_idxs)
    return [len(cell_idxs), cell_idxs]


def get_col_dof_offsets(mesh) -> dict:
    """{col_key: first global DOF of the column} in ascending key order."""
    offs, n = {}, 0
    for key in sorted(mesh.cols.keys()):
        col = mesh.cols[key]
        offs[key] = n
        nx, ny = col.ndofs
        for cell in col.cells.values():
            n += nx * ny * cell.ndofs[0]
    return offs


def get_intr_mask(mesh, **kwargs) -> np.ndarray:
    """Boolean mask, True for interior DOFs (``get_masks.py:22-129``)."""
    parts = []
    for key in sorted(mesh.cols.keys()):
        col = mesh.cols[key]
        if not col.is_lf:
            continue
        nx, ny = col.ndofs
        open_faces = [col.nhbr_keys[F][0] is None and col.nhbr_keys[F][1] is None
                      for F in range(4)]
        for ckey in sorted(col.cells.keys()):
            cell = col.cells[ckey]
            if not cell.is_lf:
                continue
            nth = cell.ndofs[0]
            m = np.ones((nx, ny, nth), dtype=bool)
            q = cell.quadrant
            if open_faces[0] and q in INFLOW_QUADRANTS[0]:
                m[nx - 1, :, :] = False
            if open_faces[1] and q in INFLOW_QUADRANTS[1]:
                m[:, ny - 1, :] = False
            if open_faces[2] and q in INFLOW_QUADRANTS[2]:
                m[0, :, :] = False
            if open_faces[3] and q in INFLOW_QUADRANTS[3]:
                m[:, 0, :] = False
            parts.append(m.ravel())
    return np.concatenate(parts) if parts else np.zeros(0, dtype=bool)


def get_bdry_mask(mesh, **kwargs) -> np.ndarray:
    return np.logical_not(get_intr_mask(mesh, **kwargs))


def split_matrix(matrix, intr_mask: np.ndarray) -> list:
    """[A[I, I], A[I, B]] as CSR."""
    A = sp.csr_matrix(matrix)
    intr = np.flatnonzero(intr_mask)
    bdry = np.flatnonzero(~np.asarray(intr_mask, dtype=bool))
    A_I = A[intr, :]
    return [A_I[:, intr].tocsr(), A_I[:, bdry].tocsr()]


def split_vector(vector, intr_mask: np.ndarray) -> list:
    v = np.asarray(vector)
    m = np.asarray(intr_mask, dtype=bool)
    return [v[m].copy(), v[~m].copy()]


def merge_vectors(vector_intr, vector_bdry, intr_mask: np.ndarray) -> np.ndarray:
    m = np.asarray(intr_mask, dtype=bool)
    out = np.zeros(m.size)
    out[m] = np.asarray(vector_intr)
    out[~m] = np.asarray(vector_bdry)
    return out
