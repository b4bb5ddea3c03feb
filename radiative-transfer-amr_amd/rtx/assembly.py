"""Assembled matrices of the reference's methods, values from the device.

The sparsity patterns (integer work) are built here from the plan tables:
  M  : diagonal                        (``src/rt/mass_matrix.py:140-173``)
  IC : cell-local, derivative lines    (``interior_convection_matrix.py:178-237``)
  S, P : same spatial node, all angle pairs of a column
         (``scattering_matrix.py:142-237``, ``preconditioner_matrix.py:225-248``)
  BC : outflow face diagonal plus upwind face blocks per link
       (``boundary_convection_matrix.py:284-524``)
and every value is evaluated by ``rtx_assemble_entries`` on the device.
Row/column order is the reference's global DOF order; the result is a
``scipy.sparse.csr_matrix`` (duplicates summed, explicit zeros removed).
"""
import ctypes as C

import numpy as np
import scipy.sparse as sp
import torch

from . import _lib
from ._lib import check
from .operator import stream_handle

PART = {"M": 0, "S": 1, "IC": 2, "BC": 3, "P": 4}


def _cells_by_shape(T):
    shapes = {}
    for k in range(T.ncell):
        c = T.cell_col[k]
        key = (int(T.col_nx[c]), int(T.col_ny[c]), int(T.cell_nth[k]))
        shapes.setdefault(key, []).append(k)
    return {s: np.asarray(v, np.int32) for s, v in shapes.items()}


def _local_ic_pattern(nx, ny, nth):
    i, j, a, q = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nth), np.arange(ny),
                             indexing="ij")
    r1 = (i * ny + q) * nth + a
    c1 = (i * ny + j) * nth + a
    i, j, a, p = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nth), np.arange(nx),
                             indexing="ij")
    r2 = (p * ny + j) * nth + a
    c2 = (i * ny + j) * nth + a
    n = nx * ny * nth
    key = np.unique(np.concatenate([r1.ravel() * n + c1.ravel(), r2.ravel() * n + c2.ravel()]))
    return (key // n).astype(np.int32), (key % n).astype(np.int32)


def pattern(T, part):
    """(kr, lr, kc, lc, link) int32 arrays of the entries of ``part``."""
    kr, lr, kc, lc, lk = [], [], [], [], []
    if part == "M":
        for (nx, ny, nth), cells in _cells_by_shape(T).items():
            n = nx * ny * nth
            kr.append(np.repeat(cells, n))
            lr.append(np.tile(np.arange(n, dtype=np.int32), cells.size))
        kc, lc = kr, lr
    elif part == "IC":
        for (nx, ny, nth), cells in _cells_by_shape(T).items():
            r, c = _local_ic_pattern(nx, ny, nth)
            kr.append(np.repeat(cells, r.size))
            lr.append(np.tile(r, cells.size))
            lc.append(np.tile(c, cells.size))
        kc = kr
    elif part in ("S", "P"):
        for ci in range(T.ncol):
            nsp = int(T.col_nx[ci] * T.col_ny[ci])
            k0, k1 = T.col_cell_start[ci], T.col_cell_start[ci + 1]
            cells = np.arange(k0, k1, dtype=np.int32)
            nts = T.cell_nth[k0:k1]
            a_cell = np.repeat(cells, nts)
            a_nt = np.repeat(nts, nts)
            a_loc = np.concatenate([np.arange(nt, dtype=np.int32) for nt in nts])
            na = a_cell.size
            s = np.arange(nsp, dtype=np.int32)[:, None, None]
            R = np.arange(na)[None, :, None]
            A = np.arange(na)[None, None, :]
            shp = (nsp, na, na)
            kr.append(np.broadcast_to(a_cell[R], shp).ravel())
            lr.append(np.broadcast_to(s * a_nt[R] + a_loc[R], shp).ravel())
            kc.append(np.broadcast_to(a_cell[A], shp).ravel())
            lc.append(np.broadcast_to(s * a_nt[A] + a_loc[A], shp).ravel())
    elif part == "BC":
        for k in range(T.ncell):
            q = int(T.cell_quad[k])
            if q < 0:
                continue
            c = T.cell_col[k]
            nx, ny, nth = int(T.col_nx[c]), int(T.col_ny[c]), int(T.cell_nth[k])
            blk = np.zeros((nx, ny, nth), bool)
            for F in (q, (q + 1) % 4):
                if F == 0:
                    blk[nx - 1] = True
                elif F == 2:
                    blk[0] = True
                elif F == 1:
                    blk[:, ny - 1] = True
                else:
                    blk[:, 0] = True
            loc = np.flatnonzero(blk.ravel()).astype(np.int32)
            kr.append(np.full(loc.size, k, np.int32))
            lr.append(loc)
            kc.append(np.full(loc.size, k, np.int32))
            lc.append(loc)
            lk.append(np.full(loc.size, -1, np.int32))
        for l in range(T.nlink):
            k, s, F = int(T.link_target[l]), int(T.link_src[l]), int(T.link_face[l])
            c0, c1 = T.cell_col[k], T.cell_col[s]
            nx, ny, nth = int(T.col_nx[c0]), int(T.col_ny[c0]), int(T.cell_nth[k])
            nx1, ny1, nt1 = int(T.col_nx[c1]), int(T.col_ny[c1]), int(T.cell_nth[s])
            n0 = nx if F % 2 else ny
            n1 = nx1 if F % 2 else ny1
            m0, r = np.meshgrid(np.arange(n0), np.arange(nth), indexing="ij")
            m1, a = np.meshgrid(np.arange(n1), np.arange(nt1), indexing="ij")
            if F == 0:
                rows = ((nx - 1) * ny + m0) * nth + r
                cols = (0 * ny1 + m1) * nt1 + a
            elif F == 2:
                rows = (0 * ny + m0) * nth + r
                cols = ((nx1 - 1) * ny1 + m1) * nt1 + a
            elif F == 1:
                rows = (m0 * ny + ny - 1) * nth + r
                cols = (m1 * ny1 + 0) * nt1 + a
            else:
                rows = (m0 * ny + 0) * nth + r
                cols = (m1 * ny1 + ny1 - 1) * nt1 + a
            rr = np.repeat(rows.ravel(), cols.size).astype(np.int32)
            cc = np.tile(cols.ravel(), rows.size).astype(np.int32)
            kr.append(np.full(rr.size, k, np.int32))
            lr.append(rr)
            kc.append(np.full(rr.size, s, np.int32))
            lc.append(cc)
            lk.append(np.full(rr.size, l, np.int32))
    else:
        raise ValueError("unknown operator part {!r}".format(part))
    cat = (lambda v: np.ascontiguousarray(np.concatenate(v), dtype=np.int32)
           if len(v) else np.zeros(0, np.int32))
    return cat(kr), cat(lr), cat(kc), cat(lc), (cat(lk) if part == "BC" else None)


def assemble(op, part):
    """CSR matrix of operator ``part`` ("M", "S" (= the reference's -S),
    "IC", "BC", "P") with values evaluated on the device."""
    T = op.t
    kr, lr, kc, lc, lk = pattern(T, part)
    n = kr.size
    ncell = T.ncell
    sizes = (T.col_nx[T.cell_col] * T.col_ny[T.cell_col] * T.cell_nth).astype(np.int64)
    if n:
        if not (kr.min() >= 0 and kr.max() < ncell and kc.min() >= 0 and kc.max() < ncell):
            raise RuntimeError("assembly pattern: cell index out of range")
        if not (np.all(lr >= 0) and np.all(lr < sizes[kr]) and np.all(lc >= 0)
                and np.all(lc < sizes[kc])):
            raise RuntimeError("assembly pattern: local index out of range")
        if lk is not None and not (lk.min() >= -1 and lk.max() < T.nlink):
            raise RuntimeError("assembly pattern: link index out of range")
    dev = op.device
    tk = [torch.from_numpy(a).to(dev) for a in (kr, lr, kc, lc)]
    tl = torch.from_numpy(lk).to(dev) if lk is not None else None
    val = torch.empty(n, dtype=torch.float64, device=dev)
    vp = lambda t: C.c_void_p(t.data_ptr()) if t is not None else None
    check(op.lib.rtx_assemble_entries(op.plan, PART[part], n, vp(tk[0]), vp(tk[1]), vp(tk[2]),
                                      vp(tk[3]), vp(tl), vp(val), stream_handle()),
          "rtx_assemble_entries")
    v = val.cpu().numpy()
    off = T.cell_dof_off.astype(np.int64)
    rows = off[kr] + lr
    cols = off[kc] + lc
    A = sp.coo_matrix((v, (rows, cols)), shape=(T.ndof, T.ndof)).tocsr()
    A.sum_duplicates()
    A.eliminate_zeros()
    return A
