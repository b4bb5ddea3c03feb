"""Interior system, right-hand side and reference solves for the oracle.

Follows ``src/rt/solve.py:49-100``: A = (BC - IC) + P, mask I from
``get_intr_mask``, A_II x_I = f_I - A_IB b_B, initial guess f_I, and the
merged solution u = merge(x_I, b_B) (``solve.py:131``).

The reference solves with PETSc GMRES(50) + bjacobi/ILU(0) at
rtol = atol = 1e-15.  PETSc is not available here, so iteration counts and
residual histories are "parity unpinned"; the solution itself is pinned by
``scipy.sparse.linalg.spsolve`` on the same assembled system (the exact
discrete solution the Krylov solve converges to).

Also holds the per-cell block-Jacobi reference (dense inverse of each cell's
interior diagonal block, identity on boundary DOFs) used to check the
device factorisation, and a numpy MLP used to check the device network.
"""
import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from .meshq import leaf_order, dof_offsets


def interior_system(A, f, b, mask):
    I = np.flatnonzero(mask)
    B = np.flatnonzero(~mask)
    A = sp.csr_matrix(A)
    AI = A[I, :]
    A_II = AI[:, I].tocsr()
    A_IB = AI[:, B].tocsr()
    rhs = f[I] - A_IB @ b[B]
    return A_II, rhs, I, B


def exact_solve(A, f, b, mask):
    A_II, rhs, I, B = interior_system(A, f, b, mask)
    x = spla.spsolve(A_II.tocsc(), rhs)
    u = np.zeros(mask.size)
    u[I] = x
    u[B] = b[B]
    return u, A_II, rhs


def cell_blocks(mesh):
    """[(first DOF, size)] of every leaf cell in global order."""
    offs, _ = dof_offsets(mesh)
    out = []
    for ck, kks in leaf_order(mesh):
        col = mesh.cols[ck]
        nx, ny = col.ndofs
        for kk in kks:
            out.append((offs[(ck, kk)], nx * ny * col.cells[kk].ndofs[0]))
    return out


def bj_block_dofs(mesh, split=None):
    """Global DOF indices of every block-Jacobi block, in cell order.  A cell
    is one block, or - when ``split`` is given and the cell has more than
    ``split`` DOFs - several blocks of g = max(1, split // (nx ny))
    consecutive angles each (all spatial nodes), in angle order."""
    offs, _ = dof_offsets(mesh)
    out = []
    for ck, kks in leaf_order(mesh):
        col = mesh.cols[ck]
        nsp = col.ndofs[0] * col.ndofs[1]
        for kk in kks:
            nth = col.cells[kk].ndofs[0]
            o = offs[(ck, kk)]
            g = max(1, split // nsp) if (split is not None and nsp * nth > split) else nth
            for a0 in range(0, nth, g):
                a = np.arange(a0, min(a0 + g, nth))
                s = np.arange(nsp)
                out.append((o + s[:, None] * nth + a[None, :]).ravel())
    return out


def block_jacobi_inverses(A, mesh, mask, split=None):
    """Dense inverse of each block-Jacobi block (``bj_block_dofs``) of A
    restricted to the interior DOFs, with identity rows/columns at boundary
    DOFs."""
    A = sp.csr_matrix(A)
    invs = []
    for idx in bj_block_dofs(mesh, split):
        blk = A[idx][:, idx].toarray()
        m = mask[idx]
        blk[~m, :] = 0.0
        blk[:, ~m] = 0.0
        blk[~m, ~m] = 1.0
        invs.append(np.linalg.inv(blk))
    return invs


def block_jacobi_apply(invs, mesh, r, split=None):
    z = np.zeros_like(r)
    for idx, inv in zip(bj_block_dofs(mesh, split), invs):
        z[idx] = inv @ r[idx]
    return z


def gmres_solve(A_II, rhs, tol=1e-8, restart=50, maxiter=5000, M=None):
    """scipy GMRES on the interior system; returns (x, iterations)."""
    its = [0]

    def cb(_):
        its[0] += 1

    x, info = spla.gmres(A_II, rhs, x0=None, rtol=tol, atol=0.0,
                         restart=restart, maxiter=maxiter, M=M,
                         callback=cb, callback_type="pr_norm")
    return x, its[0], info


# ---------------------------------------------------------------- MLP oracle
def mlp_forward(params, X):
    """params = [(W0, b0), ..., (WL, bL)], ReLU between layers, linear out.
    Returns (out, cache of pre-activations and activations)."""
    acts = [X]
    pre = []
    h = X
    for li, (W, b) in enumerate(params):
        z = h @ W.T + b
        pre.append(z)
        h = np.maximum(z, 0.0) if li < len(params) - 1 else z
        acts.append(h)
    return h, (pre, acts)


def mlp_backward(params, cache, dout):
    """Gradients of sum(dout * out) w.r.t. every W, b."""
    pre, acts = cache
    grads = [None] * len(params)
    g = dout
    for li in range(len(params) - 1, -1, -1):
        W, b = params[li]
        if li < len(params) - 1:
            g = g * (pre[li] > 0.0)
        grads[li] = (g.T @ acts[li], g.sum(axis=0))
        g = g @ W
    return grads


# ------------------------------------------------------- coarse-grid oracle
def coarse_map(mesh, n):
    """Per leaf column (global order): [(coarse cell gx * n + gy,
    area(col & cell) / area(cell), area(col & cell) / area(col))] for an
    n x n grid over the domain (brute force over all coarse cells)."""
    cols = [ck for ck, _ in leaf_order(mesh)]
    Lx = max(mesh.cols[ck].pos[2] for ck in cols)
    Ly = max(mesh.cols[ck].pos[3] for ck in cols)
    hx, hy = Lx / n, Ly / n
    out = []
    for ck in cols:
        x0, y0, x1, y1 = mesh.cols[ck].pos
        area = (x1 - x0) * (y1 - y0)
        ent = []
        for gx in range(n):
            for gy in range(n):
                ox = min(x1, (gx + 1) * hx) - max(x0, gx * hx)
                oy = min(y1, (gy + 1) * hy) - max(y0, gy * hy)
                if ox > 1e-12 and oy > 1e-12:
                    ent.append((gx * n + gy, ox * oy / (hx * hy), ox * oy / area))
        out.append(ent)
    return out


def _dof_loop(mesh):
    """Yields (column index, global DOF, cell, i, j, a, theta_a, weight) with
    weight = (dx dy dth / 8) w_i w_j w_a."""
    from .quad import lgl, lg, push
    offs, _ = dof_offsets(mesh)
    for ci, (ck, kks) in enumerate(leaf_order(mesh)):
        col = mesh.cols[ck]
        x0, y0, x1, y1 = col.pos
        nx, ny = col.ndofs
        _, wx = lgl(nx)
        _, wy = lgl(ny)
        for kk in kks:
            cell = col.cells[kk]
            nth = cell.ndofs[0]
            tb, wt = lg(nth)
            th = push(cell.pos[0], cell.pos[1], tb)
            dth = cell.pos[1] - cell.pos[0]
            o = offs[(ck, kk)]
            for i in range(nx):
                for j in range(ny):
                    for a in range(nth):
                        w = (x1 - x0) * (y1 - y0) * dth / 8.0 * wx[i] * wy[j] * wt[a]
                        yield ci, o + nth * ny * i + nth * j + a, cell, i, j, a, th[a], w


def coarse_moments(mesh, n, r):
    """[sum r, sum r cos th, sum r sin th] per coarse cell (3 n^2, channel
    major), each column's sums split over its coarse cells by area."""
    cmap = coarse_map(mesh, n)
    col_m = np.zeros((len(cmap), 3))
    for ci, b, cell, i, j, a, th, w in _dof_loop(mesh):
        col_m[ci] += r[b] * np.array([1.0, np.cos(th), np.sin(th)])
    out = np.zeros(3 * n * n)
    for ci, ent in enumerate(cmap):
        for g, fcell, fcol in ent:
            for q in range(3):
                out[q * n * n + g] += fcol * col_m[ci, q]
    return out


def coarse_average(mesh, n, v):
    """Per coarse cell: sum over columns of area(col & cell) / area(cell)
    times the quadrature-weighted column mean of v."""
    cmap = coarse_map(mesh, n)
    num, den = np.zeros(len(cmap)), np.zeros(len(cmap))
    for ci, b, cell, i, j, a, th, w in _dof_loop(mesh):
        num[ci] += w * v[b]
        den[ci] += w
    out = np.zeros(n * n)
    for ci, ent in enumerate(cmap):
        for g, fcell, fcol in ent:
            out[g] += fcell * num[ci] / den[ci]
    return out


def coarse_prolong(mesh, n, phi, scale, mask):
    """z_beta = scale * sum_e fcol_e phi[g_e] of the DOF's column, 0 where
    mask is False."""
    cmap = coarse_map(mesh, n)
    uc = [scale * sum(fcol * phi[g] for g, _, fcol in ent) for ent in cmap]
    offs, N = dof_offsets(mesh)
    z = np.zeros(N)
    for ci, b, cell, i, j, a, th, w in _dof_loop(mesh):
        z[b] = uc[ci] if mask[b] else 0.0
    return z
