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


def block_jacobi_inverses(A, mesh, mask):
    """Dense inverse of each cell's diagonal block of A restricted to the
    interior DOFs, with identity rows/columns at boundary DOFs."""
    A = sp.csr_matrix(A)
    invs = []
    for o, n in cell_blocks(mesh):
        blk = A[o:o + n, o:o + n].toarray()
        m = mask[o:o + n]
        blk[~m, :] = 0.0
        blk[:, ~m] = 0.0
        blk[~m, ~m] = 1.0
        invs.append(np.linalg.inv(blk))
    return invs


def block_jacobi_apply(invs, mesh, r):
    z = np.zeros_like(r)
    for (o, n), inv in zip(cell_blocks(mesh), invs):
        z[o:o + n] = inv @ r[o:o + n]
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
