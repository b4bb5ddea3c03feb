"""Quadrature and Lagrange basis for the oracle.

Follows ``src/dg/quadrature/quad_xyth.py:22-86`` (LGL in x/y, Gauss-Legendre
in theta, ascending, one-node rule {0},{2}), ``lag_eval.py:3-19`` (basis set
to 0 outside [-1-1e-14, 1+1e-14]) and ``lag_ddx.py:44-62`` (exact derivative).

Computed by a route independent of the product: Gauss nodes from
``numpy.polynomial.legendre.leggauss``; Lobatto interior nodes as eigenvalues
of the Jacobi matrix of the P^(1,1) Jacobi polynomials (whose roots are the
roots of P'_{n-1}), then one Newton polish.

Pins: closed-form nodes/weights for n = 2..5 (tests/test_oracle_pins.py),
exactness of n-point LGL on degree 2n-3 and LG on degree 2n-1 monomials,
SBP identity of the derivative matrix.
"""
import numpy as np
from numpy.polynomial import legendre as npleg

TOL_SUPPORT = 1.0e-14


def lg(n: int):
    if n == 1:
        return np.array([0.0]), np.array([2.0])
    x, w = npleg.leggauss(n)
    return np.asarray(x, dtype=float), np.asarray(w, dtype=float)


def lgl(n: int):
    if n == 1:
        return np.array([0.0]), np.array([2.0])
    if n == 2:
        return np.array([-1.0, 1.0]), np.array([1.0, 1.0])
    m = n - 2                      # number of interior nodes
    # Jacobi matrix of monic P^(1,1): b_k^2 = k (k + 2) / ((2k + 1)(2k + 3))
    k = np.arange(1, m, dtype=float)
    off = np.sqrt(k * (k + 2.0) / ((2.0 * k + 1.0) * (2.0 * k + 3.0)))
    J = np.diag(off, 1) + np.diag(off, -1)
    inner = np.sort(np.linalg.eigvalsh(J)) if m > 1 else np.array([0.0])
    # Newton polish on P'_{n-1}
    c = np.zeros(n)
    c[n - 1] = 1.0
    dc = npleg.legder(c)
    d2c = npleg.legder(dc)
    for _ in range(3):
        inner = inner - npleg.legval(inner, dc) / npleg.legval(inner, d2c)
    x = np.concatenate([[-1.0], inner, [1.0]])
    PN = npleg.legval(x, c)
    w = 2.0 / (n * (n - 1) * PN ** 2)
    return x, w


def lag(nodes, i, x):
    """l_i(x) on ``nodes`` with the reference's support rule."""
    if x < -1.0 - TOL_SUPPORT or x > 1.0 + TOL_SUPPORT:
        return 0.0
    r = 1.0
    for j in range(len(nodes)):
        if j != i:
            r *= (x - nodes[j]) / (nodes[i] - nodes[j])
    return r


def dlag(nodes, j, x):
    """l_j'(x), exact product-rule form."""
    n = len(nodes)
    s = 0.0
    for k in range(n):
        if k == j:
            continue
        p = 1.0
        for m in range(n):
            if m != j and m != k:
                p *= (x - nodes[m]) / (nodes[j] - nodes[m])
        s += p / (nodes[j] - nodes[k])
    return s


def lag_mat(nodes, pts):
    """M[i, k] = l_i(pts[k])."""
    return np.array([[lag(nodes, i, p) for p in pts] for i in range(len(nodes))])


def dmat(nodes):
    """D[q, j] = l_q'(nodes[j])."""
    n = len(nodes)
    return np.array([[dlag(nodes, q, nodes[j]) for j in range(n)] for q in range(n)])


def push(x0, x1, xi):
    return x0 + (x1 - x0) / 2.0 * (np.asarray(xi) + 1.0)


def pull(x0, x1, x):
    return -1.0 + 2.0 / (x1 - x0) * (np.asarray(x) - x0)
