"""1-D quadrature rules and the Lagrange nodal basis.

Interface of ``src/dg/quadrature`` in the reference:

* ``quad_xyth(nnodes_x, nnodes_y, nnodes_th)`` (``quad_xyth.py:22-86``) returns
  ascending Legendre-Gauss-Lobatto nodes/weights in x and y and ascending
  Legendre-Gauss nodes/weights in theta, on [-1, 1], memoised per order.  One
  node gives the rule {0}, {2} in every direction (``quad_xyth.py:37-41``).
* ``lgl_quad`` / ``lg_quad`` (``lgl_quad.py:3-40``, ``lg_quad.py:3-49``).
* ``lag_eval(nodes, i, x)`` (``lag_eval.py:3-19``): the i-th Lagrange
  polynomial, defined as 0 outside [-1 - 1e-14, 1 + 1e-14].
* ``lag_ddx_eval(nodes, j, x)`` / ``lag_ddx(nodes)`` (``lag_ddx.py:5-62``).

The rules are computed here by Newton iteration on the three-term Legendre
recurrence (own implementation); the vectorised helpers
``lagrange_matrix`` / ``lagrange_deriv_matrix`` are what the device plan uses.
"""
import numpy as np

__all__ = ["quad_xyth", "lgl_quad", "lg_quad", "lag_eval", "lag_ddx",
           "lag_ddx_eval", "lagrange_matrix", "lagrange_deriv_matrix",
           "SUPPORT_TOL"]

SUPPORT_TOL = 1.0e-14   # lag_eval.py:8

_lgl_cache: dict = {}
_lg_cache: dict = {}


def _legendre_and_derivs(n: int, x: np.ndarray):
    """P_n(x), P_n'(x) and P_{n-1}(x) by the Bonnet recurrence."""
    p_prev = np.ones_like(x)
    if n == 0:
        return p_prev, np.zeros_like(x), np.zeros_like(x)
    p = x.copy()
    for k in range(1, n):
        p_prev, p = p, ((2 * k + 1) * x * p - k * p_prev) / (k + 1)
    # derivative from P_n' (x^2 - 1) = n (x P_n - P_{n-1}), valid off +-1
    with np.errstate(divide="ignore", invalid="ignore"):
        dp = n * (x * p - p_prev) / (x * x - 1.0)
    return p, dp, p_prev


def lg_quad(nnodes: int):
    """Legendre-Gauss nodes (roots of P_n) and weights, ascending on [-1, 1]."""
    n = int(nnodes)
    k = np.arange(1, n + 1)
    x = -np.cos(np.pi * (k - 0.25) / (n + 0.5))     # Tricomi initial guess
    for _ in range(100):
        p, dp, _ = _legendre_and_derivs(n, x)
        dx = p / dp
        x = x - dx
        if np.max(np.abs(dx)) < 1e-16:
            break
    p, dp, _ = _legendre_and_derivs(n, x)
    w = 2.0 / ((1.0 - x * x) * dp * dp)
    order = np.argsort(x)
    return [x[order], w[order]]


def lgl_quad(nnodes: int):
    """Legendre-Gauss-Lobatto nodes (+-1 and roots of P'_{n-1}) and weights."""
    n = int(nnodes)
    N = n - 1
    x = -np.cos(np.pi * np.arange(n) / N)           # Chebyshev-Lobatto guess
    inner = x[1:-1].copy()
    for _ in range(100):
        if inner.size == 0:
            break
        p, dp, _ = _legendre_and_derivs(N, inner)
        # (1 - x^2) P'' = 2 x P' - N (N + 1) P
        d2p = (2.0 * inner * dp - N * (N + 1) * p) / (1.0 - inner * inner)
        dx = dp / d2p
        inner = inner - dx
        if np.max(np.abs(dx)) < 1e-16:
            break
    x = np.concatenate([[-1.0], np.sort(inner), [1.0]])
    pN, _, _ = _legendre_and_derivs(N, x)
    w = 2.0 / (N * (N + 1) * pN * pN)
    return [x, w]


def quad_xyth(nnodes_x: int = 1, nnodes_y: int = 1, nnodes_th: int = 1):
    """[nodes_x, w_x, nodes_y, w_y, nodes_th, w_th] (``quad_xyth.py:22-86``)."""
    out = []
    for n, cache, rule in ((nnodes_x, _lgl_cache, lgl_quad),
                           (nnodes_y, _lgl_cache, lgl_quad),
                           (nnodes_th, _lg_cache, lg_quad)):
        n = int(n)
        if n < 1:
            raise ValueError("quadrature with fewer than one node requested "
                             "(n = {})".format(n))
        if n not in cache:
            if n == 1:
                cache[n] = (np.asarray([0.0]), np.asarray([2.0]))
            else:
                nodes, weights = rule(n)
                cache[n] = (nodes, weights)
        nodes, weights = cache[n]
        out += [nodes, weights]
    return out


def lag_eval(nodes, i: int, x: float) -> float:
    """i-th Lagrange polynomial on ``nodes`` at ``x``; 0 outside [-1, 1]."""
    if not (-1.0 - SUPPORT_TOL <= x <= 1.0 + SUPPORT_TOL):
        return 0.0
    res = 1.0
    xi = nodes[i]
    for j in range(len(nodes)):
        if j != i:
            res *= (x - nodes[j]) / (xi - nodes[j])
    return res


def lag_ddx_eval(nodes, j: int, x: float) -> float:
    """Exact derivative of the j-th Lagrange polynomial at ``x``."""
    n = len(nodes)
    res = 0.0
    for k in range(n):
        if k == j:
            continue
        prod = 1.0
        for m in range(n):
            if m != j and m != k:
                prod *= (x - nodes[m]) / (nodes[j] - nodes[m])
        res += prod / (nodes[j] - nodes[k])
    return res


def lag_ddx(nodes) -> np.ndarray:
    """dmat[j, i] = l_i'(nodes[j]) (``lag_ddx.py:5-17``)."""
    n = len(nodes)
    return np.array([[lag_ddx_eval(nodes, i, nodes[j]) for i in range(n)]
                     for j in range(n)])


def lagrange_matrix(nodes, pts) -> np.ndarray:
    """L[i, k] = l_i(pts[k]) with the [-1, 1] support rule of ``lag_eval``."""
    nodes = np.asarray(nodes, dtype=np.float64)
    pts = np.asarray(pts, dtype=np.float64)
    n = nodes.size
    L = np.ones((n, pts.size))
    for i in range(n):
        for j in range(n):
            if j != i:
                L[i] *= (pts - nodes[j]) / (nodes[i] - nodes[j])
    inside = (pts >= -1.0 - SUPPORT_TOL) & (pts <= 1.0 + SUPPORT_TOL)
    L[:, ~inside] = 0.0
    return L


def lagrange_deriv_matrix(nodes) -> np.ndarray:
    """D[q, j] = l_q'(nodes[j]) (the orientation used by the convection terms,
    ``interior_convection_matrix.py:167-176``)."""
    nodes = np.asarray(nodes, dtype=np.float64)
    n = nodes.size
    D = np.zeros((n, n))
    for q in range(n):
        for j in range(n):
            D[q, j] = lag_ddx_eval(nodes, q, nodes[j])
    return D
