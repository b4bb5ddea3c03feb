"""Pins of the CPU oracle: closed forms, invariants and brute force.

Each check is chosen so a plausible slip in the oracle (a dropped factor,
a swapped operand, a wrong affine map, a missing sign) fails it.  None of
them re-types the oracle's own formula: they evaluate what the mathematics
fixes (closed-form nodes and weights, exact integrals of polynomials,
partition of unity, divergence-theorem cancellations, symmetry) or compare
two independent routes (literal loop nests vs vectorised assembly,
geometric adjacency vs stored neighbour keys).

Reference formulas being pinned: ``src/dg/quadrature/quad_xyth.py:22-86``,
``lag_eval.py:3-19``, ``lag_ddx.py:44-62``, ``src/rt/get_Ex.py:10-127``,
``get_Eth.py:9-150``, ``mass_matrix.py:140-173``,
``scattering_matrix.py:142-237``, ``interior_convection_matrix.py:161-243``,
``boundary_convection_matrix.py:258-537``, ``get_masks.py:80-110``,
``forcing_vector.py:141-153``.
"""
import math

import numpy as np
import pytest
from numpy.polynomial import polynomial as P
from scipy import integrate

from oracle import quad, emat, loops, assemble as asm, meshq, system
from rtx import problems

# ------------------------------------------------------------ helpers


def lagrange_poly(nodes, i):
    """Power-series coefficients of the i-th Lagrange polynomial (numpy
    polynomial algebra: an independent route from the oracle's products)."""
    c = np.array([1.0])
    for j, xj in enumerate(nodes):
        if j != i:
            c = P.polymul(c, np.array([-xj, 1.0]) / (nodes[i] - xj))
    return c


def poly_integral(c, a, b):
    ci = P.polyint(c)
    return P.polyval(b, ci) - P.polyval(a, ci)


# ------------------------------------------------------------ quadrature

LGL_CLOSED = {
    2: ([-1.0, 1.0], [1.0, 1.0]),
    3: ([-1.0, 0.0, 1.0], [1 / 3, 4 / 3, 1 / 3]),
    4: ([-1.0, -1 / math.sqrt(5), 1 / math.sqrt(5), 1.0], [1 / 6, 5 / 6, 5 / 6, 1 / 6]),
    5: ([-1.0, -math.sqrt(3 / 7), 0.0, math.sqrt(3 / 7), 1.0],
        [1 / 10, 49 / 90, 32 / 45, 49 / 90, 1 / 10]),
}
LG_CLOSED = {
    1: ([0.0], [2.0]),
    2: ([-1 / math.sqrt(3), 1 / math.sqrt(3)], [1.0, 1.0]),
    3: ([-math.sqrt(3 / 5), 0.0, math.sqrt(3 / 5)], [5 / 9, 8 / 9, 5 / 9]),
}


@pytest.mark.parametrize("n", sorted(LGL_CLOSED))
def test_lgl_closed_forms(n):
    x, w = quad.lgl(n)
    np.testing.assert_allclose(x, LGL_CLOSED[n][0], rtol=0, atol=2e-15)
    np.testing.assert_allclose(w, LGL_CLOSED[n][1], rtol=0, atol=2e-15)


@pytest.mark.parametrize("n", sorted(LG_CLOSED))
def test_lg_closed_forms(n):
    x, w = quad.lg(n)
    np.testing.assert_allclose(x, LG_CLOSED[n][0], rtol=0, atol=2e-15)
    np.testing.assert_allclose(w, LG_CLOSED[n][1], rtol=0, atol=2e-15)


def test_one_node_rule():
    for rule in (quad.lgl, quad.lg):
        x, w = rule(1)
        assert list(x) == [0.0] and list(w) == [2.0]


@pytest.mark.parametrize("n", range(2, 11))
def test_quadrature_exactness(n):
    """n-point LGL is exact to degree 2n-3 (and not 2n-2); n-point Gauss to
    2n-1 (and not 2n)."""
    xl, wl = quad.lgl(n)
    xg, wg = quad.lg(n)
    for k in range(0, 2 * n + 1):
        exact = (1.0 - (-1.0) ** (k + 1)) / (k + 1)
        errl = abs(np.dot(wl, xl ** k) - exact)
        errg = abs(np.dot(wg, xg ** k) - exact)
        if k <= 2 * n - 3:
            assert errl < 1e-13, (n, k, errl)
        if k == 2 * n - 2:
            assert errl > 1e-6, (n, k, errl)
        if k <= 2 * n - 1:
            assert errg < 1e-13, (n, k, errg)
        if k == 2 * n:
            assert errg > 1e-6, (n, k, errg)
    assert np.all(np.diff(xl) > 0) and np.all(np.diff(xg) > 0)


@pytest.mark.parametrize("n", range(1, 11))
def test_product_quadrature_matches_oracle(n):
    """The product computes nodes by Newton on the Legendre recurrence, the
    oracle by eigenvalues / numpy's leggauss: two independent routes."""
    from dg.quadrature import quad_xyth
    xb, wx, yb, wy, tb, wt = quad_xyth(nnodes_x=n, nnodes_y=n, nnodes_th=n)
    x, w = quad.lgl(n)
    t, v = quad.lg(n)
    np.testing.assert_allclose(xb, x, atol=4e-15)
    np.testing.assert_allclose(wx, w, atol=4e-15)
    np.testing.assert_allclose(yb, x, atol=4e-15)
    np.testing.assert_allclose(wy, w, atol=4e-15)
    np.testing.assert_allclose(tb, t, atol=4e-15)
    np.testing.assert_allclose(wt, v, atol=4e-15)


def test_lagrange_support_rule():
    """l_i is 0 outside [-1 - 1e-14, 1 + 1e-14] and the polynomial inside."""
    from dg.quadrature import lag_eval
    x, _ = quad.lgl(4)
    for f in (quad.lag, lag_eval):
        assert f(x, 3, 1.0 + 2e-14) == 0.0
        assert f(x, 0, -1.0 - 2e-14) == 0.0
        assert abs(f(x, 3, 1.0 + 5e-15) - 1.0) < 1e-13
        for i in range(4):
            c = lagrange_poly(x, i)
            for p in (-0.7, 0.1, 0.93):
                assert abs(f(x, i, p) - P.polyval(p, c)) < 1e-14


@pytest.mark.parametrize("n", range(2, 10))
def test_derivative_matrix(n):
    """D[q, j] = l_q'(xi_j): corners -n(n-1)/4 and n(n-1)/4, exact on
    polynomials of degree n-1 (transpose convention), SBP identity."""
    from dg.quadrature import lag_ddx
    x, w = quad.lgl(n)
    D = quad.dmat(x)
    assert abs(D[0, 0] + n * (n - 1) / 4.0) < 1e-11
    assert abs(D[n - 1, n - 1] - n * (n - 1) / 4.0) < 1e-11
    for k in range(n):
        u = x ** k
        du = k * x ** (k - 1) if k > 0 else np.zeros(n)
        assert np.max(np.abs(D.T @ u - du)) < 1e-11 * max(1, n * n)
    Q = np.diag(w) @ D.T
    B = np.zeros((n, n))
    B[0, 0], B[-1, -1] = -1.0, 1.0
    assert np.max(np.abs(Q + Q.T - B)) < 1e-12
    # lag_ddx keeps the reference's textbook orientation dmat[j, i] = l_i'(x_j)
    # (lag_ddx.py:5-17); the convection terms use its transpose
    from dg.quadrature import lagrange_deriv_matrix
    np.testing.assert_allclose(lag_ddx(x), D.T, atol=1e-11)
    np.testing.assert_allclose(lagrange_deriv_matrix(x), D, atol=1e-11)


# ------------------------------------------------------------ face matrices

def _half(lower):
    """Own interval [0, 1]; a child half [0, 0.5] or [0.5, 1]."""
    return (0.0, 0.5) if lower else (0.5, 1.0)


@pytest.mark.parametrize("n", [2, 3, 5, 8])
def test_E_space_same_level_equal_order_is_diag_w(n):
    E = emat.E_space(n, 0.0, 1.0, 2, n, 0.0, 1.0, 2)
    np.testing.assert_allclose(E, np.diag(quad.lgl(n)[1]), atol=1e-15)


@pytest.mark.parametrize("lower", [True, False])
@pytest.mark.parametrize("n", [2, 3, 4, 6])
def test_E_space_equal_order_level_jumps(n, lower):
    """Equal orders.  Own column finer (dlv = +1, own is a half of the
    neighbour): E[i, p] = w_p l_i((xi_p -+ 1) / 2).  Neighbour finer
    (dlv = -1): E[i, p] = 1/2 w_i l_p((xi_i -+ 1) / 2)."""
    x, w = quad.lgl(n)
    s = -1.0 if lower else 1.0
    a, b = _half(lower)
    E_fine = emat.E_space(n, a, b, 3, n, 0.0, 1.0, 2)
    E_coarse = emat.E_space(n, 0.0, 1.0, 2, n, a, b, 3)
    L = np.array([[P.polyval((x[p] + s) / 2.0, lagrange_poly(x, i)) for p in range(n)]
                  for i in range(n)])
    np.testing.assert_allclose(E_fine, L * w[None, :], atol=1e-14)
    Lc = np.array([[P.polyval((x[i] + s) / 2.0, lagrange_poly(x, p)) for p in range(n)]
                   for i in range(n)])
    np.testing.assert_allclose(E_coarse, 0.5 * w[:, None] * Lc, atol=1e-14)


@pytest.mark.parametrize("n0,n1", [(2, 3), (3, 5), (4, 7), (3, 2), (5, 3), (7, 4), (4, 4)])
@pytest.mark.parametrize("dlv", [0, 1, -1])
@pytest.mark.parametrize("lower", [True, False])
def test_E_space_row_and_column_sums(n0, n1, dlv, lower):
    """Partition of unity turns the sums of E into exact integrals of a
    single Lagrange polynomial, which numpy evaluates independently:

    * dlv in {0, +1}: column sums = w0 (n0 >= n1: collocated own weights;
      n0 < n1: the n1-point rule integrates l0_p exactly), and for n0 >= n1
      row sums = 2 int_{own part of neighbour} l1_i.
    * dlv = -1: row sums = 1/2 int l1_i = 1/2 w1_i; column sums =
      int_{child half} l0_p.
    Total 1^T E 1 = 2 for dlv in {0, +1} and 1 for dlv = -1 (the 1/2)."""
    if dlv == 0 and not lower:
        pytest.skip("one geometry for same-level faces")
    x0, w0 = quad.lgl(n0)
    x1, w1 = quad.lgl(n1)
    a, b = _half(lower)
    if dlv == 0:
        E = emat.E_space(n0, 0.0, 1.0, 2, n1, 0.0, 1.0, 2)
        seg = (-1.0, 1.0)
    elif dlv == 1:
        E = emat.E_space(n0, a, b, 3, n1, 0.0, 1.0, 2)
        seg = (-1.0, 0.0) if lower else (0.0, 1.0)   # own half in neighbour coords
    else:
        E = emat.E_space(n0, 0.0, 1.0, 2, n1, a, b, 3)
        seg = (-1.0, 0.0) if lower else (0.0, 1.0)   # neighbour half in own coords
    assert E.shape == (n1, n0)
    if dlv >= 0:
        np.testing.assert_allclose(E.sum(axis=0), w0, atol=1e-13)
        if n0 >= n1:
            rows = [2.0 / (seg[1] - seg[0]) * poly_integral(lagrange_poly(x1, i), *seg)
                    for i in range(n1)]
            np.testing.assert_allclose(E.sum(axis=1), rows, atol=1e-13)
        assert abs(E.sum() - 2.0) < 1e-13
    else:
        np.testing.assert_allclose(E.sum(axis=1), 0.5 * w1, atol=1e-13)
        cols = [poly_integral(lagrange_poly(x0, p), *seg) for p in range(n0)]
        np.testing.assert_allclose(E.sum(axis=0), cols, atol=1e-13)
        assert abs(E.sum() - 1.0) < 1e-13


@pytest.mark.parametrize("n0,n1", [(2, 3), (3, 5), (4, 7), (3, 2), (5, 3), (7, 4), (6, 5), (8, 9)])
@pytest.mark.parametrize("dlv", [0, 1, -1])
@pytest.mark.parametrize("lower", [True, False])
def test_E_space_mismatched_orders_entrywise(n0, n1, dlv, lower):
    """With n0 != n1 every branch's rule has max(n0, n1) LGL nodes, exact to
    degree 2 max - 3 >= n0 + n1 - 2, so each entry is the exact integral
    E[i, p] = (2 / |own|) int_{own & neighbour} phi1_i phi0_p dx of the
    physical Lagrange functions, computed here by numpy polynomial algebra
    on the physical nodes."""
    if dlv == 0 and not lower:
        pytest.skip("one geometry for same-level faces")
    a, b = _half(lower)
    if dlv == 0:
        own, nbr = (0.0, 1.0), (0.0, 1.0)
        E = emat.E_space(n0, 0.0, 1.0, 2, n1, 0.0, 1.0, 2)
    elif dlv == 1:
        own, nbr = (a, b), (0.0, 1.0)
        E = emat.E_space(n0, a, b, 3, n1, 0.0, 1.0, 2)
    else:
        own, nbr = (0.0, 1.0), (a, b)
        E = emat.E_space(n0, 0.0, 1.0, 2, n1, a, b, 3)
    X0 = own[0] + (own[1] - own[0]) * 0.5 * (quad.lgl(n0)[0] + 1.0)
    X1 = nbr[0] + (nbr[1] - nbr[0]) * 0.5 * (quad.lgl(n1)[0] + 1.0)
    lo, hi = max(own[0], nbr[0]), min(own[1], nbr[1])
    # a 16-point Gauss-Legendre rule on the overlap integrates degree 31
    # exactly; the Lagrange functions in product form on the physical nodes
    g, gw = np.polynomial.legendre.leggauss(16)
    xs = lo + (hi - lo) * 0.5 * (g + 1.0)
    gw = gw * 0.5 * (hi - lo)

    def lag_prod(nodes, i, x):
        v = np.ones_like(x)
        for j, xj in enumerate(nodes):
            if j != i:
                v *= (x - xj) / (nodes[i] - xj)
        return v
    ref = np.array([[2.0 / (own[1] - own[0]) *
                     np.sum(gw * lag_prod(X1, i, xs) * lag_prod(X0, p, xs))
                     for p in range(n0)] for i in range(n1)])
    np.testing.assert_allclose(E, ref, atol=1e-13)


def test_E_space_rejects_level_jump_of_two():
    with pytest.raises(ValueError):
        emat.E_space(3, 0.0, 0.25, 4, 3, 0.0, 1.0, 2)


@pytest.mark.parametrize("F", [0, 1, 2, 3])
@pytest.mark.parametrize("n", [2, 4, 6])
def test_E_theta_same_level_equal_order(F, n):
    """E^theta = diag(w_a Theta_F(theta_a)), Theta_F = cos(theta - F pi/2)."""
    a0, b0 = math.pi / 8, math.pi / 4
    t, w = quad.lg(n)
    th = a0 + (b0 - a0) * (t + 1) / 2
    E = emat.E_theta(n, a0, b0, 3, n, a0, b0, 3, F)
    d = w * np.cos(th - F * math.pi / 2)
    d[np.abs(d) < 1e-15] = 0.0
    np.testing.assert_allclose(E, np.diag(d), atol=1e-15)


@pytest.mark.parametrize("n0,n1", [(3, 5), (5, 3), (4, 4), (6, 2), (2, 6)])
@pytest.mark.parametrize("dlv", [0, 1, -1])
@pytest.mark.parametrize("F", [0, 1, 2, 3])
def test_E_theta_total_is_face_integral(n0, n1, dlv, F):
    """1^T E^theta 1 = (2 / dtheta_own) int_overlap cos(theta - F pi/2),
    up to Gauss quadrature error (tiny cells, smooth integrand).  A dropped
    1/2, a wrong Theta_F argument or a wrong child position fails this."""
    a0, b0 = 3 * math.pi / 4, math.pi          # level-3 cell
    ac, bc = 3 * math.pi / 4, 7 * math.pi / 8  # its lower child
    if dlv == 0:
        E = emat.E_theta(n0, a0, b0, 3, n1, a0, b0, 3, F)
        lo, hi, d_own = a0, b0, b0 - a0
    elif dlv == 1:   # own is the child
        E = emat.E_theta(n0, ac, bc, 4, n1, a0, b0, 3, F)
        lo, hi, d_own = ac, bc, bc - ac
    else:            # neighbour is the child
        E = emat.E_theta(n0, a0, b0, 3, n1, ac, bc, 4, F)
        lo, hi, d_own = ac, bc, b0 - a0
    exact = (2.0 / d_own) * (math.sin(hi - F * math.pi / 2) - math.sin(lo - F * math.pi / 2))
    assert abs(E.sum() - exact) < 1e-9 * max(1.0, abs(exact)), (E.sum(), exact)


def _bary(nodes, pts):
    """L[i, k] = l_i(pts[k]) by the barycentric formula (a route independent
    of the oracle's product form); exact node hits give the unit vector."""
    nodes = np.asarray(nodes, dtype=np.float64)
    wb = np.array([1.0 / np.prod([nodes[j] - nodes[k] for k in range(len(nodes)) if k != j])
                   for j in range(len(nodes))])
    L = np.zeros((len(nodes), len(pts)))
    for k, x in enumerate(pts):
        d = x - nodes
        hit = np.flatnonzero(d == 0.0)
        if hit.size:
            L[hit[0], k] = 1.0
        else:
            t = wb / d
            L[:, k] = t / t.sum()
    return L


def _eth_unified(n0, a0, b0, n1, a1, b1, F):
    """get_Eth.py:54-147 in one statement: with S the smaller of the own cell
    K = [a0, b0] and the neighbour K' = [a1, b1] and m = max(n0, n1),
    E[a, r] = (|S| / |K|) sum_p w_p Theta_F(theta_p) l1_a(xi'(theta_p))
    l0_r(xi(theta_p)), the m-point Gauss rule placed on S."""
    m = max(n0, n1)
    lo, hi = (a1, b1) if (b1 - a1) < (b0 - a0) else (a0, b0)
    xg, wg = np.polynomial.legendre.leggauss(m)
    th = lo + (hi - lo) / 2.0 * (xg + 1.0)
    x0 = np.polynomial.legendre.leggauss(n0)[0] if n0 > 1 else np.array([0.0])
    x1 = np.polynomial.legendre.leggauss(n1)[0] if n1 > 1 else np.array([0.0])
    # the cell's own nodes sit exactly on its reference rule: when S is that
    # cell and the rule has its node count, evaluate at the reference nodes
    xi0 = xg if (lo, hi) == (a0, b0) else 2.0 * (th - a0) / (b0 - a0) - 1.0
    xi1 = xg if (lo, hi) == (a1, b1) else 2.0 * (th - a1) / (b1 - a1) - 1.0
    L0, L1 = _bary(x0, xi0), _bary(x1, xi1)
    E = (hi - lo) / (b0 - a0) * (L1 * (wg * np.cos(th - F * math.pi / 2.0))) @ L0.T
    E[np.abs(E) < 1.0e-15] = 0.0
    return E


@pytest.mark.parametrize("F", [0, 1, 2, 3])
@pytest.mark.parametrize("case", ["same", "own_child_lo", "own_child_hi", "nbr_child_lo",
                                  "nbr_child_hi"])
def test_E_theta_entries_are_one_quadrature(case, F):
    """Every entry of every E^theta branch, mismatched orders included, is
    the one quadrature of _eth_unified: the reference's four branches differ
    only in which interval carries the max(n0, n1)-point rule and in the 1/2
    of a finer neighbour.  Fails for a wrong rule size, a rule on the wrong
    interval, a swapped map, a missing 1/2 or a transposed E."""
    a0, b0 = 3 * math.pi / 4, math.pi          # level-3 cell
    lo_c, mid, hi_c = a0, 7 * math.pi / 8, b0  # its children
    for n0 in range(1, 10):
        for n1 in range(1, 10):
            if case == "same":
                args = (n0, a0, b0, 3, n1, a0, b0, 3)
            elif case == "own_child_lo":
                args = (n0, lo_c, mid, 4, n1, a0, b0, 3)
            elif case == "own_child_hi":
                args = (n0, mid, hi_c, 4, n1, a0, b0, 3)
            elif case == "nbr_child_lo":
                args = (n0, a0, b0, 3, n1, lo_c, mid, 4)
            else:
                args = (n0, a0, b0, 3, n1, mid, hi_c, 4)
            E = emat.E_theta(*args, F)
            ref = _eth_unified(args[0], args[1], args[2], args[4], args[5], args[6], F)
            assert E.shape == (n1, n0)
            np.testing.assert_allclose(E, ref, rtol=0, atol=1e-14 * max(1.0, np.abs(ref).max()),
                                       err_msg=str((case, n0, n1, F)))


# ------------------------------------------------------------ matrices

def tiny_hp_mesh():
    """3x3 columns x 4 cells with spatial hp refinement (level and order
    jumps) and angular h / hp refinement: every E branch is exercised."""
    m = problems.uniform_mesh([1.0, 1.0], [2, 2, 2], 2, 1)
    keys = sorted(m.cols)
    m.ref_col(keys[0], "spt", "hp")
    last = sorted(m.cols)[-1]
    m.ref_cell(last, sorted(m.cols[last].cells)[0], "hp")
    first = sorted(m.cols)[0]
    m.ref_cell(first, sorted(m.cols[first].cells)[2], "h")
    return m


@pytest.fixture(scope="module")
def hp_mats():
    mesh = tiny_hp_mesh()
    wl = problems.two_scatterers()
    L = {
        "M": loops.mass(mesh, wl.kappa),
        "S": loops.scattering(mesh, wl.sigma, wl.Phi),
        "IC": loops.interior_conv(mesh),
        "BC": loops.boundary_conv(mesh),
        "P": loops.preconditioner(mesh, wl.kappa, wl.sigma, wl.Phi),
    }
    V = asm.assemble(mesh, wl.kappa, wl.sigma, wl.Phi)
    return mesh, wl, L, V


@pytest.mark.parametrize("part", ["M", "S", "IC", "BC", "P"])
def test_vectorised_assembly_matches_loops(hp_mats, part):
    """Two independent codings (literal loop nests vs vectorised) of every
    operator part on an hp mesh with level and order jumps."""
    _, _, L, V = hp_mats
    A, B = L[part].tocsr(), V[part].tocsr()
    assert A.shape == B.shape
    scale = abs(A).max()
    assert abs(A - B).max() <= 1e-13 * scale, part
    big = lambda X: int((abs(X) > 1e-12 * scale).sum())
    assert big(A) == big(B)


def test_vectorised_assembly_row_subset(hp_mats):
    mesh, wl, _, V = hp_mats
    keys = sorted(mesh.cols)[2:5]
    Vs = asm.assemble(mesh, wl.kappa, wl.sigma, wl.Phi, cols=keys)
    A = asm.system_matrix(V)
    As = asm.system_matrix(Vs)
    rows = np.unique(As.nonzero()[0])
    assert rows.size > 0
    assert abs(A[rows] - As[rows]).max() < 1e-14 * abs(A).max()


def test_scattering_is_symmetric(hp_mats):
    """S[(K,r),(K',a)] carries dtheta_K dtheta_K' w_r w_a Phi(theta_r,
    theta_a) with Phi symmetric: S = S^T on any angular mesh."""
    _, _, L, _ = hp_mats
    S = L["S"].tocsr()
    assert abs(S - S.T).max() < 1e-15 * abs(S).max() + 1e-300
    assert S.max() <= 0.0     # returned as -S


def test_mass_total_and_exactness():
    """Sum of M = 2 pi int int kappa, exact for polynomial kappa of degree
    <= 2n - 3 per direction."""
    mesh = problems.uniform_mesh([1.0, 2.0], [4, 4, 3], 2, 2)
    M = loops.mass(mesh, lambda x, y: 1.0 + x * y ** 2)
    exact = 2 * math.pi * (1.0 * 2.0 + (1.0 / 2.0) * (8.0 / 3.0))
    assert abs(M.diagonal().sum() - exact) < 1e-12 * exact
    assert (M - __import__("scipy.sparse", fromlist=["diags"]).diags(M.diagonal())).nnz == 0


def test_isotropic_scattering_row_sums():
    """Phi = 1/(2 pi), constant sigma: (M + S) 1 = (kappa - sigma) (M / kappa) 1
    (the angular integral of Phi is 1 on any angular mesh)."""
    mesh = tiny_hp_mesh()
    kap, sig = 1.7, 0.6
    M = loops.mass(mesh, lambda x, y: kap + 0 * x)
    S = loops.scattering(mesh, lambda x, y: sig + 0 * x, lambda a, b: 1.0 / (2 * math.pi))
    one = np.ones(M.shape[0])
    lhs = (M + S) @ one
    rhs = (kap - sig) / kap * (M @ one)
    assert np.max(np.abs(lhs - rhs)) < 1e-14 * np.max(np.abs(rhs))
    P = loops.preconditioner(mesh, lambda x, y: kap + 0 * x, lambda x, y: sig + 0 * x,
                             lambda a, b: 1.0 / (2 * math.pi))
    assert abs(P - (M + S)).max() < 1e-15


def test_hg_normalisation_and_first_moment():
    """Closed-form N_g = 4 E(m) / (1 - g) against adaptive quadrature; the
    normalised kernel integrates to 1 and its first moment is not g."""
    for g in (0.6, 0.7, 0.8, 0.9):
        ref, _ = integrate.quad(lambda t: (1 - g * g) / (1 + g * g - 2 * g * math.cos(t)) ** 1.5,
                                0, 2 * math.pi, epsabs=1e-14, epsrel=1e-14, limit=400)
        assert abs(problems.hg_norm(g) - ref) < 4e-14 * ref
        Phi = problems.henyey_greenstein(g)
        tot, _ = integrate.quad(lambda t: Phi(t, 0.3), 0, 2 * math.pi, epsabs=1e-14, limit=400)
        assert abs(tot - 1.0) < 1e-12
    assert abs(problems.hg_first_moment(0.6) - 0.785022975637) < 1e-11


def test_interior_convection_annihilates_constants_and_cell_sums():
    """sum_q l_q' = 0: IC 1 = 0 row-wise; 1^T IC = volume term of a
    constant test function."""
    mesh = tiny_hp_mesh()
    IC = loops.interior_conv(mesh)
    one = np.ones(IC.shape[0])
    # the test index q carries the derivative: sum over rows of one cell = 0
    for o, n in system.cell_blocks(mesh):
        assert abs((IC[o:o + n] @ one).sum()) < 1e-13


def test_convection_divergence_theorem_uniform_mesh():
    """Conforming uniform mesh: for every cell whose column has neighbours
    on all four faces, 1_K^T (BC - IC) 1 = 0 (opposite faces cancel)."""
    mesh = problems.uniform_mesh([1.0, 1.0], [3, 3, 3], 2, 2)
    BC, IC = loops.boundary_conv(mesh), loops.interior_conv(mesh)
    A = (BC - IC).tocsr()
    one = np.ones(A.shape[0])
    r = A @ one
    offs, _ = meshq.dof_offsets(mesh)
    checked = 0
    for ck, kks in meshq.leaf_order(mesh):
        col = mesh.cols[ck]
        if any(col.nhbr_keys[F][0] is None for F in range(4)):
            continue
        n = col.ndofs[0] * col.ndofs[1]
        for kk in kks:
            o = offs[(ck, kk)]
            assert abs(r[o:o + n * col.cells[kk].ndofs[0]].sum()) < 1e-14
            checked += 1
    assert checked == 4 * 4


def test_boundary_convection_upwinding_closed_form():
    """Uniform conforming mesh, u = indicator of one column C'.  BC u is
    (a) the own outflow face mass on C' and (b) on the neighbour C across
    face F of C (C' upwind of C), the rows of C's inflow-quadrant cells on
    the face shared with C': c_F w_j w_a Theta_F(theta_a) < 0.  Nothing
    else.  Reversing the upwind direction or the face map fails this."""
    mesh = problems.uniform_mesh([1.0, 1.0], [3, 3, 2], 2, 2)
    BC = loops.boundary_conv(mesh).tocsr()
    offs, n = meshq.dof_offsets(mesh)
    order = meshq.leaf_order(mesh)
    src = order[5][0]                                # an interior column
    u = np.zeros(n)
    for kk in order[5][1]:
        o = offs[(src, kk)]
        u[o:o + 9 * 2] = 1.0
    y = BC @ u
    expect = np.zeros(n)
    x, w = quad.lgl(3)
    t, v = quad.lg(2)
    for ck, kks in order:
        col = mesh.cols[ck]
        dx, dy = col.pos[2] - col.pos[0], col.pos[3] - col.pos[1]
        for F in range(4):
            if ck == src:
                outflow = True
            elif col.nhbr_keys[F][0] == src:
                outflow = False
            else:
                continue
            for kk in kks:
                cell = col.cells[kk]
                q = cell.quadrant
                is_out = F in (q, (q + 1) % 4)
                if is_out != outflow:
                    continue
                dth = cell.pos[1] - cell.pos[0]
                th = cell.pos[0] + dth * (t + 1) / 2
                c = (dy if F % 2 == 0 else dx) * dth / 4.0
                o = offs[(ck, kk)]
                for m_ in range(3):
                    i, j = {0: (2, m_), 1: (m_, 2), 2: (0, m_), 3: (m_, 0)}[F]
                    for a in range(2):
                        expect[o + 2 * 3 * i + 2 * j + a] += \
                            c * w[m_] * v[a] * math.cos(th[a] - F * math.pi / 2)
    assert np.max(np.abs(y - expect)) < 1e-15
    assert np.count_nonzero(y) == np.count_nonzero(expect) > 0


def test_intr_mask_geometric():
    """A DOF is a boundary (inflow) DOF iff its node lies on a domain wall
    and the direction theta_a points into the domain there."""
    from dg.matrix import get_intr_mask
    mesh = tiny_hp_mesh()
    m_loop = loops.intr_mask(mesh)
    m_vec = asm.intr_mask(mesh)
    m_prod = get_intr_mask(mesh)
    offs, n = meshq.dof_offsets(mesh)
    geo = np.ones(n, dtype=bool)
    for ck, kks in meshq.leaf_order(mesh):
        col = mesh.cols[ck]
        nx, ny = col.ndofs
        xf = quad.push(col.pos[0], col.pos[2], quad.lgl(nx)[0])
        yf = quad.push(col.pos[1], col.pos[3], quad.lgl(ny)[0])
        for kk in kks:
            cell = col.cells[kk]
            nth = cell.ndofs[0]
            th = quad.push(cell.pos[0], cell.pos[1], quad.lg(nth)[0])
            for i in range(nx):
                for j in range(ny):
                    for a in range(nth):
                        s = (math.cos(th[a]), math.sin(th[a]))
                        inflow = ((xf[i] == 0.0 and s[0] > 0) or (xf[i] == 1.0 and s[0] < 0)
                                  or (yf[j] == 0.0 and s[1] > 0) or (yf[j] == 1.0 and s[1] < 0))
                        geo[offs[(ck, kk)] + nth * ny * i + nth * j + a] = not inflow
    assert np.array_equal(m_loop, geo)
    assert np.array_equal(m_vec, geo)
    assert np.array_equal(m_prod, geo)
    assert (~geo).sum() > 0


def test_forcing_is_exact_quadrature():
    """Sum of F for a polynomial f is its exact integral over the domain."""
    mesh = problems.uniform_mesh([1.0, 1.0], [3, 3, 3], 2, 1)
    f = lambda x, y, th: 1.0 + x * y + 0.0 * th
    F = loops.forcing(mesh, f)
    exact = 2 * math.pi * (1.0 + 0.25)
    assert abs(F.sum() - exact) < 1e-13
    Fv = asm.forcing(mesh, f)
    assert np.max(np.abs(F - Fv)) < 1e-15


# ------------------------------------------------------------ mesh queries

def test_neighbour_cells_match_geometry():
    """nhbr_cells (same key, else parent, else children) equals the set of
    leaf cells of the neighbour column overlapping the cell in angle; the
    product's query agrees."""
    mesh = tiny_hp_mesh()
    for ck, kks in meshq.leaf_order(mesh):
        col = mesh.cols[ck]
        for F in range(4):
            for nk in {k for k in col.nhbr_keys[F] if k is not None}:
                for kk in kks:
                    a0, b0 = col.cells[kk].pos
                    geo = sorted(k1 for k1, c1 in mesh.cols[nk].cells.items()
                                 if c1.is_lf and min(b0, c1.pos[1]) - max(a0, c1.pos[0]) > 1e-12)
                    assert sorted(meshq.nhbr_cells(mesh, ck, kk, nk)) == geo
                    prod = [k for k in mesh.nhbr_cells_in_nhbr_col(ck, kk, nk) if k is not None]
                    assert sorted(prod) == geo


def test_stored_column_neighbours_match_geometry():
    mesh = tiny_hp_mesh()
    for ck in mesh.cols:
        geo = meshq.geometric_face_neighbours(mesh, ck)
        stored = [sorted({k for k in mesh.cols[ck].nhbr_keys[F] if k is not None})
                  for F in range(4)]
        assert stored == geo, ck


# ------------------------------------------------------------ solves

def test_block_jacobi_and_solves_small():
    mesh = problems.uniform_mesh([1.0, 1.0], [3, 3, 3], 2, 1)
    wl = problems.manufactured()
    V = asm.assemble(mesh, wl.kappa, wl.sigma, wl.Phi)
    A = asm.system_matrix(V)
    mask = asm.intr_mask(mesh)
    f = asm.forcing(mesh, wl.f)
    b = asm.bcs(mesh, wl.bcs)
    u, A_II, rhs = system.exact_solve(A, f, b, mask)
    I = np.flatnonzero(mask)
    assert np.linalg.norm(A_II @ u[I] - rhs) < 1e-12 * np.linalg.norm(rhs)
    invs = system.block_jacobi_inverses(A, mesh, mask)
    for (o, n), inv in zip(system.cell_blocks(mesh), invs):
        blk = A[o:o + n, o:o + n].toarray()
        mm = mask[o:o + n]
        blk[~mm, :] = 0.0
        blk[:, ~mm] = 0.0
        blk[~mm, ~mm] = 1.0
        assert np.max(np.abs(inv @ blk - np.eye(n))) < 1e-10
    x, its, info = system.gmres_solve(A_II, rhs, tol=1e-11)
    assert info == 0
    assert np.linalg.norm(x - u[I]) < 1e-8 * np.linalg.norm(u[I])


def test_mlp_oracle_forward_and_gradients():
    """Forward on a hand-computable net; backward against central finite
    differences of sum(dout * out)."""
    W0 = np.array([[1.0, -2.0], [0.5, 0.25]])
    b0 = np.array([0.1, -3.0])
    W1 = np.array([[2.0, -1.0]])
    b1 = np.array([0.5])
    X = np.array([[1.0, 1.0]])
    out, _ = system.mlp_forward([(W0, b0), (W1, b1)], X)
    # hidden = relu([1-2+0.1, 0.5+0.25-3]) = [0, 0]; out = 0.5
    assert out.shape == (1, 1) and abs(out[0, 0] - 0.5) < 1e-15
    X = np.array([[3.0, 1.0]])
    out, _ = system.mlp_forward([(W0, b0), (W1, b1)], X)
    # hidden = relu([1.1, -1.25]) = [1.1, 0]; out = 2.2 + 0.5
    assert abs(out[0, 0] - 2.7) < 1e-15
    rng = np.random.default_rng(3)
    sizes = [6, 16, 16, 16, 3]
    params = [(rng.normal(size=(sizes[i + 1], sizes[i])) / math.sqrt(sizes[i]),
               rng.normal(size=sizes[i + 1]) * 0.1) for i in range(4)]
    X = rng.normal(size=(9, 6))
    dout = rng.normal(size=(9, 3))
    out, cache = system.mlp_forward(params, X)
    grads = system.mlp_backward(params, cache, dout)
    h = 1e-6
    for li in range(4):
        for which in (0, 1):
            arr = params[li][which]
            flat = arr.reshape(-1)
            for idx in rng.choice(flat.size, size=min(6, flat.size), replace=False):
                old = flat[idx]
                flat[idx] = old + h
                fp = float((system.mlp_forward(params, X)[0] * dout).sum())
                flat[idx] = old - h
                fm = float((system.mlp_forward(params, X)[0] * dout).sum())
                flat[idx] = old
                fd = (fp - fm) / (2 * h)
                an = grads[li][which].reshape(-1)[idx]
                assert abs(fd - an) < 1e-6 * max(1.0, abs(fd)), (li, which, fd, an)


# ------------------------------------------------------------ split blocks, coarse grid


def _hp_cols_mesh():
    """4 x 4 columns of 3 x 3 nodes, 2 angular cells of 3 angles; one column
    p-refined and one cell h-refined so DOF counts differ by position."""
    m = problems.uniform_mesh([1.0, 1.0], [3, 3, 3], 1, 2)
    keys = sorted(m.cols)
    m.ref_col(keys[5], "spt", "p")
    last = sorted(m.cols)[-1]
    m.ref_cell(last, sorted(m.cols[last].cells)[0], "h")
    return m


@pytest.mark.parametrize("split", [None, 9, 20, 40])
def test_bj_split_blocks_partition_cells(split):
    """Blocks cover every DOF once, stay inside one cell, hold every spatial
    node of a contiguous angle range, and respect the size bound."""
    mesh = _hp_cols_mesh()
    blocks = system.bj_block_dofs(mesh, split)
    cells = system.cell_blocks(mesh)
    N = sum(n for _, n in cells)
    seen = np.concatenate(blocks)
    assert np.array_equal(np.sort(seen), np.arange(N))
    starts = np.array([o for o, _ in cells])
    for idx in blocks:
        ci = np.searchsorted(starts, idx[0], side="right") - 1
        o, n = cells[ci]
        assert idx.min() >= o and idx.max() < o + n
    # structure per cell from the geometry of the mesh (not the oracle's loop)
    offs, _ = meshq.dof_offsets(mesh)
    for ck, kks in meshq.leaf_order(mesh):
        col = mesh.cols[ck]
        nsp = col.ndofs[0] * col.ndofs[1]
        for kk in kks:
            nth = col.cells[kk].ndofs[0]
            o = offs[(ck, kk)]
            mine = [b for b in blocks if o <= b[0] < o + nsp * nth]
            if split is None or nsp * nth <= split:
                assert len(mine) == 1 and mine[0].size == nsp * nth
                continue
            angs = []
            for b in mine:
                loc = b - o
                s, a = loc // nth, loc % nth
                assert np.array_equal(np.unique(s), np.arange(nsp))        # all nodes
                ua = np.unique(a)
                assert np.array_equal(ua, np.arange(ua[0], ua[-1] + 1))    # contiguous
                assert b.size <= max(split, nsp)
                angs.append(ua[0])
            assert angs == sorted(angs)


def test_bj_split_inverses_solve_their_blocks():
    mesh = _hp_cols_mesh()
    wl = problems.manufactured()
    V = asm.assemble(mesh, wl.kappa, wl.sigma, wl.Phi)
    A = asm.system_matrix(V).tocsr()
    mask = asm.intr_mask(mesh)
    r = np.random.default_rng(0).normal(size=A.shape[0]) * mask
    for split in (None, 20):
        invs = system.block_jacobi_inverses(A, mesh, mask, split)
        z = system.block_jacobi_apply(invs, mesh, r, split)
        for idx in system.bj_block_dofs(mesh, split):
            blk = A[idx][:, idx].toarray()
            m = mask[idx]
            # interior rows of the block: blk_II z_I = r_I (backward error,
            # some blocks of this coarse mesh are badly conditioned), boundary z = 0
            B = blk[np.ix_(m, m)]
            lhs = B @ z[idx][m]
            scale = np.abs(B).sum(1) * np.abs(z[idx][m]).max() + np.abs(r[idx][m])
            assert np.all(np.abs(lhs - r[idx][m]) <= 1e-12 * scale)
            assert np.all(z[idx][~m] == 0.0)


def _col_geom(mesh):
    return [mesh.cols[ck].pos for ck, _ in meshq.leaf_order(mesh)]


@pytest.mark.parametrize("n", [4, 2, 3])
def test_coarse_map_partitions_and_positions(n):
    """fcol sums to 1 per column, fcell to 1 per coarse cell; when the grid
    matches or coarsens the columns every column lies in the coarse cell
    holding its centre (g = gx n + gy)."""
    mesh = _hp_cols_mesh()
    cmap = system.coarse_map(mesh, n)
    geo = _col_geom(mesh)
    tot_cell = np.zeros(n * n)
    for ent, (x0, y0, x1, y1) in zip(cmap, geo):
        assert abs(sum(e[2] for e in ent) - 1.0) < 1e-12
        for g, fcell, fcol in ent:
            tot_cell[g] += fcell
            # fcell / fcol = area(col) / area(coarse cell)
            assert abs(fcell / fcol - (x1 - x0) * (y1 - y0) * n * n) < 1e-12
        if n in (2, 4):
            cx, cy = 0.5 * (x0 + x1), 0.5 * (y0 + y1)
            assert len(ent) == 1 and ent[0][0] == int(cx * n) * n + int(cy * n)
    assert np.allclose(tot_cell, 1.0, atol=1e-12)


def test_coarse_moments_closed_forms():
    """On a grid that coarsens the columns: channel 0 of r = 1 counts the
    DOFs of the columns inside each coarse cell; the three channels of a
    general r sum to the global sums of r, r cos th, r sin th."""
    mesh = _hp_cols_mesh()
    n = 2
    offs, N = meshq.dof_offsets(mesh)
    geo = _col_geom(mesh)
    mom = system.coarse_moments(mesh, n, np.ones(N))
    want = np.zeros(n * n)
    for (ck, kks), (x0, y0, x1, y1) in zip(meshq.leaf_order(mesh), geo):
        col = mesh.cols[ck]
        nd = col.ndofs[0] * col.ndofs[1] * sum(col.cells[k].ndofs[0] for k in kks)
        want[int(0.5 * (x0 + x1) * n) * n + int(0.5 * (y0 + y1) * n)] += nd
    assert np.allclose(mom[:n * n], want, rtol=0, atol=1e-9)
    # global sums against the quadrature nodes of every cell
    r = np.random.default_rng(1).normal(size=N)
    th = np.zeros(N)
    for ck, kks in meshq.leaf_order(mesh):
        col = mesh.cols[ck]
        nx, ny = col.ndofs
        for kk in kks:
            c = col.cells[kk]
            nth = c.ndofs[0]
            t = c.pos[0] + (c.pos[1] - c.pos[0]) * 0.5 * (np.polynomial.legendre.leggauss(nth)[0] + 1)
            th[offs[(ck, kk)]:offs[(ck, kk)] + nx * ny * nth] = np.tile(t, nx * ny)
    for m3 in (2, 3):
        mom = system.coarse_moments(mesh, m3, r)
        k = m3 * m3
        assert abs(mom[:k].sum() - r.sum()) < 1e-10 * np.abs(r).sum()
        assert abs(mom[k:2 * k].sum() - (r * np.cos(th)).sum()) < 1e-10 * np.abs(r).sum()
        assert abs(mom[2 * k:].sum() - (r * np.sin(th)).sum()) < 1e-10 * np.abs(r).sum()


def test_coarse_average_and_prolong_closed_forms():
    """Average of a constant is that constant; of x (resp. y) it is the
    coarse cell's centre when the cell is a union of columns.  Prolongation
    of a coarse indicator is `scale` on the columns inside it, 0 elsewhere
    and on masked DOFs."""
    mesh = _hp_cols_mesh()
    n = 2
    offs, N = meshq.dof_offsets(mesh)
    xs, ys = np.zeros(N), np.zeros(N)
    colid = np.zeros(N, int)
    geo = _col_geom(mesh)
    for ci, ((ck, kks), (x0, y0, x1, y1)) in enumerate(zip(meshq.leaf_order(mesh), geo)):
        col = mesh.cols[ck]
        nx, ny = col.ndofs
        xn = x0 + (x1 - x0) * 0.5 * (quad.lgl(nx)[0] + 1)
        yn = y0 + (y1 - y0) * 0.5 * (quad.lgl(ny)[0] + 1)
        for kk in kks:
            nth = col.cells[kk].ndofs[0]
            o = offs[(ck, kk)]
            sl = slice(o, o + nx * ny * nth)
            xs[sl] = np.repeat(xn, ny * nth)
            ys[sl] = np.tile(np.repeat(yn, nth), nx)
            colid[sl] = ci
    assert np.allclose(system.coarse_average(mesh, n, np.full(N, 2.5)), 2.5, atol=1e-12)
    cen = (np.arange(n) + 0.5) / n
    ax = system.coarse_average(mesh, n, xs).reshape(n, n)
    ay = system.coarse_average(mesh, n, ys).reshape(n, n)
    assert np.allclose(ax, np.repeat(cen[:, None], n, 1), atol=1e-12)
    assert np.allclose(ay, np.repeat(cen[None, :], n, 0), atol=1e-12)
    mask = asm.intr_mask(mesh)
    for g in range(n * n):
        phi = np.zeros(n * n)
        phi[g] = 1.0
        z = system.coarse_prolong(mesh, n, phi, 0.7, mask)
        inside = np.array([int(0.5 * (x0 + x1) * n) * n + int(0.5 * (y0 + y1) * n) == g
                           for (x0, y0, x1, y1) in geo])[colid]
        assert np.allclose(z, np.where(inside & mask, 0.7, 0.0), atol=1e-14)


@pytest.mark.parametrize("pbcs", [(True, False), (False, True), (True, True)])
def test_periodic_boundary_convection_loops_vs_vectorised(pbcs):
    """Periodic meshes: the literal loops and the vectorised assembly agree
    on BC, a periodic direction has no inflow boundary, and on a conforming
    periodic mesh every cell's 1^T (BC - IC) 1 vanishes (no wall terms
    remain; opposite faces cancel)."""
    mesh = problems.uniform_mesh([1.0, 1.0], [3, 3, 3], 1, 2, pbcs=pbcs)
    wl = problems.manufactured()
    V = asm.assemble(mesh, wl.kappa, wl.sigma, wl.Phi)
    BCl = loops.boundary_conv(mesh)
    assert abs(V["BC"] - BCl).max() < 1e-13
    if all(pbcs):
        assert asm.intr_mask(mesh).all()
        T = (V["BC"] - V["IC"]).tocsr()
        for o, n in system.cell_blocks(mesh):
            tot = T[o:o + n].sum()
            assert abs(tot) < 1e-12, tot
