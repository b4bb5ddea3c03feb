"""Pins of the oracle's error indicators, projection integrals and hp
steering (oracle/indicator.py) by what the mathematics fixes.

* Jumps vanish for fields the discretisation represents continuously
  (theta-independent fields for angular jumps; low-degree polynomials in
  x, y, theta for spatial jumps on an hp mesh with level and order jumps).
* Piecewise constants give closed-form jumps.
* A field linear in theta jumps only across the 2 pi wrap.
* Integrals of 1, x and theta are closed forms.
* Legendre-mode fields give closed-form steering coefficients.
* The high-resolution error of an exactly interpolated field is zero, and a
  planted one-node difference gives w_p w_q w_r delta^2 (dx dy dth / 8).
"""
import numpy as np
import pytest
from scipy.special import eval_legendre

from oracle import indicator as oind
from oracle.quad import lgl, lg, push
from oracle.meshq import leaf_order, dof_offsets
from rtx import problems


def nodal(mesh, f):
    """f at every DOF node in the reference's order."""
    out = []
    for ck, kks in leaf_order(mesh):
        col = mesh.cols[ck]
        x0, y0, x1, y1 = col.pos
        nx, ny = col.ndofs
        xs = push(x0, x1, lgl(nx)[0])
        ys = push(y0, y1, lgl(ny)[0])
        for kk in kks:
            th0, th1 = col.cells[kk].pos
            ts = push(th0, th1, lg(col.cells[kk].ndofs[0])[0])
            out.append(np.broadcast_to(f(xs[:, None, None], ys[None, :, None],
                                         ts[None, None, :]),
                                       (nx, ny, len(ts))).ravel())
    return np.concatenate(out)


def per_cell(mesh, vals):
    """A vector constant on every cell, from {(col_key, cell_key): value}."""
    out = []
    for ck, kks in leaf_order(mesh):
        nx, ny = mesh.cols[ck].ndofs
        for kk in kks:
            out.append(np.full(nx * ny * mesh.cols[ck].cells[kk].ndofs[0], vals[(ck, kk)]))
    return np.concatenate(out)


def hp_mesh():
    m = problems.uniform_mesh([1.0, 1.0], [3, 3, 3], 2, 1)
    keys = sorted(m.cols)
    m.ref_col(keys[0], "spt", "hp")
    last = sorted(m.cols)[-1]
    m.ref_cell(last, sorted(m.cols[last].cells)[0], "hp")
    first = sorted(m.cols)[0]
    m.ref_cell(first, sorted(m.cols[first].cells)[2], "h")
    return m


def cells_of(mesh):
    return [(ck, kk) for ck, kks in leaf_order(mesh) for kk in kks]


# ----------------------------------------------------------- angular jump
def test_angular_jump_theta_independent_field_is_zero():
    mesh = hp_mesh()
    u = nodal(mesh, lambda x, y, t: 1.0 + x + y * y + 0.0 * t)
    out = oind.angular_jump(mesh, u)
    assert max(out.values()) < 1e-24


def test_angular_jump_piecewise_constant_closed_form():
    mesh = hp_mesh()
    rng = np.random.default_rng(0)
    c = {key: rng.uniform(-1, 1) for key in cells_of(mesh)}
    out = oind.angular_jump(mesh, per_cell(mesh, c))
    for (ck, kk), v in out.items():
        lo, up = mesh.cols[ck].cells[kk].nhbr_keys[:2]
        ref = 0.5 * ((c[(ck, kk)] - c[(ck, lo)]) ** 2 + (c[(ck, kk)] - c[(ck, up)]) ** 2)
        assert abs(v - ref) <= 1e-13 * max(ref, 1.0), ((ck, kk), v, ref)


def test_angular_jump_linear_theta_jumps_only_at_the_wrap():
    """u = theta is continuous between neighbouring cells except across
    theta = 0 = 2 pi, where the jump is 2 pi; each of the two wrap cells then
    carries (1/4) 4 (1/2) (2 pi)^2 and every other cell 0.  Swapped top and
    bottom traces would make every cell jump."""
    mesh = hp_mesh()
    out = oind.angular_jump(mesh, nodal(mesh, lambda x, y, t: t + 0.0 * x))
    for (ck, kk), v in out.items():
        th0, th1 = mesh.cols[ck].cells[kk].pos
        wrap = (th0 == 0.0) or abs(th1 - 2 * np.pi) < 1e-14
        ref = 0.5 * (2 * np.pi) ** 2 if wrap else 0.0
        assert abs(v - ref) <= 1e-11, ((ck, kk), v, ref)


# ----------------------------------------------------------- spatial jump
def test_spatial_jump_low_degree_polynomial_is_zero():
    """A field of degree 1 in x, y and theta is continuous across every face
    and exactly interpolated on both sides (orders 2..4, level jumps in space
    and angle), so every projected trace pair agrees; the wrong face on either
    side or a wrong angular projection would not."""
    mesh = hp_mesh()
    u = nodal(mesh, lambda x, y, t: (1.0 + x + 2.0 * y) * (1.0 + t))
    out = oind.spatial_jump(mesh, u)
    assert max(out.values()) < 1e-22


def test_spatial_jump_piecewise_constant_closed_form():
    """u = c_C on column C: every face pair contributes d_A dth_G (c_0 - c_1)^2,
    the angular pairs of a face cover 2 pi once, so column error =
    sum over faces and neighbours of d_A (c_0 - c_1)^2 / perimeter with d_A
    the shorter of the two face lengths."""
    mesh = hp_mesh()
    rng = np.random.default_rng(1)
    cc = {ck: rng.uniform(-1, 1) for ck, _ in leaf_order(mesh)}
    c = {(ck, kk): cc[ck] for ck, kk in cells_of(mesh)}
    out = oind.spatial_jump(mesh, per_cell(mesh, c))
    for ck0, _ in leaf_order(mesh):
        col0 = mesh.cols[ck0]
        x0, y0, x1, y1 = col0.pos
        ref = 0.0
        for F in range(4):
            for ck1 in set(col0.nhbr_keys[F]):
                if ck1 is None:
                    continue
                a0, b0, a1, b1 = mesh.cols[ck1].pos
                L0 = (y1 - y0) if F % 2 == 0 else (x1 - x0)
                L1 = (b1 - b0) if F % 2 == 0 else (a1 - a0)
                ref += min(L0, L1) * (cc[ck0] - cc[ck1]) ** 2
        ref /= 2 * (x1 - x0) + 2 * (y1 - y0)
        assert abs(out[ck0] - ref) <= 1e-13 * max(ref, 1.0), (ck0, out[ck0], ref)


# --------------------------------------------------------------- integrals
def test_projection_integrals_closed_forms():
    mesh = hp_mesh()
    one = nodal(mesh, lambda x, y, t: 1.0 + 0.0 * (x + y + t))
    u_x = nodal(mesh, lambda x, y, t: x + 0.0 * (y + t))
    u_t = nodal(mesh, lambda x, y, t: t + 0.0 * (x + y))
    for ck, kks in leaf_order(mesh):
        x0, y0, x1, y1 = mesh.cols[ck].pos
        assert np.allclose(oind.col_intg_th(mesh, one, ck), 2 * np.pi, rtol=1e-14)
        assert np.allclose(oind.col_intg_th(mesh, u_t, ck), 2 * np.pi ** 2, rtol=1e-14)
        for kk in kks:
            assert np.allclose(oind.cell_intg_xy(mesh, one, ck, kk), (x1 - x0) * (y1 - y0),
                               rtol=1e-14)
            assert np.allclose(oind.cell_intg_xy(mesh, u_x, ck, kk),
                               (y1 - y0) * (x1 ** 2 - x0 ** 2) / 2.0, rtol=1e-13)


# --------------------------------------------------------------- steering
@pytest.mark.parametrize("nth", [3, 4, 6, 9])
def test_cell_steer_legendre_mode(nth):
    """uh_hat = c L_{n-1}: Gauss-Legendre is exact on L_{n-1}^2, so a_n = c and
    lp = log((2n - 1) / (2 |c|)) / log(n - 1)."""
    tb, _ = lg(nth)
    for c in (1e-6, 0.37, 3.0):
        a, lp = oind.cell_steer_lp(nth, c * eval_legendre(nth - 1, tb))
        assert abs(a - c) <= 1e-13 * c
        assert abs(lp - np.log((2 * nth - 1) / (2 * c)) / np.log(nth - 1)) <= 1e-12
    # a last coefficient at rounding level: the decision is "p"
    a, lp = oind.cell_steer_lp(nth, 2.0 + tb)
    assert abs(a) < 1e-14 and lp - 0.5 >= nth


@pytest.mark.parametrize("nx,ny", [(3, 3), (4, 6), (7, 5)])
def test_col_steer_legendre_modes(nx, ny):
    """uh_hat = c1 L_{nx-1}(x) + c2 L_{ny-1}(y): the LGL rule gives
    sum w L_N^2 = 2 / N and is exact on the lower modes, so
    ax^2 = 2 (c1 (2nx - 1) / (nx - 1))^2 and likewise ay^2."""
    xb, _ = lgl(nx)
    yb, _ = lgl(ny)
    c1, c2 = 0.3, 0.02
    uh = c1 * eval_legendre(nx - 1, xb)[:, None] + c2 * eval_legendre(ny - 1, yb)[None, :]
    ax, ay, lp = oind.col_steer_lp(nx, ny, uh)
    ax_ref = 2 * (c1 * (2 * nx - 1) / (nx - 1)) ** 2
    ay_ref = 2 * (c2 * (2 * ny - 1) / (ny - 1)) ** 2
    assert abs(ax - ax_ref) <= 1e-12 * ax_ref and abs(ay - ay_ref) <= 1e-12 * ay_ref
    t0 = np.log((2 * nx - 1) / (2 * ax_ref)) / (2 * np.log(nx - 1))
    t1 = np.log((2 * ny - 1) / (2 * ay_ref)) / (2 * np.log(ny - 1))
    assert abs(lp - 0.5 * (t0 + t1)) <= 1e-12


# ------------------------------------------------------- high resolution
def test_high_resolution_exact_interpolation_and_planted_node():
    mesh = problems.uniform_mesh([1.0, 1.0], [2, 2, 2], 2, 1)
    mesh_hr = problems.uniform_mesh([1.0, 1.0], [2, 2, 2], 2, 1)
    mesh_hr.ref_mesh(kind="ang", form="p")
    mesh_hr.ref_mesh(kind="spt", form="p")
    f = lambda x, y, t: (1.0 + x) * (1.0 + 2.0 * y) * (1.0 + t)
    u, u_hr = nodal(mesh, f), nodal(mesh_hr, f)
    out, intg = oind.high_resolution(mesh, u, mesh_hr, u_hr)
    assert max(out.values()) < 1e-24
    one_hr = nodal(mesh_hr, lambda x, y, t: 1.0 + 0.0 * (x + y + t))
    _, intg1 = oind.high_resolution(mesh, nodal(mesh, lambda x, y, t: 1.0 + 0.0 * x),
                                    mesh_hr, one_hr)
    assert abs(intg1 - 2 * np.pi) <= 1e-13
    # plant delta at node (p, q, r) of one cell
    keys = cells_of(mesh_hr)
    ck, kk = keys[3]
    offs, _ = dof_offsets(mesh_hr)
    mx, my = mesh_hr.cols[ck].ndofs
    mt = mesh_hr.cols[ck].cells[kk].ndofs[0]
    p, q, r, delta = 1, 2, 0, 0.125
    v = u_hr.copy()
    v[offs[(ck, kk)] + (p * my + q) * mt + r] += delta
    out, _ = oind.high_resolution(mesh, u, mesh_hr, v)
    x0, y0, x1, y1 = mesh.cols[ck].pos
    th0, th1 = mesh.cols[ck].cells[kk].pos
    ref = (x1 - x0) * (y1 - y0) * (th1 - th0) / 8.0 * lgl(mx)[1][p] * lgl(my)[1][q] \
        * lg(mt)[1][r] * delta ** 2
    assert abs(out[(ck, kk)] - ref) <= 1e-14
    assert sum(v for key, v in out.items() if key != (ck, kk)) < 1e-24
