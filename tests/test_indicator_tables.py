"""Host tables of the device indicator kernels, checked on the CPU.

* Every spatial-jump pair record stays inside the buffers
  ``rtx_error_spatial_jump`` reads (cells, pool, node counts within the
  kernel's LDS tiles), and a numpy emulation of the kernel's arithmetic over
  the records reproduces the oracle's literal loops
  (``error_spatial_jump.py:30-227``).
* The angular trace table holds l_a(-1), l_a(+1) of every Gauss-Legendre
  basis: the traces of a polynomial of degree < n are its end values.
* The nodal-error interpolation blocks reproduce polynomials of degree < n
  at the m-point nodes.
"""
import numpy as np
import pytest

from oracle import indicator as oind
from oracle.quad import lgl, lg
from rtx import problems
from rtx._lib import NMAX
from rtx.plan import MeshTables
from rtx.error import (build_jump_pairs, trace_tables, nodal_error_tables, TAB_STRIDE, MM)
from test_oracle_indicator import hp_mesh


def _meshes():
    from test_gpu_apply import small_random_hp
    per = problems.uniform_mesh([1.0, 1.0], [3, 3, 3], 2, 2, pbcs=(True, False))
    return [("hp", hp_mesh()), ("random_hp", small_random_hp(1)), ("periodic_x", per)]


def emulate_spatial_jump(t, tabs, u):
    """The arithmetic of k_spatial_jump over the pair records, in numpy."""
    start, pair, scale, pool = tabs
    P = pair.reshape(-1, 9)
    out = np.zeros(t.ncol)
    for c in range(t.ncol):
        acc = 0.0
        for p in range(start[c], start[c + 1]):
            k0, k1, F, nB, nH, o0, o1, o2, o3 = (int(v) for v in P[p])
            c0, c1 = t.cell_col[k0], t.cell_col[k1]
            nx0, ny0, nt0 = t.col_nx[c0], t.col_ny[c0], t.cell_nth[k0]
            nx1, ny1, nt1 = t.col_nx[c1], t.col_ny[c1], t.cell_nth[k1]
            v0 = u[t.cell_dof_off[k0]:t.cell_dof_off[k0] + nx0 * ny0 * nt0].reshape(nx0, ny0, nt0)
            v1 = u[t.cell_dof_off[k1]:t.cell_dof_off[k1] + nx1 * ny1 * nt1].reshape(nx1, ny1, nt1)
            h0 = {0: v0[-1], 1: v0[:, -1], 2: v0[0], 3: v0[:, 0]}[F]
            h1 = {0: v1[0], 1: v1[:, 0], 2: v1[-1], 3: v1[:, -1]}[F]
            n0, n1 = h0.shape[0], h1.shape[0]
            psi0 = pool[o0:o0 + n0 * nB].reshape(n0, nB)
            psi1 = pool[o1:o1 + n1 * nB].reshape(n1, nB)
            xsi0 = pool[o2:o2 + nt0 * nH].reshape(nt0, nH)
            xsi1 = pool[o3:o3 + nt1 * nH].reshape(nt1, nH)
            U0 = psi0.T @ h0 @ xsi0
            U1 = psi1.T @ h1 @ xsi1
            w = lgl(nB)[1][:, None] * lg(nH)[1][None, :]
            acc += scale[p] * np.sum(w * (U0 - U1) ** 2)
        x0, y0, x1, y1 = t.col_geom[c]
        out[c] = acc / ((2 * (x1 - x0) + 2 * (y1 - y0)) * 2 * np.pi)
    return out


@pytest.mark.parametrize("case", ["hp", "random_hp", "periodic_x"])
def test_jump_pairs_in_bounds_and_match_oracle(case):
    mesh = dict(_meshes())[case]
    t = MeshTables(mesh)
    tabs = build_jump_pairs(mesh, t)
    start, pair, scale, pool = tabs
    P = pair.reshape(-1, 9)
    assert start[0] == 0 and start[-1] == P.shape[0] and (np.diff(start) >= 0).all()
    assert scale.shape[0] == P.shape[0] and (scale > 0).all()
    for rec in P:
        k0, k1, F, nB, nH, o0, o1, o2, o3 = (int(v) for v in rec)
        assert 0 <= k0 < t.ncell and 0 <= k1 < t.ncell and 0 <= F < 4
        c0, c1 = t.cell_col[k0], t.cell_col[k1]
        n0 = t.col_ny[c0] if F % 2 == 0 else t.col_nx[c0]
        n1 = t.col_ny[c1] if F % 2 == 0 else t.col_nx[c1]
        nt0, nt1 = t.cell_nth[k0], t.cell_nth[k1]
        assert nB == max(n0, n1) and nH == max(nt0, nt1) and 1 <= nB <= NMAX and nH <= NMAX
        # LDS tiles of the kernel: NMAX x NMAX doubles each
        assert n0 * nt0 <= NMAX * NMAX and n1 * nt1 <= NMAX * NMAX
        assert nB * nt0 <= NMAX * NMAX and nB * nt1 <= NMAX * NMAX
        for off, size in ((o0, n0 * nB), (o1, n1 * nB), (o2, nt0 * nH), (o3, nt1 * nH)):
            assert 0 <= off and off + size <= pool.size
    # every pair belongs to the column the kernel's workgroup owns
    for c in range(t.ncol):
        assert (t.cell_col[P[start[c]:start[c + 1], 0]] == c).all()
    u = np.random.default_rng(3).uniform(-1, 1, t.ndof)
    ref = oind.spatial_jump(mesh, u)
    got = emulate_spatial_jump(t, tabs, u)
    for c, ck in enumerate(t.col_keys):
        assert abs(got[c] - ref[ck]) <= 1e-12 * max(abs(ref[ck]), 1e-300), (ck, got[c], ref[ck])


def test_trace_table_gives_end_values():
    lend = trace_tables()
    for n in range(1, NMAX + 1):
        tb = lg(n)[0]
        for deg in range(n):
            v = tb ** deg
            bot = np.dot(v, lend[(n * 2 + 0) * NMAX:(n * 2 + 0) * NMAX + n])
            top = np.dot(v, lend[(n * 2 + 1) * NMAX:(n * 2 + 1) * NMAX + n])
            assert abs(bot - (-1.0) ** deg) <= 1e-10 and abs(top - 1.0) <= 1e-10, (n, deg)


def test_nodal_error_tables_interpolate_polynomials():
    tab = nodal_error_tables()
    for n in range(1, NMAX + 1):
        xl, xg = lgl(n)[0], lg(n)[0]
        for m in range(1, NMAX + 1):
            b = (n * (MM + 1) + m) * TAB_STRIDE
            Pl = tab[b:b + n * m].reshape(n, m)
            Pg = tab[b + NMAX * MM:b + NMAX * MM + n * m].reshape(n, m)
            yl, wl = lgl(m)
            yg, wg = lg(m)
            for deg in range(n):
                assert np.allclose(xl ** deg @ Pl, yl ** deg, atol=1e-10)
                assert np.allclose(xg ** deg @ Pg, yg ** deg, atol=1e-10)
            assert np.allclose(tab[b + 2 * NMAX * MM:b + 2 * NMAX * MM + m], wl)
            assert np.allclose(tab[b + 2 * NMAX * MM + MM:b + 2 * NMAX * MM + MM + m], wg)
    assert tab.size == (NMAX + 1) * (MM + 1) * TAB_STRIDE
