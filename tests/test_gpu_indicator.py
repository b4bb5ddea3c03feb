"""Device error indicators, projection integrals and hp steering against the
oracle (oracle/indicator.py), and the Error_Indicator methods built on them.

References: src/amr/Error_Indicator/error/error_angular_jump.py,
error_spatial_jump.py, error_high_resolution.py, cell_hp_steer.py,
col_hp_steer.py, src/dg/projection/cell_intg_xy.py, col_intg_th.py.
Tolerance: 1e-12 relative per cell / column (FP64 sums in another order).
"""
import numpy as np
import pytest

from oracle import indicator as oind
from oracle.meshq import leaf_order
from rtx import problems
from test_oracle_indicator import hp_mesh, nodal

TOL = 1e-12


def _meshes(case):
    if case == "hp":
        return hp_mesh()
    if case == "periodic_x":
        return problems.uniform_mesh([1.0, 1.0], [3, 3, 3], 2, 2, pbcs=(True, False))
    from test_gpu_apply import small_random_hp
    return small_random_hp(1)


def _close(a, b, tol=TOL):
    return abs(a - b) <= tol * max(abs(b), 1e-300)


@pytest.mark.gpu
@pytest.mark.parametrize("case", ["hp", "random_hp", "periodic_x"])
def test_indicators_match_oracle(cuda, case):
    from rtx.error import ErrorEvaluator
    mesh = _meshes(case)
    print("indicators", case, "ndof", mesh.get_ndof(), flush=True)
    ev = ErrorEvaluator(mesh)
    t = ev.t
    u = np.random.default_rng(5).uniform(-1, 1, t.ndof)
    # angular jumps per cell
    ang = ev.angular_jump(u)
    ref = oind.angular_jump(mesh, u)
    for k, key in enumerate(t.cell_keys):
        assert _close(ang[k], ref[key]), (key, ang[k], ref[key])
    # spatial jumps per column
    spt = ev.spatial_jump(u)
    ref = oind.spatial_jump(mesh, u)
    for c, ck in enumerate(t.col_keys):
        assert _close(spt[c], ref[ck]), (ck, spt[c], ref[ck])
    # projection integrals
    xy, th, n2 = ev.moments(u)
    for c, ck in enumerate(t.col_keys):
        nx, ny = int(t.col_nx[c]), int(t.col_ny[c])
        s0 = int(t.col_sp_off[c])
        assert np.allclose(th[s0:s0 + nx * ny].reshape(nx, ny), oind.col_intg_th(mesh, u, ck),
                           rtol=TOL, atol=TOL * np.abs(th).max())
    for k, (ck, kk) in enumerate(t.cell_keys):
        c = int(t.cell_col[k])
        a0 = int(t.col_ang_start[c]) + int(t.cell_ang_off[k])
        ref = oind.cell_intg_xy(mesh, u, ck, kk)
        assert np.allclose(xy[a0:a0 + ref.size], ref, rtol=TOL, atol=TOL * np.abs(ref).max())
    # the nodal norm is the high-resolution normalisation of a mesh against itself
    _, intg = oind.high_resolution(mesh, u, mesh, u) if case == "hp" else (None, None)
    if intg is not None:
        assert _close(float(np.sum(n2)), intg)


@pytest.mark.gpu
def test_high_resolution_errors_match_oracle(cuda):
    from rtx.error import ErrorEvaluator
    from rtx.plan import MeshTables
    mesh = hp_mesh()
    mesh_hr = hp_mesh()
    mesh_hr.ref_mesh(kind="ang", form="p")
    mesh_hr.ref_mesh(kind="spt", form="p")
    ev = ErrorEvaluator(mesh)
    t_hr = MeshTables(mesh_hr)
    rng = np.random.default_rng(6)
    u = rng.uniform(-1, 1, ev.t.ndof)
    u_hr = rng.uniform(-1, 1, t_hr.ndof)
    err, n2 = ev.nodal_errors(u, u_hr, t_hr)
    ref, intg = oind.high_resolution(mesh, u, mesh_hr, u_hr)
    for k, key in enumerate(ev.t.cell_keys):
        assert _close(err[k], ref[key]), (key, err[k], ref[key])
    assert _close(float(np.sum(n2)), intg)


def _oracle_flags(ind_kind, mesh, u, tol):
    """Errors, flags and forms the reference's rules give for the oracle's
    indicator values with ref_kind "all", ref_form "hp"."""
    if ind_kind == "angular":
        e = oind.angular_jump(mesh, u)
        cells = {key: np.sqrt(v) for key, v in e.items()}
        cols = {}
        for (ck, kk), v in e.items():
            cols[ck] = cols.get(ck, 0.0) + v
        cols = {ck: np.sqrt(v) for ck, v in cols.items()}
    else:
        cols = {ck: np.sqrt(v) for ck, v in oind.spatial_jump(mesh, u).items()}
        cells = {(ck, kk): -np.finfo(np.float64).max for ck, kks in leaf_order(mesh) for kk in kks}
    cell_max, col_max = max(cells.values()), max(cols.values())
    col_flag = {ck: v >= tol * col_max for ck, v in cols.items()}
    cell_flag = {key: v >= tol * cell_max for key, v in cells.items()}
    return cols, cells, col_flag, cell_flag


@pytest.mark.gpu
@pytest.mark.parametrize("kind", ["angular", "spatial"])
def test_error_indicator_jump_methods(cuda, kind):
    """Error_Indicator.error_angular_jump / error_spatial_jump: errors and
    flags equal the reference's rules on the oracle's values; steered forms
    equal the oracle's steering on the oracle's integrals."""
    from dg.projection import Projection
    from amr.error_indicator import Error_Indicator
    from rtx.plan import MeshTables
    mesh = hp_mesh()
    t = MeshTables(mesh)
    u = np.random.default_rng(7).uniform(-1, 1, t.ndof)
    proj = Projection(mesh)
    proj.from_vector(u)
    tol = 0.5
    ind = Error_Indicator(proj, ref_kind="all", ref_form="hp", ref_tol=[tol, tol])
    (ind.error_angular_jump if kind == "angular" else ind.error_spatial_jump)()
    cols, cells, col_flag, cell_flag = _oracle_flags(kind, mesh, u, tol)
    for ck, col in ind.cols.items():
        assert _close(col.error, cols[ck])
        assert col.do_ref == col_flag[ck]
        if col.do_ref:
            nx, ny = mesh.cols[ck].ndofs
            _, _, lp = oind.col_steer_lp(nx, ny, oind.col_intg_th(mesh, u, ck))
            assert col.ref_form == ("p" if lp - 0.5 >= 0.5 * (nx + ny) else "h")
        for kk, cell in col.cells.items():
            if kind == "angular":
                assert _close(cell.error, cells[(ck, kk)])
            else:
                assert cell.error == cells[(ck, kk)]
            assert cell.do_ref == cell_flag[(ck, kk)]
            if cell.do_ref:
                nth = mesh.cols[ck].cells[kk].ndofs[0]
                _, lp = oind.cell_steer_lp(nth, oind.cell_intg_xy(mesh, u, ck, kk))
                assert cell.ref_form == ("p" if lp - 0.5 >= nth else "h")
    assert any(c.do_ref for c in ind.cols.values())
    # the refined mesh has more DOFs, and its refined entities are the flagged ones
    new = ind.ref_by_ind()
    assert new.get_ndof() > mesh.get_ndof()


@pytest.mark.gpu
def test_error_high_resolution_method(cuda):
    """error_high_resolution on a manufactured problem: the solve runs on the
    p-refined mesh and the errors equal the oracle's against that solution."""
    from rt import Problem
    from dg.projection import Projection
    from amr.error_indicator import Error_Indicator
    wl = problems.manufactured()
    mesh = problems.uniform_mesh([1.0, 1.0], [3, 3, 3], 2, 2)
    prob = Problem(wl.kappa, wl.sigma, wl.Phi, wl.f, wl.bcs_dirac)
    uh, conv, _ = prob.solve(mesh, rtol=1e-10, atol=0.0)
    ind = Error_Indicator(uh, ref_kind="all", ref_form="h", ref_tol=[0.9, 0.9])
    uh_hr, info, _ = ind.error_high_resolution(prob, ang_ref_offset=1, spt_ref_offset=1,
                                               rtol=1e-10, atol=0.0)
    assert info["converged_reason"] > 0
    assert uh_hr.mesh.cols[sorted(uh_hr.mesh.cols)[0]].ndofs == [4, 4]
    ref, intg = oind.high_resolution(mesh, uh.to_vector(), uh_hr.mesh, uh_hr.to_vector())
    tot = sum(ref.values())
    assert _close(ind.error, np.sqrt(tot / intg))
    for (ck, kk), v in ref.items():
        assert _close(ind.cols[ck].cells[kk].error, np.sqrt(v / intg))
    # a smooth solution: a small relative error against one order more, and
    # some cells flagged at 0.9 of the maximum
    assert 0 < ind.error < 0.5
    assert any(c.do_ref for col in ind.cols.values() for c in col.cells.values())
