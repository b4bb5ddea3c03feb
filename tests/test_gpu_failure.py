"""Failure handling and the reference's default call of ``Problem.solve``.

Reference: ``src/rt/solve.py:85-104`` (rtol = atol = EPS, max_it 5000, the
PETSc ``converged_reason`` codes), ``src/rt/class_Problem.py:11-32`` (a
``Problem`` rebuilds every matrix per call, so an in-place refinement of the
mesh is always seen).  Every case here is an ordinary call that returns or
raises an error; no kernel is made to fault or hang (a non-finite block
ends its workgroup through the same pivot test as any singular block).
"""

import numpy as np
import pytest

from oracle import assemble as asm, loops, system as osys
from rtx import problems


def _c1_small():
    return problems.uniform_mesh([1.0, 1.0], [4, 4, 4], 2, 2), problems.manufactured()


def _problem(wl, **over):
    from rt import Problem
    fns = {"kappa": wl.kappa, "sigma": wl.sigma, "Phi": wl.Phi, "f": wl.f}
    fns.update(over)
    return Problem(fns["kappa"], fns["sigma"], fns["Phi"], fns["f"], wl.bcs_dirac)


def _exact(mesh, wl):
    V = asm.assemble(mesh, wl.kappa, wl.sigma, wl.Phi)
    A = asm.system_matrix(V)
    u_ref, _, _ = osys.exact_solve(A, asm.forcing(mesh, wl.f), asm.bcs(mesh, wl.bcs),
                                   asm.intr_mask(mesh))
    return u_ref


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_reference_default_tolerances_converge(cuda, precision):
    """solve(mesh, ksp_type, pc_type) with the reference's rtol = atol =
    1e-15: the true residual cannot reach that on this mesh, so the solve
    stops at its rounding floor with a positive reason after a bounded
    number of iterations, at the exact discrete solution."""
    mesh, wl = _c1_small()
    prob = _problem(wl)
    kw = {"ksp_type": "gmres", "pc_type": "bjacobi"}
    if precision != "fp64":
        kw["precision"] = precision
    uh, conv, minfo = prob.solve(mesh, **kw)
    assert conv["converged_reason"] > 0, conv["converged_reason"]
    assert conv["iteration_number"] < (400 if precision == "fp64" else 600), conv["iteration_number"]
    u_ref = _exact(mesh, wl)
    assert np.linalg.norm(uh.to_vector() - u_ref) / np.linalg.norm(u_ref) < 1e-10
    assert minfo["pc_bytes"] > 0


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_max_it_gives_diverged_its(cuda, precision):
    mesh, wl = _c1_small()
    uh, conv, _ = _problem(wl).solve(mesh, rtol=1e-14, atol=0.0, max_it=5, precision=precision)
    assert conv["converged_reason"] == -3          # KSP_DIVERGED_ITS
    assert conv["iteration_number"] == 5           # inner iterations in all outer steps


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_nonfinite_right_hand_side_stops(cuda, precision):
    mesh, wl = _c1_small()

    def f_nan(x, y, th):
        v = wl.f(x, y, th)
        return np.where((np.asarray(x) > 0.5) & (np.asarray(y) > 0.5), np.nan, v)

    uh, conv, _ = _problem(wl, f=f_nan).solve(mesh, rtol=1e-10, atol=0.0, precision=precision)
    assert conv["converged_reason"] == -9          # KSP_DIVERGED_NANORINF
    assert conv["iteration_number"] == 0


@pytest.mark.gpu
def test_nonfinite_coefficient_is_a_singular_block(cuda):
    """A NaN absorption in one column makes its cells' diagonal blocks
    non-invertible: rtx_bjacobi_factor returns RTX_ERR_NUMERIC (-4) naming a
    cell of that column, and the binding raises."""
    from rtx import _lib
    from rtx.krylov import BlockJacobi
    from rtx.operator import TransportOperator
    mesh, wl = _c1_small()

    def kappa_nan(x, y):
        k = wl.kappa(x, y)
        return np.where((np.asarray(x) < 0.2) & (np.asarray(y) < 0.2), np.nan, k)

    op = TransportOperator(mesh, kappa_nan, wl.sigma, wl.Phi)
    for method in (1, 0):                           # register sweeps, then LDS Gauss-Jordan
        _lib.check(op.lib.rtx_bjacobi_set_method(op.plan, method), "set method")
        with pytest.raises(_lib.RtxError) as ei:
            BlockJacobi(op)
        msg = str(ei.value)
        assert "(-4)" in msg and "singular diagonal block in cell" in msg, msg
        cell = int(msg.rsplit("cell", 1)[1].split()[0])
        # the first column (x, y < 0.25) holds cells 0..3
        assert 0 <= cell < 4, cell


@pytest.mark.gpu
def test_kaczmarz_is_reported_not_silent(cuda):
    from rt.problem import KaczmarzSubstituted
    mesh, wl = _c1_small()
    with pytest.warns(KaczmarzSubstituted):
        _, conv, _ = _problem(wl).solve(mesh, ksp_type="fgmres", pc_type="kaczmarz",
                                        rtol=1e-10, atol=0.0)
    assert conv["pc_type_used"] == "bjacobi"
    assert conv["converged_reason"] > 0


@pytest.mark.gpu
def test_operator_follows_in_place_refinement(cuda):
    """Vectors and matrices after ``ref_mesh`` on the same mesh object have
    the refined mesh's size and values (the reference rebuilds per call)."""
    mesh, wl = _c1_small()
    prob = _problem(wl)
    f0 = prob.forcing_vector(mesh)
    assert f0.size == mesh.get_ndof() == 4096
    mesh.ref_mesh(kind="spt", form="h")
    f1 = prob.forcing_vector(mesh)
    assert f1.size == mesh.get_ndof() == 16384
    ref = loops.forcing(mesh, wl.f)
    assert np.abs(f1 - ref).max() <= 1e-14 * np.abs(ref).max()
    M = prob.mass_matrix(mesh)
    assert M.shape == (16384, 16384)
    # an order change on the same keys is seen too
    for col in mesh.cols.values():
        if col.is_lf:
            col.ndofs = [3, 3]
    f2 = prob.forcing_vector(mesh)
    assert f2.size == mesh.get_ndof() == 9216
    assert np.abs(f2 - loops.forcing(mesh, wl.f)).max() <= 1e-14 * np.abs(f2).max()
