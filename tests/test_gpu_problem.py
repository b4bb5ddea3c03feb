"""rt.Problem (the reference's interface) on the device path.

Assembled matrices are compared with the oracle's literal-loop restatement of
the reference's assembly (same loops, same drop thresholds), vectors with the
oracle's, and ``solve`` with the exact discrete solution (spsolve) on the
same system.  Reference: ``src/rt/class_Problem.py:11-32`` and the modules it
binds.
"""
import numpy as np
import pytest

from oracle import loops, assemble as asm, system as osys
from rtx import problems


def _rel(a, b):
    d = a - b
    na = abs(d).max() if hasattr(d, "toarray") or hasattr(d, "max") else np.abs(d).max()
    return float(na)


@pytest.mark.gpu
def test_problem_matrices_match_reference_loops(cuda):
    from rt import Problem
    from test_gpu_apply import hp_mesh
    mesh = hp_mesh()
    wl = problems.two_scatterers()
    prob = Problem(wl.kappa, wl.sigma, wl.Phi, wl.f, wl.bcs_dirac)
    ref = {"M": loops.mass(mesh, wl.kappa), "S": loops.scattering(mesh, wl.sigma, wl.Phi),
           "IC": loops.interior_conv(mesh), "BC": loops.boundary_conv(mesh),
           "P": loops.preconditioner(mesh, wl.kappa, wl.sigma, wl.Phi)}
    got = {"M": prob.mass_matrix(mesh), "S": prob.scattering_matrix(mesh),
           "IC": prob.interior_convection_matrix(mesh),
           "BC": prob.boundary_convection_matrix(mesh),
           "P": prob.preconditioner_matrix(mesh)[0]}
    for k in ref:
        d = abs(got[k] - ref[k]).max()
        scale = abs(ref[k]).max()
        assert d <= 1e-13 * scale, (k, d, scale)
        # same nonzero structure, ignoring entries that are rounding-level
        # (e.g. l'_c(0) of a centre node: exactly 0 in one quadrature route,
        # ~1e-17 in another)
        big_g = int((abs(got[k]) > 1e-12 * scale).sum())
        big_r = int((abs(ref[k]) > 1e-12 * scale).sum())
        assert big_g == big_r, (k, big_g, big_r)


@pytest.mark.gpu
def test_problem_vectors_and_solve(cuda):
    from rt import Problem
    mesh = problems.uniform_mesh([1.0, 1.0], [4, 4, 4], 2, 2)
    wl = problems.manufactured()
    prob = Problem(wl.kappa, wl.sigma, wl.Phi, wl.f, wl.bcs_dirac)
    f_ref = loops.forcing(mesh, wl.f)
    b_ref = loops.bcs(mesh, wl.bcs)
    f = prob.forcing_vector(mesh)
    b = prob.boundary_conditions_vector(mesh)
    assert np.abs(f - f_ref).max() <= 1e-14 * np.abs(f_ref).max()
    assert np.abs(b - b_ref).max() <= 1e-15 * np.abs(b_ref).max()
    V = asm.assemble(mesh, wl.kappa, wl.sigma, wl.Phi)
    A = asm.system_matrix(V)
    mask = asm.intr_mask(mesh)
    u_ref, _, _ = osys.exact_solve(A, f_ref, b_ref, mask)
    uh, conv, minfo = prob.solve(mesh, rtol=1e-12, atol=0.0)
    u = uh.to_vector()
    assert conv["converged_reason"] > 0
    for key in ("converged_reason", "convergence_history", "iteration_number", "res_f",
                "res_best", "rtol", "atol", "divtol", "max_it"):
        assert key in conv
    assert np.linalg.norm(u - u_ref) / np.linalg.norm(u_ref) < 1e-8
    # FP64 refinement around reduced-precision inner solves ("mixed": FP32
    # convection, BF16 scattering, FP8 block-Jacobi; "fp32": FP32 throughout)
    # reaches the same solution
    for precision in ("mixed", "fp32"):
        uh2, conv2, _ = prob.solve(mesh, rtol=1e-11, atol=0.0, precision=precision)
        assert conv2["converged_reason"] > 0, precision
        assert conv2["precision"] == precision
        assert np.linalg.norm(uh2.to_vector() - u_ref) / np.linalg.norm(u_ref) < 1e-8, precision


def _intg_u2(u, n=60):
    """int u^2 over [0, 1]^2 x [0, 2 pi) by a 60^3 Gauss rule (the smooth
    manufactured u is resolved to rounding)."""
    xg, wg = np.polynomial.legendre.leggauss(n)
    x, wx = 0.5 * (xg + 1.0), 0.5 * wg
    th, wt = np.pi * (xg + 1.0), np.pi * wg
    X, Y, T = np.meshgrid(x, x, th, indexing="ij")
    W = wx[:, None, None] * wx[None, :, None] * wt[None, None, :]
    return float(np.sum(W * u(X, Y, T) ** 2))


def _anl_error(mesh, wl, intg):
    """Solve on the device and measure the relative L2 error the way the
    reference's experiment does (``Error_Indicator.error_analytic``,
    ``experiments/manufactured-solution/experiment.py:152-153``)."""
    from rt import Problem
    from amr.error_indicator import Error_Indicator
    prob = Problem(wl.kappa, wl.sigma, wl.Phi, wl.f, wl.bcs_dirac)
    uh, conv, _ = prob.solve(mesh, rtol=1e-12, atol=0.0)
    assert conv["converged_reason"] > 0
    return Error_Indicator(uh).error_analytic(wl.u, intg=intg).error


@pytest.mark.gpu
@pytest.mark.parametrize("n, levels", [(2, (3, 4, 5)), (3, (2, 3, 4))])
def test_manufactured_h_convergence_rate(cuda, n, levels):
    """Uniform h-refinement in space and angle together with n nodes per
    direction: the analytic L2 error falls at the order n of the degree n - 1
    basis, each observed rate within 0.5 of it (reference experiment:
    ``experiments/manufactured-solution/experiment.py:104-168``, error by
    ``error_analytic.py:22-149``).  A dropped face term or a wrong E entry
    lowers the order and fails this."""
    wl = problems.manufactured()
    intg = _intg_u2(wl.u)
    errs = []
    for lv in levels:
        mesh = problems.uniform_mesh([1.0, 1.0], [n, n, n], lv + 1, lv)
        errs.append(_anl_error(mesh, wl, intg))
    rates = [float(np.log2(errs[i] / errs[i + 1])) for i in range(len(errs) - 1)]
    assert all(abs(r - n) < 0.5 for r in rates), (errs, rates)


@pytest.mark.gpu
def test_manufactured_p_convergence(cuda):
    """p-refinement on a fixed 4 x 4 x 4-cell mesh: the error falls by a factor
    of at least 4 per added node (spectral convergence of the smooth u)."""
    wl = problems.manufactured()
    intg = _intg_u2(wl.u)
    errs = [_anl_error(problems.uniform_mesh([1.0, 1.0], [n, n, n], 2, 2), wl, intg)
            for n in (2, 3, 4, 5, 6)]
    ratios = [errs[i] / errs[i + 1] for i in range(len(errs) - 1)]
    assert all(r > 4.0 for r in ratios), (errs, ratios)
    assert errs[-1] < 5e-4, errs


@pytest.mark.gpu
def test_c1_stated_size_solve(cuda):
    """C1 at its stated 65,536 DOF: 16 x 16 x 4 cells at 4 nodes per direction
    (``Mesh(ndofs=[4, 4, 4])``, two angular and four spatial refinements),
    the manufactured problem, GMRES + block-Jacobi to 1e-10 as the
    configuration asks; FP64 and mixed refinement both match spsolve on the
    oracle's assembled interior system."""
    from rt import Problem
    mesh = problems.uniform_mesh([1.0, 1.0], [4, 4, 4], 2, 4)
    assert mesh.get_ndof() == 65536
    wl = problems.manufactured()
    V = asm.assemble(mesh, wl.kappa, wl.sigma, wl.Phi)
    A = asm.system_matrix(V)
    u_ref, _, _ = osys.exact_solve(A, asm.forcing(mesh, wl.f), asm.bcs(mesh, wl.bcs),
                                   asm.intr_mask(mesh))
    prob = Problem(wl.kappa, wl.sigma, wl.Phi, wl.f, wl.bcs_dirac)
    for precision in ("fp64", "mixed"):
        uh, conv, _ = prob.solve(mesh, rtol=1e-10, atol=0.0, precision=precision)
        assert conv["converged_reason"] > 0, precision
        err = np.linalg.norm(uh.to_vector() - u_ref) / np.linalg.norm(u_ref)
        assert err < 1e-8, (precision, err)
