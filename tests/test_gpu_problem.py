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


@pytest.mark.gpu
def test_manufactured_solution_converges(cuda):
    """Discrete solution approaches the manufactured u under h-refinement at a
    rate consistent with the order (reference experiment:
    experiments/manufactured-solution)."""
    from rt import Problem
    from dg.projection import Projection
    wl = problems.manufactured()
    errs = []
    for nref in (1, 2, 3):
        mesh = problems.uniform_mesh([1.0, 1.0], [3, 3, 3], 2 + nref, nref + 1)
        prob = Problem(wl.kappa, wl.sigma, wl.Phi, wl.f, wl.bcs_dirac)
        uh, conv, _ = prob.solve(mesh, rtol=1e-12, atol=0.0)
        u_ex = Projection(mesh, wl.u).to_vector()
        errs.append(np.linalg.norm(uh.to_vector() - u_ex) / np.linalg.norm(u_ex))
    rates = [np.log2(errs[i] / errs[i + 1]) for i in range(len(errs) - 1)]
    assert errs[-1] < errs[0]
    assert rates[-1] > 1.5, (errs, rates)
