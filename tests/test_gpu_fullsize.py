"""Full-size solves of the apply and hp workloads (``src/rt/solve.py:49-149``):
the uniform absorber (32 x 32 x 16 cells, p = 4, 2,048,000 DOF) and the
batched uneven-order hp mesh with two strong scatterers (8,263,762 DOF).

Plain FP64 GMRES(50) + block-Jacobi and FP64 iterative refinement around the
reduced-precision inner solves both reach a true FP64 relative residual of
1e-8 (recomputed here with the FP64 operator, independently of the solver's
own bookkeeping), and the two solutions agree to 1e-6 in relative L2.  The
smaller meshes of these workloads are pinned against spsolve in
``test_gpu_solve.py``.
"""
import numpy as np
import pytest

from rtx import problems


def _solve_pair(mesh, wl):
    import torch
    from rtx import learned
    from rtx.krylov import BlockJacobi
    from rtx.operator import TransportOperator, OP_QUADW, OP_SYSTEM, OP_RESTRICT
    op = TransportOperator(mesh, wl.kappa, wl.sigma, wl.Phi)
    f = op.apply(torch.from_numpy(op.nodal(wl.f)).cuda(), ops=OP_QUADW)
    b = torch.from_numpy(op.nodal(wl.bcs)).cuda()
    bb = op.mask(b.clone(), keep_interior=False)
    rhs = op.mask(f - op.apply(bb, ops=OP_SYSTEM), keep_interior=True)
    bj = BlockJacobi(op, keep_f64=True, keep_fp8=True)
    out = {}
    for prec in ("fp64", "mixed"):
        print("solve", prec, op.ndof, flush=True)
        x, info = learned.solve(op, rhs, None, bj=bj, precision=prec, rtol=1e-8)
        r = rhs - op.apply(x, ops=OP_SYSTEM | OP_RESTRICT)
        rel = float(r.norm() / rhs.norm())
        assert info["converged_reason"] > 0, prec
        assert rel <= 1.05e-8, (prec, rel)
        out[prec] = x
    err = float((out["mixed"] - out["fp64"]).norm() / out["fp64"].norm())
    assert err < 1e-6, err
    return op.ndof


@pytest.mark.gpu
def test_c2_full_size_solves(cuda):
    mesh = problems.uniform_mesh([1.0, 1.0], [5, 5, 5], 4, 5)
    assert _solve_pair(mesh, problems.absorber_c2()) == 2_048_000


@pytest.mark.gpu
def test_c3_full_size_solves(cuda):
    mesh = problems.random_hp_mesh(0)
    assert _solve_pair(mesh, problems.two_scatterers_c3()) == 8_263_762
