"""The learned-preconditioner workload at full size: 64 x 64 columns x 16
angular cells, p = 4 (8,192,000 DOF), log kappa a Gaussian random field
(length scale 0.15, mean 0, variance 1), sigma / kappa ~ U(0.9, 0.999),
Henyey-Greenstein g = 0.5, unit source, vacuum inflow; the unseen
evaluation field (seed 10,000).

* Apply parity: rows of four sampled columns against the oracle's
  assembled matrix ((BC - IC) + M + S of ``src/rt/*_matrix.py``) at 1e-12,
  and linearity over the whole vector.
* Solves (``src/rt/solve.py:49-149``): plain FP64 GMRES(50) + block-Jacobi,
  mixed refinement, and mixed refinement with the learned coarse correction
  applied in every inner iteration (the FP64 judge off, so the correction
  runs through the whole solve) all reach a true FP64 relative residual of
  1e-8 and agree to 1e-6 in relative L2.
"""
import numpy as np
import pytest

from oracle import assemble as asm
from rtx import problems

TOL64 = 1e-12


def _mesh():
    return problems.uniform_mesh([1.0, 1.0], [5, 5, 5], 4, 6)


@pytest.fixture(scope="module")
def c5():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from rtx.operator import TransportOperator
    mesh = _mesh()
    wl = problems.grf_scatterer(10_000, n=64)
    op = TransportOperator(mesh, wl.kappa, wl.sigma, wl.Phi)
    assert op.ndof == 8_192_000
    return mesh, wl, op


@pytest.mark.gpu
def test_c5_apply_rows_and_linearity(cuda, c5):
    import torch
    from rtx.operator import OP_SYSTEM
    mesh, wl, op = c5
    rng = np.random.default_rng(5)
    x = rng.uniform(-1, 1, op.ndof)
    y = op.apply(torch.from_numpy(x).cuda(), ops=OP_SYSTEM).cpu().numpy()
    keys = sorted(mesh.cols)
    sample = [keys[0], keys[63], keys[2080], keys[-1]]
    V = asm.assemble(mesh, wl.kappa, wl.sigma, wl.Phi, cols=sample)
    A = (V["BC"] - V["IC"] + V["M"] + V["S"]).tocsr()
    rows = np.unique(A.nonzero()[0])
    assert rows.size == 4 * 25 * 80
    ref = (A @ x)[rows]
    err = np.linalg.norm(y[rows] - ref) / np.linalg.norm(ref)
    assert err < TOL64, err
    x1 = torch.from_numpy(x).cuda()
    x2 = torch.from_numpy(rng.uniform(-1, 1, op.ndof)).cuda()
    lhs = op.apply(2.0 * x1 - 3.0 * x2, ops=OP_SYSTEM)
    rhs = 2.0 * op.apply(x1, ops=OP_SYSTEM) - 3.0 * op.apply(x2, ops=OP_SYSTEM)
    assert float((lhs - rhs).norm() / rhs.norm()) < 1e-13


@pytest.mark.gpu
def test_c5_fp64_mixed_and_learned_solves_agree(cuda, c5):
    import torch
    from rtx import learned
    from rtx.krylov import BlockJacobi
    from rtx.operator import OP_QUADW
    from rtx._lib import PREC_MIXED
    mesh, wl, op = c5
    op.set_coeffs(wl.kappa, wl.sigma)
    f = op.apply(torch.from_numpy(op.nodal(wl.f)).cuda(), ops=OP_QUADW)
    rhs = op.mask(f, keep_interior=True)
    bj = BlockJacobi(op, keep_f64=True, keep_fp8=True)
    print("fp64 solve", flush=True)
    x64, i64 = learned.solve(op, rhs, None, bj=bj, precision="fp64", rtol=1e-8)
    print("mixed solve", flush=True)
    xm, im = learned.solve(op, rhs, None, bj=bj, precision="mixed", rtol=1e-8)
    for info in (i64, im):
        assert info["converged_reason"] > 0 and info["rel_residual"] <= 1e-8
    assert float((xm - x64).norm() / x64.norm()) < 1e-6
    # a coarse model trained on solver pairs of four other fields of the family
    print("training pairs", flush=True)
    grid = learned.CoarseGrid(op, 64)
    field = lambda s: problems.grf_scatterer(s, n=64)
    X, Y = learned.generate_pairs(op, grid, range(4), cycles=20, restart=10, skip=1,
                                  field=field, direction_stride=2)
    nc = grid.nc
    mlp = learned.MLP([5 * nc, 128, 128, 128, nc], op.device, seed=0)
    mlp.n_moments = 3 * nc
    xs, ys, hist = learned.train(mlp, X, Y, epochs=40, batch=64, lr=1e-3)
    del X, Y
    op.set_coeffs(wl.kappa, wl.sigma)
    model = learned.CoarseModel(mlp, 3 * nc, xs, ys)
    model.set_coefficients(grid.coefficient_channels(wl.kappa, wl.sigma))
    pc = learned.TwoLevelPC(op, bj, grid, model, apply_precision=PREC_MIXED)
    print("learned + mixed solve", flush=True)
    xl, il = learned.solve(op, rhs, None, pc=pc, bj=bj, precision="mixed", rtol=1e-8,
                           judge=False)
    assert il["converged_reason"] > 0 and il["rel_residual"] <= 1e-8
    assert not il["fallback"] and pc.calls > 0
    assert float((xl - x64).norm() / x64.norm()) < 1e-6
