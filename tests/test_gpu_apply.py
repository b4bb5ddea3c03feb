"""Device operator apply vs the oracle's assembled matrices (FP64 1e-12).

Reference: the assembled operator of ``src/rt/solve.py:49-52``
(A = (BC - IC) + M + S_ret) and its parts
(``mass_matrix.py``, ``interior_convection_matrix.py``,
``boundary_convection_matrix.py``, ``scattering_matrix.py``).
"""
import numpy as np
import pytest

from oracle import assemble as asm
from rtx import problems

TOL64 = 1e-12      # relative 2-norm, FP64 (matrix-free vs assembled)
TOL32 = 2e-6       # relative 2-norm, FP32 storage and arithmetic


def hp_mesh():
    m = problems.uniform_mesh([1.0, 1.0], [3, 3, 3], 2, 1)
    keys = sorted(m.cols)
    m.ref_col(keys[0], "spt", "hp")
    last = sorted(m.cols)[-1]
    m.ref_cell(last, sorted(m.cols[last].cells)[0], "hp")
    first = sorted(m.cols)[0]
    m.ref_cell(first, sorted(m.cols[first].cells)[2], "h")
    return m


def _rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def _dev(x):
    import torch
    return torch.from_numpy(np.ascontiguousarray(x)).to("cuda")


@pytest.mark.gpu
@pytest.mark.parametrize("path", [0, 1, 2], ids=["generic", "fused", "uniform"])
@pytest.mark.parametrize("case", ["hp", "c1"])
def test_apply_parts_match_oracle(cuda, case, path):
    from rtx.operator import (TransportOperator, OP_MASS, OP_SCAT, OP_ICONV, OP_BCONV,
                              OP_SYSTEM, OP_RESTRICT)
    if case == "hp" and path == 2:
        pytest.skip("the hp mesh is not a uniform order bin")
    if case == "hp":
        mesh = hp_mesh()
        wl = problems.two_scatterers()
    else:
        mesh = problems.uniform_mesh([1.0, 1.0], [4, 4, 4], 2, 2)
        wl = problems.manufactured()
    print("case", case, "path", path, "ndof", mesh.get_ndof(), flush=True)
    V = asm.assemble(mesh, wl.kappa, wl.sigma, wl.Phi)
    A = (V["BC"] - V["IC"] + V["M"] + V["S"]).tocsr()
    op = TransportOperator(mesh, wl.kappa, wl.sigma, wl.Phi)
    op.set_path(path)
    x = np.random.default_rng(0).uniform(-1.0, 1.0, op.ndof)
    xd = _dev(x)
    for name, ops, ref in [("M", OP_MASS, V["M"] @ x), ("S", OP_SCAT, V["S"] @ x),
                           ("IC", OP_ICONV, -(V["IC"] @ x)), ("BC", OP_BCONV, V["BC"] @ x),
                           ("A", OP_SYSTEM, A @ x)]:
        y = op.apply(xd, ops=ops).cpu().numpy()
        assert _rel(y, ref) < TOL64, (name, _rel(y, ref))
    mask = asm.intr_mask(mesh)
    y = op.apply(xd, ops=OP_SYSTEM | OP_RESTRICT).cpu().numpy()
    ref = mask * (A @ (mask * x))
    assert _rel(y, ref) < TOL64


TOLMIX = 5e-3     # relative 2-norm, BF16 scattering operands with FP32 accumulation


@pytest.mark.gpu
@pytest.mark.parametrize("path", [0, 1, 2], ids=["generic", "fused", "uniform"])
def test_apply_accumulate_and_reduced_precision(cuda, path):
    import torch
    from rtx.operator import (TransportOperator, OP_SYSTEM, OP_ACCUM, OP_SCAT, OP_RESTRICT,
                              PREC_F32, PREC_MIXED)
    if path == 2:
        mesh = problems.uniform_mesh([1.0, 1.0], [5, 5, 5], 4, 2)   # bin (5, 5, 16)
    else:
        mesh = hp_mesh()
    wl = problems.two_scatterers()
    V = asm.assemble(mesh, wl.kappa, wl.sigma, wl.Phi)
    A = (V["BC"] - V["IC"] + V["M"] + V["S"]).tocsr()
    op = TransportOperator(mesh, wl.kappa, wl.sigma, wl.Phi)
    op.set_path(path)
    rng = np.random.default_rng(1)
    x, y0 = rng.uniform(-1, 1, op.ndof), rng.uniform(-1, 1, op.ndof)
    yd = _dev(y0)
    op.apply(_dev(x), yd, ops=OP_SYSTEM | OP_ACCUM)
    assert _rel(yd.cpu().numpy(), y0 + A @ x) < TOL64
    x32 = _dev(x).to(torch.float32)
    y32 = op.apply(x32, precision=PREC_F32).cpu().double().numpy()
    assert _rel(y32, A @ x) < TOL32
    if path >= 1:
        ym = op.apply(x32, precision=PREC_MIXED).cpu().double().numpy()
        assert _rel(ym, A @ x) < TOLMIX, _rel(ym, A @ x)
        ys = op.apply(x32, ops=OP_SCAT, precision=PREC_MIXED).cpu().double().numpy()
        assert _rel(ys, V["S"] @ x) < 2e-2, _rel(ys, V["S"] @ x)
        mask = asm.intr_mask(mesh)
        yr = op.apply(x32, ops=OP_SYSTEM | OP_RESTRICT, precision=PREC_MIXED).cpu().double().numpy()
        assert np.all(yr[~mask] == 0.0)
        assert _rel(yr, mask * (A @ (mask * x))) < TOLMIX


@pytest.mark.gpu
def test_apply_c2_rows_and_linearity(cuda):
    """Full-size uniform 32x32x16, p=4 (2,048,000 DOF): rows of sampled
    columns against the oracle, plus linearity over the whole vector."""
    import torch
    from rtx.operator import TransportOperator, OP_SYSTEM
    mesh = problems.uniform_mesh([1.0, 1.0], [5, 5, 5], 4, 5)
    wl = problems.absorber_c2()
    op = TransportOperator(mesh, wl.kappa, wl.sigma, wl.Phi)
    assert op.ndof == 2_048_000
    rng = np.random.default_rng(0)
    x = rng.uniform(-1, 1, op.ndof)
    y = op.apply(_dev(x), ops=OP_SYSTEM).cpu().numpy()
    keys = sorted(mesh.cols)
    sample = [keys[0], keys[17], keys[500], keys[-1]]
    V = asm.assemble(mesh, wl.kappa, wl.sigma, wl.Phi, cols=sample)
    A = (V["BC"] - V["IC"] + V["M"] + V["S"]).tocsr()
    rows = np.unique(A.nonzero()[0])
    assert rows.size == 4 * 25 * 80
    assert _rel(y[rows], (A @ x)[rows]) < TOL64
    x2 = torch.from_numpy(rng.uniform(-1, 1, op.ndof)).cuda()
    x1 = _dev(x)
    lhs = op.apply(2.0 * x1 - 3.0 * x2, ops=OP_SYSTEM)
    rhs = 2.0 * op.apply(x1, ops=OP_SYSTEM) - 3.0 * op.apply(x2, ops=OP_SYSTEM)
    assert float((lhs - rhs).norm() / rhs.norm()) < 1e-13
