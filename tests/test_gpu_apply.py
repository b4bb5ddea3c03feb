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
@pytest.mark.parametrize("path", [0, 1, 2, 3], ids=["generic", "fused", "uniform", "column"])
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
@pytest.mark.parametrize("path", [0, 1, 2, 3], ids=["generic", "fused", "uniform", "column"])
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


def small_random_hp(seed):
    """A small mesh of the batched uneven-order kind: 4 x 4 x 4 base, one
    random spatial and one angular refinement round, node counts 3..9 per
    column and per cell (about 60,000 DOF)."""
    return problems.random_hp_mesh(seed, base_spt=2, base_ang=2, extra_spt=1, extra_ang=1,
                                   p_col=0.5, p_cell=0.5)


@pytest.mark.gpu
@pytest.mark.parametrize("path", [0, 1, 3], ids=["generic", "fused", "column"])
@pytest.mark.parametrize("seed", [1, 2])
def test_apply_random_hp_matches_oracle(cuda, seed, path):
    """Uneven orders on both sides of every face and level jumps in space and
    angle: every operator part against the oracle's assembled matrices."""
    from rtx.operator import (TransportOperator, OP_MASS, OP_SCAT, OP_ICONV, OP_BCONV,
                              OP_SYSTEM, OP_RESTRICT)
    mesh = small_random_hp(seed)
    wl = problems.two_scatterers_c3()
    print("random hp seed", seed, "path", path, "ndof", mesh.get_ndof(), flush=True)
    V = asm.assemble(mesh, wl.kappa, wl.sigma, wl.Phi)
    A = (V["BC"] - V["IC"] + V["M"] + V["S"]).tocsr()
    op = TransportOperator(mesh, wl.kappa, wl.sigma, wl.Phi)
    op.set_path(path)
    x = np.random.default_rng(seed).uniform(-1.0, 1.0, op.ndof)
    xd = _dev(x)
    for name, ops, ref in [("M", OP_MASS, V["M"] @ x), ("S", OP_SCAT, V["S"] @ x),
                           ("IC", OP_ICONV, -(V["IC"] @ x)), ("BC", OP_BCONV, V["BC"] @ x),
                           ("A", OP_SYSTEM, A @ x)]:
        y = op.apply(xd, ops=ops).cpu().numpy()
        assert _rel(y, ref) < TOL64, (name, _rel(y, ref))
    mask = asm.intr_mask(mesh)
    y = op.apply(xd, ops=OP_SYSTEM | OP_RESTRICT).cpu().numpy()
    assert _rel(y, mask * (A @ (mask * x))) < TOL64


@pytest.mark.gpu
@pytest.mark.parametrize("path", [-1, 3], ids=["auto", "column"])
def test_apply_c3_rows_and_linearity(cuda, path):
    """Full-size batched uneven-order hp mesh (16 x 16 x 8 base, random
    refinement, node counts 3..9: 8,263,762 DOF): rows of sampled columns
    (the largest column among them) against the oracle, linearity over the
    whole vector."""
    import torch
    from rtx.operator import TransportOperator, OP_SYSTEM
    mesh = problems.random_hp_mesh(0)
    wl = problems.two_scatterers_c3()
    op = TransportOperator(mesh, wl.kappa, wl.sigma, wl.Phi)
    op.set_path(path)
    print("c3 path", path, flush=True)
    assert op.ndof == 8_263_762
    rng = np.random.default_rng(0)
    x = rng.uniform(-1, 1, op.ndof)
    y = op.apply(_dev(x), ops=OP_SYSTEM).cpu().numpy()
    keys = sorted(mesh.cols)
    size = {k: mesh.cols[k].ndofs[0] ** 2 * sum(c.ndofs[0] for c in mesh.cols[k].cells.values())
            for k in keys}
    sample = [keys[0], keys[len(keys) // 2], keys[-1], max(keys, key=lambda k: size[k])]
    # the exact discrete operator (no drop thresholds): the device path applies it
    V = asm.assemble(mesh, wl.kappa, wl.sigma, wl.Phi, cols=sample, drop=False)
    A = (V["BC"] - V["IC"] + V["M"] + V["S"]).tocsr()
    rows = np.unique(A.nonzero()[0])
    assert rows.size == sum(size[k] for k in set(sample))
    assert _rel(y[rows], (A @ x)[rows]) < TOL64
    # the reference's assembled matrix drops |S| <= 1e-14 and |BC| <= 1e-15;
    # at this mesh's smallest cells genuine entries fall below that, and the
    # whole difference is those dropped entries
    Vr = asm.assemble(mesh, wl.kappa, wl.sigma, wl.Phi, cols=sample)
    Ar = (Vr["BC"] - Vr["IC"] + Vr["M"] + Vr["S"]).tocsr()
    Dd = (A - Ar).tocsr()
    Dd.eliminate_zeros()
    assert Dd.nnz > 0 and abs(Dd).max() <= 1e-14
    assert _rel(y[rows] - (Dd @ x)[rows], (Ar @ x)[rows]) < TOL64
    x2 = torch.from_numpy(rng.uniform(-1, 1, op.ndof)).cuda()
    x1 = _dev(x)
    lhs = op.apply(2.0 * x1 - 3.0 * x2, ops=OP_SYSTEM)
    rhs = 2.0 * op.apply(x1, ops=OP_SYSTEM) - 3.0 * op.apply(x2, ops=OP_SYSTEM)
    assert float((lhs - rhs).norm() / rhs.norm()) < 1e-13


@pytest.mark.gpu
@pytest.mark.parametrize("n,nref_ang,nref_spt", [(5, 4, 3), (4, 2, 2)], ids=["p4", "p3"])
def test_uniform_apply_bit_reproducible(cuda, n, nref_ang, nref_spt):
    """The uniform kernel adds every output in a fixed order (face fluxes
    staged, no atomics): repeated FP64 and mixed applies are bit-identical, and
    so are the FP64 applies of every valid part count (the per-DOF arithmetic
    does not depend on how a column is split)."""
    import torch
    from rtx.operator import TransportOperator, OP_SYSTEM, OP_RESTRICT, PREC_MIXED
    from rtx import RtxError
    mesh = problems.uniform_mesh([1.0, 1.0], [n, n, n], nref_ang, nref_spt)
    wl = problems.absorber_c2()
    op = TransportOperator(mesh, wl.kappa, wl.sigma, wl.Phi)
    assert op.plan_info()["uniform"]
    x = _dev(np.random.default_rng(9).uniform(-1.0, 1.0, op.ndof))
    ops = OP_SYSTEM | OP_RESTRICT
    ref = None
    ran = 0
    for H in (1, 2, 4):
        try:
            op.set_uniform_parts(H)
        except RtxError:
            continue
        ys = [op.apply(x, ops=ops).clone() for _ in range(3)]
        for y in ys[1:]:
            assert torch.equal(y, ys[0]), H
        if ref is None:
            ref = ys[0]
        assert torch.equal(ys[0], ref), H
        ym = [op.apply(x.float(), ops=ops, precision=PREC_MIXED).clone() for _ in range(3)]
        for y in ym[1:]:
            assert torch.equal(y, ym[0]), H
        ran += 1
    assert ran >= 2
    op.set_uniform_parts(0)


@pytest.mark.gpu
@pytest.mark.parametrize("n,nref_ang,nref_spt", [(5, 4, 2), (4, 2, 2), (3, 3, 2), (5, 3, 2)],
                         ids=["5-5-16", "4-4-4", "3-3-8", "5-5-8"])
def test_uniform_output_layout_kernel(cuda, n, nref_ang, nref_spt):
    """The default FP64 uniform kernel (every term of an output added where
    the scattering MFMA leaves it) against the line-pass kernel (one part per
    column) on compiled shapes: the same operations in the same order, so the
    results agree to the last bits (1e-15 relative; equality is reported);
    both match the oracle at 1e-12; repeated applies are bit-identical, also
    into a y that is only 8-byte aligned (element-wise write-back instead of
    16-byte copies); accumulate mode adds exactly.  FP32 and mixed default to
    the line-pass kernel with one part."""
    import torch
    from rtx.operator import (TransportOperator, OP_SYSTEM, OP_RESTRICT, OP_ACCUM, PREC_F32,
                              PREC_MIXED)
    mesh = problems.uniform_mesh([1.0, 1.0], [n, n, n], nref_ang, nref_spt)
    wl = problems.absorber_c2()
    op = TransportOperator(mesh, wl.kappa, wl.sigma, wl.Phi)
    info = op.plan_info()
    assert info["uniform"] and info["uniform_output_layout"]
    print("shape", n, nref_ang, "ndof", op.ndof, flush=True)
    x = np.random.default_rng(21).uniform(-1.0, 1.0, op.ndof)
    xd = _dev(x)
    ops = OP_SYSTEM | OP_RESTRICT
    ye = [op.apply(xd, ops=ops).clone() for _ in range(3)]
    for yy in ye[1:]:
        assert torch.equal(yy, ye[0])
    # y only 8-byte aligned: the element-wise write-back, the same values
    buf = torch.full((op.ndof + 1,), float("nan"), dtype=torch.float64, device="cuda")
    op.apply(xd, buf[1:], ops=ops)
    assert torch.equal(buf[1:], ye[0]) and torch.isnan(buf[0])
    op.set_uniform_parts(1)
    assert not op.plan_info()["uniform_output_layout"]
    yl = op.apply(xd, ops=ops)
    yl32 = op.apply(xd.float(), ops=ops, precision=PREC_F32)
    ylm = op.apply(xd.float(), ops=ops, precision=PREC_MIXED)
    op.set_uniform_parts(0)
    err = float((ye[0] - yl).norm() / yl.norm())
    print("output-layout vs line-pass FP64: rel", err, "equal", bool(torch.equal(ye[0], yl)), flush=True)
    assert err < 1e-15, err
    V = asm.assemble(mesh, wl.kappa, wl.sigma, wl.Phi)
    A = (V["BC"] - V["IC"] + V["M"] + V["S"]).tocsr()
    mask = asm.intr_mask(mesh)
    ref = mask * (A @ (mask * x))
    assert _rel(ye[0].cpu().numpy(), ref) < TOL64
    y32 = op.apply(xd.float(), ops=ops, precision=PREC_F32)
    assert float((y32 - yl32).norm() / yl32.norm()) < 1e-6
    ym = op.apply(xd.float(), ops=ops, precision=PREC_MIXED)
    assert float((ym - ylm).norm() / ylm.norm()) < 1e-6
    y0 = torch.rand_like(xd)
    y1 = y0.clone()
    op.apply(xd, y1, ops=OP_SYSTEM | OP_ACCUM)
    assert torch.allclose(y1 - y0, op.apply(xd, ops=OP_SYSTEM), rtol=1e-13, atol=1e-13)


@pytest.mark.gpu
def test_hp_default_apply_bit_reproducible(cuda):
    """On an hp mesh the default launch shapes (the hp column kernel: one
    workgroup per column, every output written by one lane; the generic
    kernels for columns it does not take) give bit-identical repeated
    applies, in FP64 and mixed precision."""
    import torch
    from rtx.operator import TransportOperator, OP_SYSTEM, OP_RESTRICT, PREC_MIXED
    mesh = small_random_hp(2)
    wl = problems.two_scatterers_c3()
    op = TransportOperator(mesh, wl.kappa, wl.sigma, wl.Phi)
    info = op.plan_info()
    assert info["fused_bins"] == [] and info["column_kernel"]
    assert info["column_kernel_cols"] == op.t.ncol
    x = _dev(np.random.default_rng(11).uniform(-1.0, 1.0, op.ndof))
    for xx, prec in ((x, 0), (x.float(), PREC_MIXED)):
        ys = [op.apply(xx, ops=OP_SYSTEM | OP_RESTRICT, precision=prec).clone() for _ in range(3)]
        for y in ys[1:]:
            assert torch.equal(y, ys[0]), prec


@pytest.mark.gpu
def test_column_kernel_bit_reproducible(cuda):
    """The hp column kernel adds every output in a fixed order (one owner per
    output, no atomics): repeated FP64, FP32 and mixed applies are
    bit-identical, and accumulate mode adds to y exactly."""
    import torch
    from rtx.operator import (TransportOperator, OP_SYSTEM, OP_RESTRICT, OP_ACCUM, PREC_F32,
                              PREC_MIXED)
    mesh = small_random_hp(2)
    wl = problems.two_scatterers_c3()
    op = TransportOperator(mesh, wl.kappa, wl.sigma, wl.Phi)
    op.set_path(3)
    x = _dev(np.random.default_rng(12).uniform(-1.0, 1.0, op.ndof))
    for xx, prec in ((x, 0), (x.float(), PREC_F32), (x.float(), PREC_MIXED)):
        ys = [op.apply(xx, ops=OP_SYSTEM | OP_RESTRICT, precision=prec).clone() for _ in range(3)]
        for y in ys[1:]:
            assert torch.equal(y, ys[0]), prec
    y0 = torch.rand_like(x)
    y1 = y0.clone()
    op.apply(x, y1, ops=OP_SYSTEM | OP_ACCUM)
    assert torch.allclose(y1 - y0, op.apply(x, ops=OP_SYSTEM), rtol=1e-13, atol=1e-13)


@pytest.mark.gpu
def test_uniform_parts_and_autotune(cuda):
    """Every valid part count of the uniform kernel, and the shapes the
    autotuner picks for the uniform kernel and the fused order bins, give the
    oracle's operator."""
    from rtx.operator import TransportOperator, OP_SYSTEM, OP_RESTRICT
    from rtx import RtxError
    mesh = problems.uniform_mesh([1.0, 1.0], [4, 4, 4], 2, 2)
    wl = problems.manufactured()
    V = asm.assemble(mesh, wl.kappa, wl.sigma, wl.Phi)
    A = (V["BC"] - V["IC"] + V["M"] + V["S"]).tocsr()
    op = TransportOperator(mesh, wl.kappa, wl.sigma, wl.Phi)
    assert op.plan_info()["uniform"]
    x = np.random.default_rng(3).uniform(-1.0, 1.0, op.ndof)
    xd = _dev(x)
    ran = []
    for H in (1, 2, 4):
        try:
            op.set_uniform_parts(H)
        except RtxError:
            continue
        print("uniform parts", H, flush=True)
        assert op.plan_info()["uniform_parts"] == H
        y = op.apply(xd, ops=OP_SYSTEM).cpu().numpy()
        assert _rel(y, A @ x) < TOL64, (H, _rel(y, A @ x))
        ran.append(H)
    assert len(ran) >= 2
    op.set_uniform_parts(0)
    info = op.autotune(reps=3)
    assert info["uniform_parts"] in ran
    y = op.apply(xd, ops=OP_SYSTEM).cpu().numpy()
    assert _rel(y, A @ x) < TOL64

    mesh = small_random_hp(1)
    wl = problems.two_scatterers_c3()
    V = asm.assemble(mesh, wl.kappa, wl.sigma, wl.Phi)
    A = (V["BC"] - V["IC"] + V["M"] + V["S"]).tocsr()
    op = TransportOperator(mesh, wl.kappa, wl.sigma, wl.Phi)
    before = op.plan_info()
    # by default every order bin runs on the generic kernels (no atomics)
    assert before["fused_bins"] == [] and before["bins_on_generic"]
    info = op.autotune(reps=3)
    print("hp bins", before, "->", info, flush=True)
    assert sorted([n for _, n in info["fused_bins"]] + info["bins_on_generic"]) == \
        sorted(before["bins_on_generic"])
    x = np.random.default_rng(4).uniform(-1.0, 1.0, op.ndof)
    mask = asm.intr_mask(mesh)
    ref = mask * (A @ (mask * x))
    for path in (-1, 0, 1, 3):   # tuned, all generic, all fused, column kernel
        op.set_path(path)
        y = op.apply(_dev(x), ops=OP_SYSTEM | OP_RESTRICT).cpu().numpy()
        assert _rel(y, ref) < TOL64, (path, _rel(y, ref))
    # bins moved to the generic kernels by hand: the generic set is rebuilt
    op.set_path(-1)
    for flags in ([1] * 6, [0, 1] * 3, [0] * 6):
        for b, g in enumerate(flags):
            op.set_bin_generic(b, g)
        y = op.apply(_dev(x), ops=OP_SYSTEM | OP_RESTRICT).cpu().numpy()
        assert _rel(y, ref) < TOL64, (flags, _rel(y, ref))
    assert op.plan_info()["bins_on_generic"] == []


@pytest.mark.gpu
@pytest.mark.parametrize("pbcs", [(True, False), (True, True)], ids=["periodic-x", "periodic-xy"])
@pytest.mark.parametrize("path", [-1, 0, 3], ids=["auto", "generic", "column"])
def test_apply_periodic_matches_oracle(cuda, pbcs, path):
    """Periodic meshes (``Mesh(pbcs=...)``, class_Mesh.py:46-53): columns on
    opposite walls are face neighbours and no wall in a periodic direction
    is an inflow boundary; every part against the oracle's matrices."""
    from rtx.operator import (TransportOperator, OP_MASS, OP_SCAT, OP_ICONV, OP_BCONV,
                              OP_SYSTEM, OP_RESTRICT)
    mesh = problems.uniform_mesh([1.0, 1.0], [4, 4, 4], 2, 2, pbcs=pbcs)
    if path in (0, 3):
        keys = sorted(mesh.cols)
        mesh.ref_col(keys[0], "spt", "p")
    wl = problems.manufactured()
    V = asm.assemble(mesh, wl.kappa, wl.sigma, wl.Phi)
    A = (V["BC"] - V["IC"] + V["M"] + V["S"]).tocsr()
    op = TransportOperator(mesh, wl.kappa, wl.sigma, wl.Phi)
    op.set_path(path)
    x = np.random.default_rng(9).uniform(-1.0, 1.0, op.ndof)
    xd = _dev(x)
    for name, ops, ref in [("BC", OP_BCONV, V["BC"] @ x), ("A", OP_SYSTEM, A @ x)]:
        y = op.apply(xd, ops=ops).cpu().numpy()
        assert _rel(y, ref) < TOL64, (name, _rel(y, ref))
    mask = asm.intr_mask(mesh)
    y = op.apply(xd, ops=OP_SYSTEM | OP_RESTRICT).cpu().numpy()
    assert _rel(y, mask * (A @ (mask * x))) < TOL64


@pytest.mark.gpu
def test_copy_f64_measurement_helper(cuda):
    """rtx_copy_f64 (the counter-calibration copy) copies exactly, for sizes
    that are not a multiple of its workgroup, and does nothing for n = 0."""
    import ctypes as C
    import torch
    from rtx import _lib
    lib = _lib.load()
    for n in (1, 255, 257, 1 << 20):
        src = torch.rand(n, dtype=torch.float64, device="cuda")
        dst = torch.full((n + 1,), -7.0, dtype=torch.float64, device="cuda")
        _lib.check(lib.rtx_copy_f64(C.c_void_p(src.data_ptr()), C.c_void_p(dst.data_ptr()), n,
                                    None), "rtx_copy_f64")
        torch.cuda.synchronize()
        assert torch.equal(dst[:n], src) and float(dst[n]) == -7.0, n
    _lib.check(lib.rtx_copy_f64(C.c_void_p(src.data_ptr()), C.c_void_p(dst.data_ptr()), 0, None),
               "rtx_copy_f64")
    assert lib.rtx_copy_f64(C.c_void_p(src.data_ptr()), C.c_void_p(dst.data_ptr()), -1, None) != 0
