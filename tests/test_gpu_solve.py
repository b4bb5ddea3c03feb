"""Device block-Jacobi and GMRES vs the oracle's exact interior solve.

Reference: ``src/rt/solve.py:49-149`` (A_II x_I = f_I - A_IB b_B, GMRES +
block-Jacobi).  The solution is pinned by ``spsolve`` on the oracle's
assembled interior system; iteration counts are not compared (PETSc absent).
"""
import numpy as np
import pytest

from oracle import assemble as asm, system as osys
from rtx import problems

TOL_SOL = 1e-8     # relative 2-norm of the solution error vs spsolve


def _dev(x):
    import torch
    return torch.from_numpy(np.ascontiguousarray(x)).to("cuda")


def _c1_small():
    mesh = problems.uniform_mesh([1.0, 1.0], [4, 4, 4], 2, 2)
    return mesh, problems.manufactured()


def _oracle_system(mesh, wl):
    V = asm.assemble(mesh, wl.kappa, wl.sigma, wl.Phi)
    A = asm.system_matrix(V)
    f = asm.forcing(mesh, wl.f)
    b = asm.bcs(mesh, wl.bcs)
    mask = asm.intr_mask(mesh)
    return A, f, b, mask


@pytest.mark.gpu
@pytest.mark.parametrize("method", [0, 1], ids=["lds", "register"])
@pytest.mark.parametrize("case", ["hp", "p4"])
def test_block_jacobi_matches_oracle(cuda, method, case):
    """Both build methods (LDS Gauss-Jordan with row interchanges, register
    sweeps with implicit pivoting) against the oracle's dense inverses, on
    an hp mesh (blocks of 27..64 DOFs) and a p = 4 mesh (125 DOFs)."""
    from rtx.operator import TransportOperator
    from rtx.krylov import BlockJacobi
    from rtx import _lib
    from test_gpu_apply import hp_mesh
    if case == "hp":
        mesh = hp_mesh()
    else:
        mesh = problems.uniform_mesh([1.0, 1.0], [5, 5, 5], 2, 2)
    wl = problems.two_scatterers()
    A, f, b, mask = _oracle_system(mesh, wl)
    op = TransportOperator(mesh, wl.kappa, wl.sigma, wl.Phi)
    _lib.check(op.lib.rtx_bjacobi_set_method(op.plan, method), "set method")
    pc = BlockJacobi(op, restrict=True, keep_f32=True)
    invs = osys.block_jacobi_inverses(A, mesh, mask)
    r = np.random.default_rng(3).uniform(-1, 1, op.ndof)
    z_ref = osys.block_jacobi_apply(invs, mesh, r)
    z = pc.apply(_dev(r), _dev(np.zeros_like(r))).cpu().numpy()
    assert np.linalg.norm(z - z_ref) / np.linalg.norm(z_ref) < 1e-11
    import torch
    z32 = pc.apply(_dev(r).float(), torch.zeros(op.ndof, dtype=torch.float32, device="cuda"),
                   precision=1).cpu().double().numpy()
    assert np.linalg.norm(z32 - z_ref) / np.linalg.norm(z_ref) < 1e-5


@pytest.mark.gpu
@pytest.mark.parametrize("case", ["c1", "hp"])
def test_gmres_solution_matches_spsolve(cuda, case):
    from rtx.operator import TransportOperator
    from rtx.krylov import BlockJacobi, InteriorSystem, gmres
    if case == "c1":
        mesh, wl = _c1_small()
    else:
        from test_gpu_apply import hp_mesh
        mesh, wl = hp_mesh(), problems.two_scatterers()
    A, f, b, mask = _oracle_system(mesh, wl)
    u_ref, A_II, rhs = osys.exact_solve(A, f, b, mask)
    op = TransportOperator(mesh, wl.kappa, wl.sigma, wl.Phi)
    # device right-hand side: f_I - A_IB b_B in full-length form
    bb = np.where(mask, 0.0, b)
    rhs_full = np.where(mask, f - A @ bb, 0.0)
    sysd = InteriorSystem(op, pc=BlockJacobi(op))
    x, info = gmres(sysd, _dev(rhs_full), rtol=1e-12, restart=30, max_it=2000)
    u = np.where(mask, x.cpu().numpy(), b)
    assert info["converged_reason"] > 0, info["converged_reason"]
    assert info["rel_residual"] <= 1e-12
    err = np.linalg.norm(u - u_ref) / np.linalg.norm(u_ref)
    assert err < TOL_SOL, err


@pytest.mark.gpu
def test_refinement_mixed_reaches_fp64_residual(cuda):
    from rtx.operator import TransportOperator, PREC_F32
    from rtx.krylov import BlockJacobi, InteriorSystem, refine
    mesh, wl = _c1_small()
    A, f, b, mask = _oracle_system(mesh, wl)
    u_ref, _, _ = osys.exact_solve(A, f, b, mask)
    op = TransportOperator(mesh, wl.kappa, wl.sigma, wl.Phi)
    pc = BlockJacobi(op, keep_f32=True)
    hi = InteriorSystem(op, pc=pc)
    lo = InteriorSystem(op, precision=PREC_F32, pc=pc, pc_precision=PREC_F32)
    bb = np.where(mask, 0.0, b)
    rhs_full = np.where(mask, f - A @ bb, 0.0)
    x, info = refine(hi, lo, _dev(rhs_full), rtol=1e-10)
    assert info["converged_reason"] > 0
    assert info["rel_residual"] <= 1e-10
    u = np.where(mask, x.cpu().numpy(), b)
    assert np.linalg.norm(u - u_ref) / np.linalg.norm(u_ref) < TOL_SOL


@pytest.mark.gpu
@pytest.mark.parametrize("variant", ["plain", "flexible", "mixed", "mixed_f64_basis",
                                     "mixed_f16_basis", "mixed_f16_basis_fp4", "hp_plain",
                                     "hp_mixed", "hp_mixed_f16_basis", "hp_mixed_f16_basis_fp4"])
def test_graph_replay_matches_eager(cuda, variant):
    """Chunks of GMRES iterations replayed as HIP graphs reproduce the eager
    run exactly (same kernels, same order): iterate, residual history and
    iteration count, over many restarts (so every chunk graph is replayed
    several times, with eager residual / update work and host syncs between
    the replays)."""
    import torch
    from rtx.operator import TransportOperator
    from rtx.krylov import BlockJacobi, InteriorSystem, gmres, refine
    from rtx._lib import PREC_F64, PREC_F32, PREC_MIXED, PREC_FP8, PREC_FP4
    fp4 = variant.endswith("_fp4")
    variant = variant[:-4] if fp4 else variant
    if variant.startswith("hp"):
        # hp mesh on the default (generic, atomic-free) launch shapes
        from test_gpu_apply import small_random_hp
        mesh, wl = small_random_hp(1), problems.two_scatterers_c3()
        variant = variant[3:]
        tol, max_it = 1e-6, 4000
    else:
        mesh, wl = _c1_small()
        tol, max_it = 1e-10, 500
    op = TransportOperator(mesh, wl.kappa, wl.sigma, wl.Phi)
    f16 = variant == "mixed_f16_basis"
    pc = BlockJacobi(op, keep_f32=not f16, keep_fp8=f16 and not fp4, keep_fp4=f16 and fp4)
    pc_lo = (PREC_FP4 if fp4 else PREC_FP8) if f16 else PREC_F32
    rhs = op.mask(_dev(np.random.default_rng(4).uniform(-1, 1, op.ndof)))
    runs = []
    for graphs in (False, True, True):
        if variant.startswith("mixed"):
            hi = InteriorSystem(op, PREC_F64, None, PREC_F64)
            lo = InteriorSystem(op, PREC_MIXED, pc, pc_lo)
            basis = {"mixed_f64_basis": "f64", "mixed_f16_basis": "f16"}.get(variant, "f32")
            x, info = refine(hi, lo, rhs, None, rtol=tol, inner_rtol=1e-2, restart=7,
                             max_inner=max_it, chunk=3, graphs=graphs, basis=basis)
            assert info["converged_reason"] > 0
            runs.append((x.cpu().numpy(), info["residual_history"], info["iteration_number"]))
        else:
            system = InteriorSystem(op, PREC_F64, pc, PREC_F64)
            x, info = gmres(system, rhs, None, rtol=tol, restart=7, max_it=max_it,
                            flexible=variant == "flexible", chunk=3, graphs=graphs)
            assert info["converged_reason"] > 0
            if graphs:
                assert info["graphs"] >= 2
            runs.append((x.cpu().numpy(), info["convergence_history"], info["iteration_number"]))
        torch.cuda.synchronize()
    for other in runs[1:]:
        assert other[2] == runs[0][2]
        assert np.array_equal(other[0], runs[0][0])
        assert list(other[1]) == list(runs[0][1])


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_random_hp_solve_matches_spsolve(cuda, precision):
    """Batched uneven-order hp mesh (node counts 3..9, cells of up to 729
    DOFs): the interior solve against spsolve of the oracle's system."""
    import torch
    from rtx.operator import TransportOperator, OP_SYSTEM, OP_QUADW
    from rtx import learned
    from test_gpu_apply import small_random_hp
    mesh = small_random_hp(1)
    wl = problems.two_scatterers_c3()
    A, f, b, mask = _oracle_system(mesh, wl)
    u_ref, A_II, rhs_ref = osys.exact_solve(A, f, b, mask)
    op = TransportOperator(mesh, wl.kappa, wl.sigma, wl.Phi)
    fd = op.apply(torch.from_numpy(op.nodal(wl.f)).cuda(), ops=OP_QUADW)
    bd = op.mask(torch.from_numpy(op.nodal(wl.bcs)).cuda(), keep_interior=False)
    rhs = op.mask(fd - op.apply(bd, ops=OP_SYSTEM), keep_interior=True)
    x, info = learned.solve(op, rhs, None, precision=precision, rtol=1e-10, inner_rtol=1e-3)
    assert info["converged_reason"] > 0
    u = (x + bd).cpu().numpy()
    assert np.linalg.norm(u - u_ref) / np.linalg.norm(u_ref) < TOL_SOL


@pytest.mark.gpu
def test_block_jacobi_split_cells_match_oracle(cuda):
    """Cells of more than 128 DOFs are split into groups of angles (every
    spatial node, g = 128 // (nx ny) angles per block); each block's dense
    inverse against the oracle's on a mesh with node counts 3..9."""
    import torch
    from rtx.operator import TransportOperator
    from rtx.krylov import BlockJacobi
    from test_gpu_apply import small_random_hp
    mesh = small_random_hp(2)
    wl = problems.two_scatterers_c3()
    A, f, b, mask = _oracle_system(mesh, wl)
    op = TransportOperator(mesh, wl.kappa, wl.sigma, wl.Phi)
    pc = BlockJacobi(op, restrict=True, keep_f32=True, keep_bf16=True)
    invs = osys.block_jacobi_inverses(A, mesh, mask, split=128)
    assert len(invs) > len(osys.cell_blocks(mesh))          # some cells were split
    r = np.random.default_rng(5).uniform(-1, 1, op.ndof)
    z_ref = osys.block_jacobi_apply(invs, mesh, r, split=128)
    z = pc.apply(_dev(r), _dev(np.zeros_like(r))).cpu().numpy()
    assert np.linalg.norm(z - z_ref) / np.linalg.norm(z_ref) < 1e-11
    r32 = _dev(r).float()
    z16 = pc.apply(r32, torch.zeros_like(r32), precision=2).cpu().double().numpy()
    assert np.linalg.norm(z16 - z_ref) / np.linalg.norm(z_ref) < 2e-2


@pytest.mark.gpu
def test_arnoldi_fp32_basis(cuda):
    """Arnoldi steps on an FP32 basis: the basis stays orthonormal to FP32
    accuracy, every input vector is reproduced by its Hessenberg column
    (w_j = sum_i h_ij v_i), and the columns agree with the FP64 step's."""
    import ctypes as C
    import torch
    from rtx import _lib
    from rtx.krylov import _vp
    lib = _lib.load()
    n, k = 200_003, 12
    rng = np.random.default_rng(5)
    W = rng.normal(size=(k, n))
    ws = C.c_int64()
    _lib.check(lib.rtx_arnoldi_workspace(k + 1, C.byref(ws)), "workspace")
    work = torch.zeros(int(ws.value), dtype=torch.float64, device="cuda")
    out = {}
    for dt, step in ((torch.float32, lib.rtx_arnoldi_step_f32),
                     (torch.float64, lib.rtx_arnoldi_step)):
        V = torch.zeros((k + 1, n), dtype=dt, device="cuda")
        V[0] = torch.from_numpy(W[0] / np.linalg.norm(W[0])).to(dt)
        Hs = []
        for j in range(1, k):
            w = torch.from_numpy(W[j]).to(dt).cuda()
            hcol = torch.zeros(k + 2, dtype=torch.float64, device="cuda")
            _lib.check(step(_vp(V), n, j, n, _vp(w), _vp(hcol), _vp(work), 0, None), "step")
            Hs.append(hcol[:j + 1].cpu().numpy())
        torch.cuda.synchronize()
        Vh = V[:k].double().cpu().numpy()
        out[dt] = (Vh, Hs)
    Vh, Hs = out[torch.float32]
    G = Vh @ Vh.T
    assert np.abs(G - np.eye(k)).max() < 1e-5
    for j in range(1, k):
        rec = Hs[j - 1] @ Vh[:j + 1]
        assert np.linalg.norm(rec - W[j]) < 1e-5 * np.linalg.norm(W[j])
        ref = out[torch.float64][1][j - 1]
        assert np.abs(Hs[j - 1] - ref).max() < 1e-5 * np.abs(ref).max()


@pytest.mark.gpu
def test_arnoldi_fp16_basis(cuda):
    """Arnoldi steps on an FP16 basis (FP32 w, FP64 dot products and
    Hessenberg column, new column rounded to FP16 after the normalisation):
    the basis is orthonormal to FP16 rounding (unit roundoff 2^-11), every
    input vector is reproduced by its Hessenberg column to that accuracy,
    and the columns agree with the FP64 step's."""
    import ctypes as C
    import torch
    from rtx import _lib
    from rtx.krylov import _vp
    lib = _lib.load()
    n, k = 200_003, 12
    rng = np.random.default_rng(5)
    W = rng.normal(size=(k, n))
    ws = C.c_int64()
    _lib.check(lib.rtx_arnoldi_workspace(k + 1, C.byref(ws)), "workspace")
    work = torch.zeros(int(ws.value), dtype=torch.float64, device="cuda")
    _lib.check(lib.rtx_gmres_state_size(k, C.byref(ws)), "state size")
    state = torch.zeros(int(ws.value), dtype=torch.float64, device="cuda")
    _lib.check(lib.rtx_gmres_begin(k, 1.0, _vp(state), None), "begin")
    V = torch.zeros((k + 1, n), dtype=torch.float16, device="cuda")
    V[0] = torch.from_numpy(W[0] / np.linalg.norm(W[0])).to(torch.float16)
    Hs = []
    for j in range(1, k):
        w = torch.from_numpy(W[j]).to(torch.float32).cuda()
        hcol = torch.zeros(k + 2, dtype=torch.float64, device="cuda")
        # step j - 1 of the cycle orthogonalises against j vectors
        _lib.check(lib.rtx_gmres_arnoldi(2, _vp(V), n, n, _vp(w), _vp(hcol), _vp(work), 0,
                                         j - 1, k, _vp(state), None), "fp16 step")
        Hs.append(hcol[:j + 1].cpu().numpy())
    torch.cuda.synchronize()
    Vh = V[:k].double().cpu().numpy()
    G = Vh @ Vh.T
    assert np.abs(G - np.eye(k)).max() < 3e-3, np.abs(G - np.eye(k)).max()
    # the FP64 step on the same (FP16-rounded) basis vectors and inputs
    Vd = torch.from_numpy(Vh).cuda()
    for j in range(1, k):
        rec = Hs[j - 1] @ Vh[:j + 1]
        assert np.linalg.norm(rec - W[j]) < 3e-3 * np.linalg.norm(W[j])
        V64 = torch.zeros((j + 1, n), dtype=torch.float64, device="cuda")
        V64[:j] = Vd[:j]
        w64 = torch.from_numpy(W[j].astype(np.float32).astype(np.float64)).cuda()
        h64 = torch.zeros(j + 2, dtype=torch.float64, device="cuda")
        _lib.check(lib.rtx_arnoldi_step(_vp(V64), n, j, n, _vp(w64), _vp(h64), _vp(work), 0,
                                        None), "fp64 step")
        ref = h64[:j + 1].cpu().numpy()
        assert np.abs(Hs[j - 1][:j] - ref[:j]).max() < 1e-9 * np.abs(ref).max()
        assert abs(Hs[j - 1][j] - ref[j]) < 1e-6 * np.abs(ref).max()


@pytest.mark.gpu
def test_fp8_apply_fp16_input(cuda):
    """The FP8 block-Jacobi apply from an FP16 r equals the apply from the
    same values in FP32 bit for bit (one conversion of each entry, the same
    arithmetic), on whole cells and on split cells."""
    import torch
    from rtx.operator import TransportOperator
    from rtx.krylov import BlockJacobi
    from rtx._lib import PREC_FP8
    from test_gpu_apply import small_random_hp
    for mesh, wl in ((_c1_small()[0], problems.manufactured()),
                     (small_random_hp(2), problems.two_scatterers_c3())):
        op = TransportOperator(mesh, wl.kappa, wl.sigma, wl.Phi)
        pc = BlockJacobi(op, keep_f64=False, keep_fp8=True)
        r16 = torch.from_numpy(np.random.default_rng(9).uniform(-1, 1, op.ndof)).to(
            torch.float16).cuda()
        z_h = pc.apply_fp8_h(r16, torch.zeros(op.ndof, dtype=torch.float32, device="cuda"))
        z_f = pc.apply(r16.float(), torch.zeros(op.ndof, dtype=torch.float32, device="cuda"),
                       PREC_FP8)
        assert torch.equal(z_h, z_f)


@pytest.mark.gpu
@pytest.mark.parametrize("bj", ["fp8", "fp4"])
def test_refine_fp16_basis_matches_spsolve(cuda, bj):
    """Mixed iterative refinement with an FP16 inner basis and FP8 or FP4
    block-Jacobi (FP4 is the default of mixed precision) reaches the same
    discrete solution as spsolve (FP64 outer residual 1e-10)."""
    from rtx.operator import TransportOperator
    from rtx.krylov import BlockJacobi, InteriorSystem, refine, pick_basis
    from rtx._lib import PREC_F64, PREC_MIXED, PREC_FP8, PREC_FP4
    mesh, wl = _c1_small()
    A, f, b, mask = _oracle_system(mesh, wl)
    u_ref, _, _ = osys.exact_solve(A, f, b, mask)
    op = TransportOperator(mesh, wl.kappa, wl.sigma, wl.Phi)
    pc = BlockJacobi(op, keep_f64=False, keep_fp8=bj == "fp8", keep_fp4=bj == "fp4")
    hi = InteriorSystem(op, PREC_F64, None, PREC_F64)
    lo = InteriorSystem(op, PREC_MIXED, pc, PREC_FP8 if bj == "fp8" else PREC_FP4)
    assert pick_basis(lo, False) == "f16" and pick_basis(lo, True) == "f32"
    bb = np.where(mask, 0.0, b)
    rhs_full = np.where(mask, f - A @ bb, 0.0)
    x, info = refine(hi, lo, _dev(rhs_full), rtol=1e-10, inner_rtol=1e-3, basis="f16")
    assert info["converged_reason"] > 0 and info["rel_residual"] <= 1e-10
    u = np.where(mask, x.cpu().numpy(), b)
    assert np.linalg.norm(u - u_ref) / np.linalg.norm(u_ref) < TOL_SOL


@pytest.mark.gpu
def test_refine_fp32_basis_matches_spsolve(cuda):
    """Mixed iterative refinement with an FP32 inner basis reaches the same
    discrete solution as spsolve (FP64 outer residual 1e-10)."""
    from rtx.operator import TransportOperator
    from rtx.krylov import BlockJacobi, InteriorSystem, refine
    from rtx._lib import PREC_F64, PREC_MIXED
    mesh, wl = _c1_small()
    A, f, b, mask = _oracle_system(mesh, wl)
    u_ref, _, _ = osys.exact_solve(A, f, b, mask)
    op = TransportOperator(mesh, wl.kappa, wl.sigma, wl.Phi)
    pc = BlockJacobi(op, keep_f64=False, keep_bf16=True)
    hi = InteriorSystem(op, PREC_F64, None, PREC_F64)
    lo = InteriorSystem(op, PREC_MIXED, pc, PREC_MIXED)
    bb = np.where(mask, 0.0, b)
    rhs_full = np.where(mask, f - A @ bb, 0.0)
    x, info = refine(hi, lo, _dev(rhs_full), rtol=1e-10, inner_rtol=1e-3, basis="f32")
    assert info["converged_reason"] > 0 and info["rel_residual"] <= 1e-10
    u = np.where(mask, x.cpu().numpy(), b)
    assert np.linalg.norm(u - u_ref) / np.linalg.norm(u_ref) < TOL_SOL


@pytest.mark.gpu
def test_inner_solve_shortcuts(cuda):
    """The inner-solve shortcuts of iterative refinement.  With ``x_zero``
    and ``bnorm`` (r = b, no operator apply) GMRES gives the iterate, the
    estimate history and the iteration count of the plain call from x = 0
    bit for bit.  With ``trust_estimate`` as well it stops after the first
    cycle whose estimate meets the target: reason 2, ``res_f`` that estimate,
    no more iterations than the plain call, and the plain call's iterate
    when both stop after the same iteration."""
    import torch
    from rtx.operator import TransportOperator
    from rtx.krylov import BlockJacobi, InteriorSystem, GMRES, _norm
    from rtx._lib import PREC_MIXED, PREC_FP4
    mesh, wl = _c1_small()
    A, f, b, mask = _oracle_system(mesh, wl)
    op = TransportOperator(mesh, wl.kappa, wl.sigma, wl.Phi)
    pc = BlockJacobi(op, keep_f64=False, keep_fp4=True)
    lo = InteriorSystem(op, PREC_MIXED, pc, PREC_FP4)
    bb = np.where(mask, 0.0, b)
    r = _dev(np.where(mask, f - A @ bb, 0.0))
    gm = GMRES(lo, op.ndof, op.device, restart=6, basis="f16")
    rn = _norm(lo.lib, r, gm.work, gm.scal)
    x1, i1 = gm.solve(r, torch.zeros_like(r), rtol=1e-4, max_it=300)
    assert i1["converged_reason"] > 0 and i1["iteration_number"] > 6    # restarts happen
    x2, i2 = gm.solve(r, torch.zeros_like(r), rtol=1e-4, max_it=300, bnorm=rn, x_zero=True)
    assert torch.equal(x1, x2)
    assert i2["convergence_history"] == i1["convergence_history"]
    assert i2["iteration_number"] == i1["iteration_number"]
    x3, i3 = gm.solve(r, torch.zeros_like(r), rtol=1e-4, max_it=300, bnorm=rn, x_zero=True,
                      trust_estimate=True)
    assert i3["converged_reason"] == 2
    assert i3["res_f"] == i3["convergence_history"][-1] <= 1e-4 * rn
    assert i3["iteration_number"] <= i1["iteration_number"]
    if i3["iteration_number"] == i1["iteration_number"]:
        assert torch.equal(x3, x1)


def _bf16_round(a):
    """Round float32 values to BF16 (nearest even), returned as float64."""
    u = np.ascontiguousarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16 << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


@pytest.mark.gpu
@pytest.mark.parametrize("method", [0, 1], ids=["whole-cells", "split"])
def test_block_jacobi_bf16_packed_apply(cuda, method):
    """The packed BF16 block-Jacobi apply equals the oracle's inverses
    rounded to BF16 applied in high precision (to FP32 accumulation), on an
    hp mesh of 80..150-DOF cells (split into angle groups for method 1)."""
    import torch
    from rtx.operator import TransportOperator
    from rtx.krylov import BlockJacobi
    from rtx._lib import PREC_MIXED, check
    mesh = problems.uniform_mesh([1.0, 1.0], [4, 4, 5], 1, 2)
    keys = sorted(mesh.cols)
    mesh.ref_col(keys[3], "spt", "p")
    mesh.ref_cell(keys[-1], sorted(mesh.cols[keys[-1]].cells)[0], "p")
    wl = problems.two_scatterers()
    A, f, b, mask = _oracle_system(mesh, wl)
    op = TransportOperator(mesh, wl.kappa, wl.sigma, wl.Phi)
    check(op.lib.rtx_bjacobi_set_method(op.plan, method), "set method")
    pc = BlockJacobi(op, keep_f32=True, keep_bf16=True)
    split = 128 if method == 1 else None
    invs = osys.block_jacobi_inverses(A, mesh, mask, split)
    r = np.random.default_rng(7).uniform(-1, 1, op.ndof).astype(np.float32)
    z_ref = np.zeros(op.ndof)
    for idx, inv in zip(osys.bj_block_dofs(mesh, split), invs):
        z_ref[idx] = _bf16_round(inv.astype(np.float32)) @ r[idx].astype(np.float64)
    z = pc.apply(_dev(r), torch.zeros(op.ndof, dtype=torch.float32, device="cuda"),
                 PREC_MIXED).cpu().numpy()
    # the device rounds its own FP32 inverse: the two roundings may differ by
    # one BF16 ulp where the FP32 values sit on a rounding boundary
    assert np.linalg.norm(z - z_ref) / np.linalg.norm(z_ref) < 2e-4
    assert pc.bytes_per_apply(PREC_MIXED) >= 2 * pc.entries


@pytest.mark.gpu
@pytest.mark.parametrize("basis", [0, 1], ids=["f64", "f32"])
def test_fused_arnoldi_givens_matches_separate(cuda, basis):
    """rtx_gmres_arnoldi (Arnoldi step with the Givens update folded into the
    normalisation) leaves the basis, the Hessenberg column and the whole
    least-squares state equal (basis bit for bit, state to rounding of the
    rotation recurrence) to rtx_arnoldi_step + rtx_gmres_givens over a full
    cycle of m steps; also on a length not divisible by 4 (the scalar
    path)."""
    import ctypes as C
    import torch
    from rtx import _lib
    from rtx.krylov import _vp
    lib = _lib.load()
    m = 9
    for n in (100_004, 77_777):
        rng = np.random.default_rng(11)
        W = rng.normal(size=(m, n))
        dt = torch.float32 if basis else torch.float64
        ws, ss = C.c_int64(), C.c_int64()
        _lib.check(lib.rtx_arnoldi_workspace(m + 1, C.byref(ws)), "ws")
        _lib.check(lib.rtx_gmres_state_size(m, C.byref(ss)), "ss")
        out = []
        for fused in (False, True):
            V = torch.zeros((m + 1, n), dtype=dt, device="cuda")
            V[0] = torch.from_numpy(W[0] / np.linalg.norm(W[0])).to(dt)
            work = torch.zeros(int(ws.value), dtype=torch.float64, device="cuda")
            state = torch.zeros(int(ss.value), dtype=torch.float64, device="cuda")
            hcol = torch.zeros(m + 2, dtype=torch.float64, device="cuda")
            _lib.check(lib.rtx_gmres_begin(m, 2.5, _vp(state), None), "begin")
            for j in range(m - 1):
                w = torch.from_numpy(W[j + 1]).to(dt).cuda()
                if fused:
                    _lib.check(lib.rtx_gmres_arnoldi(basis, _vp(V), n, n, _vp(w), _vp(hcol),
                                                     _vp(work), 0, j, m, _vp(state), None), "f")
                else:
                    step = lib.rtx_arnoldi_step_f32 if basis else lib.rtx_arnoldi_step
                    _lib.check(step(_vp(V), n, j + 1, n, _vp(w), _vp(hcol), _vp(work), 0, None),
                               "step")
                    _lib.check(lib.rtx_gmres_givens(j, m, _vp(hcol), _vp(state), None), "g")
            torch.cuda.synchronize()
            out.append((V.cpu().numpy(), state.cpu().numpy()))
        assert np.array_equal(out[0][0], out[1][0])
        np.testing.assert_allclose(out[1][1], out[0][1], rtol=1e-13, atol=1e-15)


@pytest.mark.gpu
def test_c4_mixed_refinement_matches_fp64(cuda):
    """The batched uneven-order hp generator scaled to 12,397,421 DOF
    (p_col = 0.4, p_cell = 0.2): mixed refinement (BF16 scattering and
    block-Jacobi, FP32 convection, FP32 inner basis) reaches an FP64
    residual of 1e-8 and the same solution as plain FP64 GMRES(50) to the
    accuracy the residual allows (relative L2 difference below 1e-6)."""
    import torch
    from rtx.operator import TransportOperator, OP_SYSTEM, OP_QUADW
    from rtx.krylov import BlockJacobi
    from rtx import learned
    mesh = problems.random_hp_mesh(0, p_col=0.4, p_cell=0.2)
    wl = problems.two_scatterers_c3()
    op = TransportOperator(mesh, wl.kappa, wl.sigma, wl.Phi)
    assert op.ndof == 12_397_421
    op.autotune(reps=3)
    fd = op.apply(torch.from_numpy(op.nodal(wl.f)).cuda(), ops=OP_QUADW)
    bd = op.mask(torch.from_numpy(op.nodal(wl.bcs)).cuda(), keep_interior=False)
    rhs = op.mask(fd - op.apply(bd, ops=OP_SYSTEM), keep_interior=True)
    bj = BlockJacobi(op, keep_f64=True, keep_bf16=True)
    x64, i64 = learned.Session(op, bj=bj, precision="fp64").solve(rhs, rtol=1e-8)
    xm, im = learned.Session(op, bj=bj, precision="mixed").solve(rhs, rtol=1e-8)
    print("c4 fp64 its", i64["iteration_number"], "mixed its", im["iteration_number"], flush=True)
    assert i64["converged_reason"] > 0 and im["converged_reason"] > 0
    assert im["rel_residual"] <= 1e-8
    assert float((xm - x64).norm() / x64.norm()) < 1e-6


def _e4m3_values():
    """Magnitudes of the finite OCP e4m3 codes 0 .. 0x7E (ascending, 0 .. 448):
    exponent bits 6-3 with bias 7, mantissa bits 2-0, subnormals at exponent 0."""
    vals = []
    for c in range(127):
        e, m = c >> 3, c & 7
        vals.append(m / 512.0 if e == 0 else (1 + m / 8.0) * 2.0 ** (e - 7))
    return np.array(vals)


def _e4m3_code(a):
    """Nearest e4m3 code (ties to the even code), |a| <= 448: sign in bit 7."""
    vals = _e4m3_values()
    x = np.abs(np.asarray(a, dtype=np.float64))
    i = np.clip(np.searchsorted(vals, x), 1, len(vals) - 1)
    lo, hi = vals[i - 1], vals[i]
    pick_hi = (x - lo > hi - x) | ((x - lo == hi - x) & (i % 2 == 0))
    mag = np.where(pick_hi, i, i - 1)
    return (mag | np.where((np.asarray(a) < 0) & (mag > 0), 0x80, 0)).astype(np.uint8)


def _e4m3_decode(codes):
    codes = np.asarray(codes, dtype=np.int64)
    mag = codes & 0x7F
    assert (mag < 127).all(), "NaN code in the pool"
    return np.where(codes & 0x80, -1.0, 1.0) * _e4m3_values()[mag]


def _e4m3_round(a):
    """Nearest FP8 e4m3 value (ties to even), |a| <= 448, as float64."""
    return _e4m3_decode(_e4m3_code(a))


def test_e4m3_helpers():
    """The numpy e4m3 reference: a few code points of the OCP table, ties to
    even, sign, and round trips of every finite code."""
    vals = _e4m3_values()
    assert vals[0x7E] == 448.0 and vals[0x08] == 2.0 ** -6 and vals[0x01] == 2.0 ** -9
    assert vals[0x38] == 1.0 and vals[0x3F] == 1.875
    assert (np.diff(vals) > 0).all()
    codes = np.arange(256, dtype=np.uint8)
    codes = codes[(codes & 0x7F) != 0x7F]
    back = _e4m3_code(_e4m3_decode(codes))
    assert ((back == codes) | ((codes & 0x7F) == 0)).all()       # -0 encodes as +0
    # 1 + 1/16 lies halfway between 1 (code 0x38) and 1.125 (0x39): even wins
    assert _e4m3_code(1.0625) == 0x38 and _e4m3_code(1.1875) == 0x3A
    assert _e4m3_code(-3.0) == 0x80 | 0x44 and _e4m3_round(-3.0) == -3.0


def _device_inverses(pc, blocks, invs):
    """The device's FP32 inverses (column-major per block, blocks back to
    back) as FP64 [row][column] arrays, each checked against the oracle's
    FP64 inverse to the accuracy of an FP32 elimination: relative error
    (Frobenius) at most 8 n u cond(B) with u the FP32 unit roundoff, and at
    most 1e-3 whatever the condition number."""
    inv32 = pc.inv32.cpu().numpy().astype(np.float64)
    out, off = [], 0
    for idx, inv in zip(blocks, invs):
        n = len(idx)
        B = inv32[off:off + n * n].reshape(n, n).T
        off += n * n
        err = np.linalg.norm(B - inv) / np.linalg.norm(inv)
        bound = 8 * n * 2.0 ** -24 * np.linalg.cond(inv)
        assert err <= min(bound, 1e-3), (n, err, bound)
        out.append(B)
    assert off == inv32.size
    return out


@pytest.mark.gpu
def test_block_jacobi_fp8_apply(cuda):
    """FP8 block-Jacobi (e4m3 codes, one FP32 scale max|column| / 448 per
    column) against the oracle's inverses quantised the same way in numpy
    and applied in FP64 (to FP32 accumulation), on an hp mesh with split
    cells; the device's FP32-eliminated inverses against the oracle's, and
    the pool against those inverses quantised in numpy; and a mixed
    refinement solve with it reaches 1e-8."""
    import torch
    from rtx.operator import TransportOperator, OP_SYSTEM, OP_QUADW
    from rtx.krylov import BlockJacobi
    from rtx._lib import PREC_FP8, check
    from rtx import learned
    from test_gpu_apply import small_random_hp
    mesh = small_random_hp(2)
    wl = problems.two_scatterers_c3()
    A, f, b, mask = _oracle_system(mesh, wl)
    op = TransportOperator(mesh, wl.kappa, wl.sigma, wl.Phi)
    pc = BlockJacobi(op, keep_f64=False, keep_f32=True, keep_fp8=True)
    invs = osys.block_jacobi_inverses(A, mesh, mask, split=128)
    dinvs = _device_inverses(pc, osys.bj_block_dofs(mesh, 128), invs)
    r = np.random.default_rng(8).uniform(-1, 1, op.ndof).astype(np.float32)
    z_ref = np.zeros(op.ndof)
    for idx, inv in zip(osys.bj_block_dofs(mesh, 128), invs):
        B = inv.astype(np.float32).astype(np.float64)
        sc = np.abs(B).max(axis=0) / 448.0
        sc = np.where(sc > 0, sc, 1.0).astype(np.float32).astype(np.float64)
        q = _e4m3_round(B / sc[None, :])
        z_ref[idx] = (q * sc[None, :]) @ r[idx].astype(np.float64)
    z = pc.apply(_dev(r), torch.zeros(op.ndof, dtype=torch.float32, device="cuda"),
                 PREC_FP8).cpu().numpy()
    # codes on a rounding boundary may differ by one step between the FP32
    # device scaling and the numpy FP64 one
    assert np.linalg.norm(z - z_ref) / np.linalg.norm(z_ref) < 5e-3
    # the pool itself: codes of block b at off8[b] + j ldp (column j, rows
    # padded to ldp = 16 ceil(n / 16) with code 0), then the FP32 scales
    # after the codes rounded up to 16 bytes, at soff[b] + j
    pool = pc.inv8.cpu().numpy()
    blocks = osys.bj_block_dofs(mesh, 128)
    ns = np.array([len(i) for i in blocks])
    ldps = (ns + 15) // 16 * 16
    off8 = np.concatenate([[0], np.cumsum(ns * ldps)])
    soff = np.concatenate([[0], np.cumsum(ns)])
    codes8 = (off8[-1] + 15) // 16 * 16
    assert pool.size == codes8 + 4 * soff[-1]
    scales = pool[codes8:].view(np.float32).astype(np.float64)
    z_pool = np.zeros(op.ndof)
    n_diff = n_all = 0
    for b, (idx, B) in enumerate(zip(blocks, dinvs)):
        n, ldp = ns[b], ldps[b]
        cd = pool[off8[b]:off8[b + 1]].reshape(n, ldp)           # [column][row]
        assert not cd[:, n:].any(), "padding rows must hold code 0"
        sc = scales[soff[b]:soff[b + 1]]
        sc_ref = np.abs(B).max(axis=0) / 448.0
        sc_ref = np.where(sc_ref > 0, sc_ref, 1.0)
        assert np.allclose(sc, sc_ref, rtol=1e-6, atol=0)
        c_dev = cd[:, :n].T.astype(np.int64)                       # [row][column]
        c_ref = _e4m3_code(np.clip(B / sc[None, :], -448.0, 448.0)).astype(np.int64)
        # signed code index: the codes ordered along the real line
        signed = lambda c: np.where(c & 0x80, -1, 1) * (c & 0x7F)
        step = np.abs(signed(c_dev) - signed(c_ref))
        # a code may differ from the numpy one only by one step, where the
        # device's FP32 quotient sits on a rounding boundary
        assert (step <= 1).all()
        n_diff += int((step != 0).sum())
        n_all += step.size
        z_pool[idx] = (_e4m3_decode(c_dev) * sc[None, :]) @ r[idx].astype(np.float64)
    assert n_diff <= max(2, n_all // 1000), (n_diff, n_all)
    # the apply against the device's own pool: FP32 accumulation only
    assert np.linalg.norm(z - z_pool) / np.linalg.norm(z_pool) < 2e-6
    fd = op.apply(torch.from_numpy(op.nodal(wl.f)).cuda(), ops=OP_QUADW)
    bd = op.mask(torch.from_numpy(op.nodal(wl.bcs)).cuda(), keep_interior=False)
    rhs = op.mask(fd - op.apply(bd, ops=OP_SYSTEM), keep_interior=True)
    x, info = learned.Session(op, bj=pc, precision="mixed", bj_precision=PREC_FP8).solve(
        rhs, rtol=1e-8)
    assert info["converged_reason"] > 0 and info["rel_residual"] <= 1e-8


def _e2m1_values():
    """Magnitudes of the OCP FP4 e2m1 codes 0 .. 7 (exponent bits 2-1 with
    bias 1, mantissa bit 0, a subnormal at exponent 0)."""
    return np.array([(c & 1) / 2.0 if c < 2 else (1 + (c & 1) / 2.0) * 2.0 ** ((c >> 1) - 1)
                     for c in range(8)])


def _e2m1_code(a):
    """Nearest e2m1 code (ties to the even code), saturating at 6: sign in bit 3."""
    vals = _e2m1_values()
    x = np.minimum(np.abs(np.asarray(a, dtype=np.float64)), 6.0)
    i = np.clip(np.searchsorted(vals, x), 1, len(vals) - 1)
    lo, hi = vals[i - 1], vals[i]
    pick_hi = (x - lo > hi - x) | ((x - lo == hi - x) & (i % 2 == 0))
    mag = np.where(pick_hi, i, i - 1)
    return (mag | np.where((np.asarray(a) < 0) & (mag > 0), 8, 0)).astype(np.uint8)


def _e2m1_decode(codes):
    codes = np.asarray(codes, dtype=np.int64)
    return np.where(codes & 8, -1.0, 1.0) * _e2m1_values()[codes & 7]


def test_e2m1_helpers():
    """The numpy e2m1 reference: the OCP table, ties to the even code,
    saturation, sign and round trips."""
    assert list(_e2m1_values()) == [0.0, 0.5, 1.0, 1.5, 2.0, 3.0, 4.0, 6.0]
    codes = np.arange(16, dtype=np.uint8)
    back = _e2m1_code(_e2m1_decode(codes))
    assert ((back == codes) | (codes == 8)).all()                 # -0 encodes as +0
    # midpoints: 0.25 -> 0, 0.75 -> 1.0, 1.25 -> 1.0, 1.75 -> 2, 2.5 -> 2, 3.5 -> 4, 5 -> 4
    mids = [0.25, 0.75, 1.25, 1.75, 2.5, 3.5, 5.0]
    assert list(_e2m1_decode(_e2m1_code(mids))) == [0.0, 1.0, 1.0, 2.0, 2.0, 4.0, 4.0]
    assert _e2m1_decode(_e2m1_code(7.9)) == 6.0 and _e2m1_decode(_e2m1_code(-2.9)) == -3.0


@pytest.mark.gpu
def test_block_jacobi_fp4_apply(cuda):
    """FP4 block-Jacobi (e2m1 codes, two per byte with the even row in the
    low nibble, one FP32 scale max|column| / 6 per column): the device's
    FP32 inverses against the oracle's, the pool against the device
    inverses quantised the same way in numpy, the apply
    against that pool decoded in numpy (FP32 accumulation only), the apply
    from an FP16 r against the FP32 one bit for bit, on an hp mesh with
    split cells and on whole 125-DOF cells; and a mixed refinement solve with
    it reaches 1e-8.  The first two cases eliminate in FP32 (register
    build); the third keeps cells of up to 150 DOFs whole (the LDS build,
    in FP64 then rounded), so the apply runs the slab kernel."""
    import torch
    from rtx.operator import TransportOperator, OP_SYSTEM, OP_QUADW
    from rtx.krylov import BlockJacobi
    from rtx._lib import PREC_FP4, check
    from rtx import learned
    from test_gpu_apply import small_random_hp
    big = problems.uniform_mesh([1.0, 1.0], [4, 4, 5], 1, 2)
    keys = sorted(big.cols)
    big.ref_col(keys[3], "spt", "p")
    big.ref_cell(keys[-1], sorted(big.cols[keys[-1]].cells)[0], "p")
    cases = ((small_random_hp(2), problems.two_scatterers_c3(), 128, -1),
             (problems.uniform_mesh([1.0, 1.0], [5, 5, 5], 2, 2), problems.absorber_c2(), None,
              -1),
             (big, problems.two_scatterers(), None, 0))
    for mesh, wl, split, method in cases:
        A, f, b, mask = _oracle_system(mesh, wl)
        op = TransportOperator(mesh, wl.kappa, wl.sigma, wl.Phi)
        check(op.lib.rtx_bjacobi_set_method(op.plan, method), "set method")
        pc = BlockJacobi(op, keep_f64=False, keep_f32=True, keep_fp4=True)
        blocks = osys.bj_block_dofs(mesh, split)
        invs = osys.block_jacobi_inverses(A, mesh, mask, split=split)
        dinvs = _device_inverses(pc, blocks, invs)
        r = np.random.default_rng(8).uniform(-1, 1, op.ndof).astype(np.float32)
        z = pc.apply(_dev(r), torch.zeros(op.ndof, dtype=torch.float32, device="cuda"),
                     PREC_FP4).cpu().numpy()
        r16 = _dev(r).to(torch.float16)
        z_h = pc.apply_h(r16, torch.zeros(op.ndof, dtype=torch.float32, device="cuda"), PREC_FP4)
        z_f = pc.apply(r16.float(), torch.zeros(op.ndof, dtype=torch.float32, device="cuda"),
                       PREC_FP4)
        assert torch.equal(z_h, z_f)
        # the pool: block b's codes at off8[b] / 2 + j ldp / 2 (column j,
        # ldp = 16 ceil(n / 16) rows, padding code 0), then the FP32 scales
        # after the codes rounded up to 16 bytes, at soff[b] + j
        pool = pc.inv4.cpu().numpy()
        ns = np.array([len(i) for i in blocks])
        ldps = (ns + 15) // 16 * 16
        off4 = np.concatenate([[0], np.cumsum(ns * ldps // 2)])
        soff = np.concatenate([[0], np.cumsum(ns)])
        codes4 = (off4[-1] + 15) // 16 * 16
        assert pool.size == codes4 + 4 * soff[-1]
        scales = pool[codes4:].view(np.float32).astype(np.float64)
        z_pool = np.zeros(op.ndof)
        n_diff = n_all = 0
        for bk, (idx, B) in enumerate(zip(blocks, dinvs)):
            n, ldp = ns[bk], ldps[bk]
            by = pool[off4[bk]:off4[bk + 1]].reshape(n, ldp // 2).astype(np.int64)
            cd = np.stack([by & 15, by >> 4], axis=2).reshape(n, ldp)   # [column][row]
            assert not cd[:, n:].any(), "padding rows must hold code 0"
            sc = scales[soff[bk]:soff[bk + 1]]
            sc_ref = np.abs(B).max(axis=0) / 6.0
            sc_ref = np.where(sc_ref > 0, sc_ref, 1.0)
            assert np.allclose(sc, sc_ref, rtol=1e-6, atol=0)
            c_dev = cd[:, :n].T                                         # [row][column]
            c_ref = _e2m1_code(B / sc[None, :]).astype(np.int64)
            signed = lambda c: np.where(c & 8, -1, 1) * (c & 7)
            step = np.abs(signed(c_dev) - signed(c_ref))
            # one step at most, where the device's FP32 quotient sits on a
            # rounding boundary
            assert (step <= 1).all()
            n_diff += int((step != 0).sum())
            n_all += step.size
            z_pool[idx] = (_e2m1_decode(c_dev) * sc[None, :]) @ r[idx].astype(np.float64)
        assert n_diff <= max(2, n_all // 1000), (n_diff, n_all)
        assert np.linalg.norm(z - z_pool) / np.linalg.norm(z_pool) < 2e-6
        fd = op.apply(torch.from_numpy(op.nodal(wl.f)).cuda(), ops=OP_QUADW)
        bd = op.mask(torch.from_numpy(op.nodal(wl.bcs)).cuda(), keep_interior=False)
        rhs = op.mask(fd - op.apply(bd, ops=OP_SYSTEM), keep_interior=True)
        x, info = learned.Session(op, bj=pc, precision="mixed").solve(rhs, rtol=1e-8)
        assert info["converged_reason"] > 0 and info["rel_residual"] <= 1e-8
