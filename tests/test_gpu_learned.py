"""Learned coarse correction: kernels against the oracle, and what the
trained network gains over block-Jacobi GMRES.

No reference counterpart (the reference's solve has only PETSc bjacobi /
kaczmarz, ``src/rt/solve.py:86-95``); the oracle is ``oracle/system.py``
(numpy MLP forward / backward, literal-loop coarse transfers).
Tolerances: FP32 GEMMs and MLP passes 1e-5 relative to the FP64 numpy
result (max-norm over the output); BF16 GEMV 1e-5 against numpy with the
same BF16-rounded weights; coarse transfers 1e-6 (FP32 outputs of FP64
sums).
"""
import numpy as np
import pytest

from oracle import system as osys, assemble as asm
from rtx import problems

TOL32 = 1e-5


def _t(a, dtype=None):
    import torch
    t = torch.from_numpy(np.ascontiguousarray(a))
    if dtype is not None:
        t = t.to(dtype)
    return t.cuda()


def _relmax(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-30))


def _bf16_round(a):
    u = np.ascontiguousarray(a, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32)


@pytest.mark.gpu
@pytest.mark.parametrize("M,N,K", [(64, 128, 5000), (128, 3000, 64), (64, 4096, 128),
                                   (1, 128, 700), (37, 53, 129), (200, 70, 9)])
@pytest.mark.parametrize("layout", ["NT", "TN", "NN"])
def test_gemm_layouts_and_epilogues(cuda, M, N, K, layout):
    import torch
    from rtx import _lib
    from rtx.learned import _vp
    lib = _lib.load()
    rng = np.random.default_rng(M * 7 + N + K)
    Am = rng.standard_normal((M, K)).astype(np.float32)
    Bm = rng.standard_normal((K, N)).astype(np.float32)
    bias = rng.standard_normal(N).astype(np.float32)
    Z = rng.standard_normal((M, N)).astype(np.float32)
    # store A and B in the layout under test and pass strides
    if layout == "NT":      # A row-major, B given as B^T row-major (N x K)
        A_st, sam, sak = Am, K, 1
        B_st, sbk, sbn = np.ascontiguousarray(Bm.T), 1, K
    elif layout == "TN":    # A given as A^T row-major (K x M), B row-major
        A_st, sam, sak = np.ascontiguousarray(Am.T), 1, M
        B_st, sbk, sbn = Bm, N, 1
    else:
        A_st, sam, sak = Am, K, 1
        B_st, sbk, sbn = Bm, N, 1
    dA, dB = _t(A_st), _t(B_st)
    ref = Am.astype(np.float64) @ Bm.astype(np.float64)
    w = __import__("ctypes").c_int64()
    _lib.check(lib.rtx_mlp_gemm_workspace(M, N, K, __import__("ctypes").byref(w)), "ws")
    work = torch.empty(max(1, int(w.value)), dtype=torch.float32, device="cuda")
    for epi in (0, 1, 2):
        C = torch.empty((M, N), dtype=torch.float32, device="cuda")
        H = torch.empty_like(C)
        dbias, dZ = _t(bias), _t(Z)
        _lib.check(lib.rtx_mlp_gemm(M, N, K, _vp(dA), sam, sak, _vp(dB), sbk, sbn, _vp(C), N, epi,
                                    _vp(dbias) if epi == 1 else None,
                                    _vp(H) if epi == 1 else None,
                                    _vp(dZ) if epi == 2 else None, N if epi == 2 else 0,
                                    _vp(work), work.numel(), None), "gemm")
        out = C.cpu().numpy().astype(np.float64)
        exp = ref + bias if epi == 1 else (ref * (Z > 0) if epi == 2 else ref)
        assert _relmax(out, exp) < TOL32, (layout, epi, _relmax(out, exp))
        if epi == 1:
            assert _relmax(H.cpu().numpy(), np.maximum(exp, 0)) < TOL32


def _mlp(dims, seed=0):
    from rtx.learned import MLP
    return MLP(dims, "cuda", seed=seed)


@pytest.mark.gpu
@pytest.mark.parametrize("dims,B", [([40, 32, 32, 32, 24], 16), ([3000, 128, 128, 128, 300], 8)])
def test_mlp_forward_backward_match_oracle(cuda, dims, B):
    mlp = _mlp(dims, seed=3)
    rng = np.random.default_rng(5)
    # non-zero biases so the bias paths are exercised
    import torch
    for l in range(mlp.L):
        mlp.b(l).copy_(torch.from_numpy(rng.normal(0, 0.1, dims[l + 1]).astype(np.float32)))
    X = rng.standard_normal((B, dims[0])).astype(np.float32)
    dout = rng.standard_normal((B, dims[-1])).astype(np.float32)
    out = mlp.forward(_t(X)).cpu().numpy()
    params = mlp.params_numpy()
    ref, cache = osys.mlp_forward(params, X.astype(np.float64))
    assert _relmax(out, ref) < TOL32
    g = mlp.backward(_t(dout))
    grads = osys.mlp_backward(params, cache, dout.astype(np.float64))
    gh = g.cpu().numpy().astype(np.float64)
    for l in range(mlp.L):
        w0, b0 = mlp.offs[l]
        N, K = dims[l + 1], dims[l]
        assert _relmax(gh[w0:w0 + N * K].reshape(N, K), grads[l][0]) < TOL32, ("W", l)
        assert _relmax(gh[b0:b0 + N], grads[l][1]) < TOL32, ("b", l)


@pytest.mark.gpu
def test_adam_step_matches_formula(cuda):
    mlp = _mlp([10, 8, 6])
    rng = np.random.default_rng(1)
    p0 = mlp.p.cpu().numpy().astype(np.float64)
    m = np.zeros_like(p0)
    v = np.zeros_like(p0)
    lr, b1, b2, eps, wd = 1e-2, 0.9, 0.999, 1e-8, 1e-3
    for step in range(1, 4):
        g = rng.standard_normal(p0.size).astype(np.float32)
        mlp.g.copy_(_t(g))
        mlp.adam(lr=lr, beta1=b1, beta2=b2, eps=eps, weight_decay=wd)
        m = b1 * m + (1 - b1) * g
        v = b2 * v + (1 - b2) * g.astype(np.float64) ** 2
        p0 = p0 - lr * ((m / (1 - b1 ** step)) / (np.sqrt(v / (1 - b2 ** step)) + eps) + wd * p0)
    assert _relmax(mlp.p.cpu().numpy(), p0) < 1e-5


@pytest.mark.gpu
@pytest.mark.parametrize("N,K,relu", [(128, 5000, 1), (300, 128, 0), (7, 3, 1)])
def test_gemv_bf16(cuda, N, K, relu):
    import torch
    from rtx import _lib
    from rtx.learned import _vp
    lib = _lib.load()
    rng = np.random.default_rng(N + K)
    W = rng.standard_normal((N, K + 5)).astype(np.float32)    # row stride K + 5
    x = rng.standard_normal(K).astype(np.float32)
    b = rng.standard_normal(N).astype(np.float32)
    Wd = _t(W.reshape(-1))
    Wb = torch.empty(W.size, dtype=torch.int16, device="cuda")
    _lib.check(lib.rtx_mlp_to_bf16(W.size, _vp(Wd), _vp(Wb), None), "to_bf16")
    y = torch.empty(N, dtype=torch.float32, device="cuda")
    dx, db = _t(x), _t(b)       # keep the device buffers alive across the call
    _lib.check(lib.rtx_mlp_gemv_bf16(N, K, _vp(Wb), K + 5, _vp(dx), _vp(db), relu,
                                     _vp(y), None), "gemv")
    Wr = _bf16_round(W)[:, :K].astype(np.float64)
    ref = Wr @ x.astype(np.float64) + b
    if relu:
        ref = np.maximum(ref, 0)
    assert _relmax(y.cpu().numpy(), ref) < TOL32
    # the BF16 conversion itself: round to nearest even
    back = (Wb.cpu().numpy().astype(np.int32).astype(np.uint32) << 16).view(np.float32)
    assert np.array_equal(back.reshape(W.shape), _bf16_round(W))


def hp_mesh():
    m = problems.uniform_mesh([1.0, 1.0], [3, 3, 3], 2, 1)
    keys = sorted(m.cols)
    m.ref_col(keys[0], "spt", "hp")
    m.ref_col(sorted(m.cols)[0], "spt", "h")
    last = sorted(m.cols)[-1]
    m.ref_cell(last, sorted(m.cols[last].cells)[0], "hp")
    return m


@pytest.mark.gpu
@pytest.mark.parametrize("n", [2, 4, 8])
def test_coarse_transfers_match_oracle(cuda, n):
    """Columns coarser than, equal to and finer than the coarse cells."""
    import torch
    from rtx.operator import TransportOperator
    from rtx.learned import CoarseGrid
    mesh = hp_mesh()
    wl = problems.two_scatterers()
    op = TransportOperator(mesh, wl.kappa, wl.sigma, wl.Phi)
    grid = CoarseGrid(op, n)
    rng = np.random.default_rng(n)
    v = rng.uniform(-1, 1, op.ndof)
    for dtype in (torch.float64, torch.float32):
        vd = _t(v, dtype)
        m0 = torch.empty(3 * n * n, dtype=torch.float32, device="cuda")
        grid.restrict(vd, m0, mode=0)
        vv = vd.double().cpu().numpy()
        assert _relmax(m0.cpu().numpy(), osys.coarse_moments(mesh, n, vv)) < 1e-6
        a = torch.empty(n * n, dtype=torch.float32, device="cuda")
        grid.restrict(vd, a, mode=1)
        assert _relmax(a.cpu().numpy(), osys.coarse_average(mesh, n, vv)) < 1e-6
        phi = rng.standard_normal(n * n).astype(np.float32)
        scale = torch.tensor([1.7], dtype=torch.float32, device="cuda")
        z = torch.zeros(op.ndof, dtype=dtype, device="cuda")
        grid.prolong(_t(phi), scale, z)
        ref = osys.coarse_prolong(mesh, n, phi.astype(np.float64), 1.7, asm.intr_mask(mesh))
        assert _relmax(z.double().cpu().numpy(), ref) < 1e-6


@pytest.mark.gpu
def test_line_search_and_device_axpy(cuda):
    import torch
    from rtx import _lib
    from rtx.learned import _vp
    lib = _lib.load()
    rng = np.random.default_rng(2)
    n = 100003
    r, q = rng.standard_normal(n), rng.standard_normal(n)
    q += 0.3 * r
    w = __import__("ctypes").c_int64()
    _lib.check(lib.rtx_line_search_workspace(__import__("ctypes").byref(w)), "ws")
    work = torch.empty(int(w.value), dtype=torch.float64, device="cuda")
    alpha = torch.zeros(4, dtype=torch.float64, device="cuda")
    dr, dq = _t(r), _t(q)
    _lib.check(lib.rtx_line_search(0, n, _vp(dr), _vp(dq), _vp(alpha), _vp(work), None), "ls")
    a = alpha.cpu().numpy()
    assert abs(a[0] - (r @ q) / (q @ q)) < 1e-12 * abs(a[0])
    assert abs(a[3] - r @ r) < 1e-12 * (r @ r)
    y = dr.clone()
    _lib.check(lib.rtx_axpy_dev(0, n, _vp(alpha), -1.0, _vp(dq), _vp(y), None), "axpy")
    assert _relmax(y.cpu().numpy(), r - a[0] * q) < 1e-14
    # a direction that increases the residual gets alpha = 0
    dmq = _t(-q)
    _lib.check(lib.rtx_line_search(0, n, _vp(dr), _vp(dmq), _vp(alpha), _vp(work), None), "ls")
    assert alpha[0].item() == 0.0


# ------------------------------------------------------------ gain test
def _diffusive(seed, n):
    return problems.grf_scatterer(seed, n=n, kappa_scale=30.0, ratio_range=(0.99, 0.999), g=0.3)


@pytest.fixture(scope="module")
def trained():
    """16 x 16 columns x 8 cells, p = 2 (55,296 DOF); coarse grid = the
    columns; network 1280 -> 128 -> 128 -> 128 -> 256 trained on pairs
    from 48 random fields (restart residuals and Arnoldi directions)."""
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from rtx.operator import TransportOperator
    from rtx import learned
    mesh = problems.uniform_mesh([1.0, 1.0], [3, 3, 3], 3, 4)
    wl0 = _diffusive(0, 16)
    op = TransportOperator(mesh, wl0.kappa, wl0.sigma, wl0.Phi)
    grid = learned.CoarseGrid(op, 16)
    print("generating pairs", flush=True)
    X, Y = learned.generate_pairs(op, grid, range(48), cycles=26, restart=8, skip=2,
                                  field=lambda s: _diffusive(s, 16), direction_stride=1)
    nc = grid.nc
    mlp = learned.MLP([5 * nc, 128, 128, 128, nc], "cuda", seed=0)
    mlp.n_moments = 3 * nc
    print("training on", X.shape[0], "pairs", flush=True)
    xs, ys, hist = learned.train(mlp, X, Y, epochs=150, batch=64, lr=1e-3, log=print)
    return mesh, op, grid, mlp, xs, ys, hist


def _rhs(op, wl):
    import torch
    from rtx.operator import OP_QUADW
    f = op.apply(torch.from_numpy(op.nodal(wl.f)).cuda(), ops=OP_QUADW)
    return op.mask(f, keep_interior=True)


@pytest.mark.gpu
def test_learned_preconditioner_gain(cuda, trained):
    """On an unseen field the two-level preconditioner needs at least 10 %
    fewer GMRES iterations than block-Jacobi alone, and both reach the same
    solution (FP64 true residual 1e-8).  (The exact Galerkin coarse solve on
    this problem, 215 -> 116 iterations, bounds what any coarse scalar-flux
    correction of this form can gain; the network reaches about 170.)"""
    from rtx import learned
    from rtx.krylov import BlockJacobi
    mesh, op, grid, mlp, xs, ys, hist = trained
    assert hist[-1][1] is not None and hist[-1][1] < hist[0][1]      # validation error fell
    wl = _diffusive(10_000, 16)
    op.set_coeffs(wl.kappa, wl.sigma)
    rhs = _rhs(op, wl)
    bj = BlockJacobi(op, keep_f32=True, keep_f64=True, keep_bf16=True)
    x_bj, info_bj = learned.solve(op, rhs, None, pc=None, bj=bj, precision="fp64", rtol=1e-8)
    model = learned.CoarseModel(mlp, 3 * grid.nc, xs, ys)
    model.set_coefficients(grid.coefficient_channels(wl.kappa, wl.sigma))
    pc = learned.TwoLevelPC(op, bj, grid, model)
    x_l, info_l = learned.solve(op, rhs, None, pc=pc, bj=bj, precision="fp64", rtol=1e-8,
                                judge=False)
    print("block-Jacobi: {} its; learned: {} its (fallback {})".format(
        info_bj["iteration_number"], info_l["iteration_number"], info_l["fallback"]))
    assert info_bj["converged_reason"] > 0 and info_l["converged_reason"] > 0
    assert info_l["rel_residual"] <= 1e-8
    assert float((x_l - x_bj).norm() / x_bj.norm()) < 1e-6
    assert not info_l["fallback"]
    assert info_l["iteration_number"] < 0.9 * info_bj["iteration_number"]
    # mixed precision with the learned correction reaches the FP64 target too
    pc.enabled = True
    x_m, info_m = learned.solve(op, rhs, None, pc=pc, bj=bj, precision="mixed", rtol=1e-8,
                                judge=False)
    assert info_m["rel_residual"] <= 1e-8
    assert float((x_m - x_bj).norm() / x_bj.norm()) < 1e-6


@pytest.mark.gpu
def test_untrained_network_falls_back_safely(cuda, trained):
    """A network that was never trained must not break the solve: the line
    search keeps each coarse step from increasing the residual and the FP64
    judge switches the correction off when it stops helping."""
    from rtx import learned
    from rtx.krylov import BlockJacobi
    mesh, op, grid, _, _, _, _ = trained
    wl = _diffusive(10_001, 16)
    op.set_coeffs(wl.kappa, wl.sigma)
    rhs = _rhs(op, wl)
    bj = BlockJacobi(op, keep_f32=True, keep_f64=True, keep_bf16=True)
    raw = learned.MLP([5 * grid.nc, 128, 128, 128, grid.nc], "cuda", seed=7)
    model = learned.CoarseModel(raw, 3 * grid.nc, 1.0, -50.0)     # wrong sign and scale
    model.set_coefficients(grid.coefficient_channels(wl.kappa, wl.sigma))
    pc = learned.TwoLevelPC(op, bj, grid, model)
    x, info = learned.solve(op, rhs, None, pc=pc, bj=bj, precision="mixed", rtol=1e-8)
    assert info["converged_reason"] > 0 and info["rel_residual"] <= 1e-8
    assert info["fallback"]
    x_bj, _ = learned.solve(op, rhs, None, pc=None, bj=bj, precision="fp64", rtol=1e-10)
    assert float((x - x_bj).norm() / x_bj.norm()) < 1e-6


@pytest.mark.gpu
def test_model_with_large_validation_error_is_gated(cuda, trained):
    """A coarse model whose held-out error exceeds the gate is not used: the
    session solves with block-Jacobi alone (no trial step) and says so; a
    model under the gate goes to the FP64 judge as before."""
    from rtx import learned
    from rtx.krylov import BlockJacobi
    mesh, op, grid, _, _, _, _ = trained
    wl = _diffusive(10_002, 16)
    op.set_coeffs(wl.kappa, wl.sigma)
    rhs = _rhs(op, wl)
    bj = BlockJacobi(op, keep_f64=True, keep_bf16=True)
    raw = learned.MLP([5 * grid.nc, 128, 128, 128, grid.nc], "cuda", seed=7)
    model = learned.CoarseModel(raw, 3 * grid.nc, 1.0, 1.0)
    model.set_coefficients(grid.coefficient_channels(wl.kappa, wl.sigma))
    model.validation_error = 0.9
    pc = learned.TwoLevelPC(op, bj, grid, model)
    sess = learned.Session(op, pc=pc, bj=bj, precision="mixed")
    x, info = sess.solve(rhs, rtol=1e-8)
    assert info["gated"] and info["fallback"]
    assert info["converged_reason"] > 0 and info["rel_residual"] <= 1e-8
    assert pc.calls == 0
    ref = learned.Session(op, bj=bj, precision="mixed").solve(rhs, rtol=1e-8)[0]
    assert float((x - ref).norm() / ref.norm()) < 1e-7
    model.validation_error = 0.1
    x, info = learned.Session(op, pc=pc, bj=bj, precision="mixed").solve(rhs, rtol=1e-8)
    assert not info["gated"] and info["converged_reason"] > 0
