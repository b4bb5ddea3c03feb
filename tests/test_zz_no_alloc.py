"""rtx_gmres_arnoldi_static (no workspace argument, nothing allocated) against
rtx_gmres_arnoldi with a caller-allocated workspace, through the C ABI.

Both run the same kernels; the basis column, the Hessenberg column and the
whole least-squares state must agree bit for bit.  Two calls in a row, the
second on a smaller cycle and a shorter basis (the static array keeps the
first call's contents), in FP64 and FP32 bases, with the default one-pass
Gram-Schmidt and with the device-decided second pass.  The limit on the
cycle length is rejected as an ordinary argument error.
"""
import ctypes as C

import numpy as np
import pytest


def _cycle(lib, vp, fn, basis, m, n, W, refine, work=None):
    import torch
    from rtx import _lib
    dt = torch.float32 if basis else torch.float64
    ss = C.c_int64()
    _lib.check(lib.rtx_gmres_state_size(m, C.byref(ss)), "state size")
    V = torch.zeros((m + 1, n), dtype=dt, device="cuda")
    V[0] = torch.from_numpy(W[0] / np.linalg.norm(W[0])).to(dt)
    state = torch.zeros(int(ss.value), dtype=torch.float64, device="cuda")
    hcol = torch.zeros(m + 2, dtype=torch.float64, device="cuda")
    _lib.check(lib.rtx_gmres_begin(m, 1.75, vp(state), None), "begin")
    hcols = []
    for j in range(m):
        w = torch.from_numpy(W[j + 1]).to(dt).cuda()
        if work is None:
            rc = fn(basis, vp(V), n, n, vp(w), vp(hcol), refine, j, m, vp(state), None)
        else:
            rc = fn(basis, vp(V), n, n, vp(w), vp(hcol), vp(work), refine, j, m, vp(state),
                    None)
        _lib.check(rc, "arnoldi step")
        hcols.append(hcol[:j + 2].cpu().numpy().copy())
    torch.cuda.synchronize()
    return V.cpu().numpy(), state.cpu().numpy(), hcols


@pytest.mark.gpu
@pytest.mark.parametrize("refine", [0, 1])
@pytest.mark.parametrize("basis", [0, 1], ids=["f64", "f32"])
def test_static_arnoldi_matches_workspace_arnoldi(cuda, basis, refine):
    import torch
    from rtx import _lib
    from rtx.krylov import _vp
    lib = _lib.load()
    ws = C.c_int64()
    _lib.check(lib.rtx_arnoldi_workspace(200, C.byref(ws)), "workspace")
    # first call: a long cycle on a long basis; second: shorter on both
    for m, n, seed in ((12, 300_004, 5), (5, 77_777, 6)):
        print("static arnoldi: basis {} refine {} m {} n {}".format(basis, refine, m, n),
              flush=True)
        W = np.random.default_rng(seed).normal(size=(m + 1, n))
        work = torch.zeros(int(ws.value), dtype=torch.float64, device="cuda")
        ref = _cycle(lib, _vp, lib.rtx_gmres_arnoldi, basis, m, n, W, refine, work=work)
        got = _cycle(lib, _vp, lib.rtx_gmres_arnoldi_static, basis, m, n, W, refine)
        assert np.array_equal(got[0], ref[0])
        assert np.array_equal(got[1], ref[1])
        for a, b in zip(got[2], ref[2]):
            assert np.array_equal(a, b)
        # the cycle is a real Arnoldi process: orthonormal columns
        Vd = got[0].astype(np.float64)
        G = Vd @ Vd.T
        assert np.abs(G - np.eye(m + 1)).max() < (1e-5 if basis else 1e-12)


@pytest.mark.gpu
def test_static_arnoldi_rejects_cycles_beyond_its_limit(cuda):
    import torch
    from rtx import _lib
    from rtx.krylov import _vp
    lib = _lib.load()
    n = 1024
    V = torch.zeros((2, n), dtype=torch.float64, device="cuda")
    w = torch.zeros(n, dtype=torch.float64, device="cuda")
    hcol = torch.zeros(4, dtype=torch.float64, device="cuda")
    state = torch.zeros(64, dtype=torch.float64, device="cuda")
    rc = lib.rtx_gmres_arnoldi_static(0, _vp(V), n, n, _vp(w), _vp(hcol), 0, 0, 300,
                                      _vp(state), None)
    assert rc == -1
    assert b"static workspace" in lib.rtx_last_error()
