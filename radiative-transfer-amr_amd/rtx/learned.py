"""Learned coarse scalar-flux correction: a two-level preconditioner.

No counterpart in the reference (its solve offers PETSc bjacobi / kaczmarz,
``src/rt/solve.py:86-95``).  Scattering-dominated problems leave GMRES with
per-cell block-Jacobi stuck on smooth, nearly isotropic error modes; a coarse
correction of the scalar flux removes them.  Here that coarse solve is a
small MLP:

* restriction: per coarse cell, the angular moments (int r, int r cos th,
  int r sin th) of the weak-form residual summed over the columns covering
  it (``rtx_coarse_restrict`` mode 0), normalised by their 2-norm s;
* network: 4 dense layers (ReLU between), input [moments / s, coefficient
  channels log kappa and log(kappa - sigma) on the coarse grid], output an
  isotropic correction per coarse cell; trained in FP32 (MFMA GEMMs, AdamW),
  inferred with BF16 weights (GEMV kernels);
* prolongation: the correction times s on every interior DOF of each column;
* two-level step: z_c = P(N(R r)), q = A z_c, alpha >= 0 minimising the
  zeroth coarse moments of r - alpha q on the device (alpha = 1 for an exact
  Galerkin coarse solve), then z = alpha z_c + BJ(r - alpha q).  The network
  is nonlinear, so the outer Krylov method is flexible GMRES.

Training pairs come from the solver itself: GMRES + block-Jacobi on
A x = 0 from a random start, for random kappa / sigma fields.  There the
error of every iterate is exactly -x and its residual -A x, so each restart
yields a (residual, correction) pair without a reference solution.

Fallback: ``Session.solve`` judges outer steps by the FP64 true residual:
one step with the correction and one without, keeping whichever reduced the
residual faster per second; afterwards a step with the correction that does
not cut the residual by 10 % switches it off, and the solve continues as
plain block-Jacobi GMRES.
"""
import ctypes as C
import time

import numpy as np
import torch

from . import _lib, problems
from ._lib import (check, PREC_F64, PREC_F32, PREC_MIXED, PREC_FP8, PREC_FP4, OP_SYSTEM,
                   OP_RESTRICT)
from .operator import stream_handle
from .krylov import (BlockJacobi, InteriorSystem, GMRES, gmres, refine, pick_basis, _norm,
                     INNER_RESTART, INNER_CHUNK)


def _vp(t):
    return C.c_void_p(t.data_ptr()) if t is not None else None


def _ip(a):
    return a.ctypes.data_as(C.c_void_p)


# --------------------------------------------------------------------- grid
class CoarseGrid:
    """n x n coarse grid over the domain; every column maps to the coarse
    cells it overlaps (``rtx_coarse_setup``).  Coarse cell g = gx * n + gy."""

    def __init__(self, op, n=64):
        t = op.t
        self.op, self.n, self.nc = op, int(n), int(n) * int(n)
        self.Lx = float(t.col_geom[:, 2].max())
        self.Ly = float(t.col_geom[:, 3].max())
        hx, hy = self.Lx / n, self.Ly / n
        starts, cells, fcell, fcol = [0], [], [], []
        for c in range(t.ncol):
            x0, y0, x1, y1 = t.col_geom[c]
            area = (x1 - x0) * (y1 - y0)
            gx0, gx1 = int(np.floor(x0 / hx + 1e-9)), int(np.ceil(x1 / hx - 1e-9))
            gy0, gy1 = int(np.floor(y0 / hy + 1e-9)), int(np.ceil(y1 / hy - 1e-9))
            for gx in range(max(gx0, 0), min(max(gx1, gx0 + 1), n)):
                ox = min(x1, (gx + 1) * hx) - max(x0, gx * hx)
                if ox <= 0:
                    continue
                for gy in range(max(gy0, 0), min(max(gy1, gy0 + 1), n)):
                    oy = min(y1, (gy + 1) * hy) - max(y0, gy * hy)
                    if oy <= 0:
                        continue
                    cells.append(gx * n + gy)
                    fcell.append(ox * oy / (hx * hy))
                    fcol.append(ox * oy / area)
            starts.append(len(cells))
        self.starts = np.asarray(starts, np.int32)
        self.cells = np.asarray(cells, np.int32)
        self.fcell = np.asarray(fcell, np.float64)
        self.fcol = np.asarray(fcol, np.float64)
        check(op.lib.rtx_coarse_setup(op.plan, self.nc, self.cells.size, _ip(self.starts),
                                      _ip(self.cells), _ip(self.fcell), _ip(self.fcol)),
              "rtx_coarse_setup")

    def centres(self):
        h = (np.arange(self.n) + 0.5) / self.n
        return np.meshgrid(h * self.Lx, h * self.Ly, indexing="ij")

    def coefficient_channels(self, kappa, sigma):
        """[log kappa, log(kappa - sigma)] at the coarse cell centres, scaled
        so the block has unit 2-norm per unit-variance entry."""
        X, Y = self.centres()
        k = np.asarray(kappa(X, Y), dtype=np.float64) * np.ones_like(X)
        s = np.asarray(sigma(X, Y), dtype=np.float64) * np.ones_like(X)
        a = np.maximum(k - s, 1e-12 * k)
        ch = np.concatenate([np.log(k).ravel(), (np.log(a).ravel() + 4.0) / 2.0])
        return (ch / np.sqrt(ch.size)).astype(np.float32)

    def restrict(self, v, out, mode=0):
        prec = PREC_F64 if v.dtype == torch.float64 else PREC_F32
        check(self.op.lib.rtx_coarse_restrict(self.op.plan, mode, prec, _vp(v), _vp(out),
                                              stream_handle()), "rtx_coarse_restrict")
        return out

    def prolong(self, phi, scale, z, accumulate=False):
        prec = PREC_F64 if z.dtype == torch.float64 else PREC_F32
        check(self.op.lib.rtx_coarse_prolong(self.op.plan, prec, _vp(phi), _vp(scale),
                                             1 if accumulate else 0, _vp(z), stream_handle()),
              "rtx_coarse_prolong")
        return z


# ---------------------------------------------------------------------- MLP
class MLP:
    """Dense ReLU network; parameters packed [W_0 (N_0 x K_0), b_0, W_1, ...]
    in one FP32 device tensor.  Training uses the MFMA GEMM kernels."""

    def __init__(self, dims, device, seed=0):
        self.lib = _lib.load()
        self.dims = [int(d) for d in dims]
        self.L = len(dims) - 1
        self.offs = []
        o = 0
        for l in range(self.L):
            K, N = self.dims[l], self.dims[l + 1]
            self.offs.append((o, o + N * K))
            o += N * K + N
        self.nparam = o
        rng = np.random.default_rng(seed)
        p = np.zeros(o, np.float32)
        for l in range(self.L):
            K, N = self.dims[l], self.dims[l + 1]
            w0, b0 = self.offs[l]
            p[w0:w0 + N * K] = rng.normal(0.0, np.sqrt(2.0 / K), N * K)
        self.device = torch.device(device)
        self.p = torch.from_numpy(p).to(self.device)
        self.g = torch.zeros_like(self.p)
        self.m = torch.zeros_like(self.p)
        self.v = torch.zeros_like(self.p)
        self.step = 0
        self._B = None

    def W(self, l, p=None):
        p = self.p if p is None else p
        w0, b0 = self.offs[l]
        return p[w0:b0].view(self.dims[l + 1], self.dims[l])

    def b(self, l, p=None):
        p = self.p if p is None else p
        w0, b0 = self.offs[l]
        return p[b0:b0 + self.dims[l + 1]]

    def params_numpy(self):
        h = self.p.cpu().numpy().astype(np.float64)
        return [(self.W(l, torch.from_numpy(h)).numpy(), self.b(l, torch.from_numpy(h)).numpy())
                for l in range(self.L)]

    def _alloc(self, B):
        """Activation buffers for batches of up to B rows; self.Z / self.H /
        self.dZ are views of the first B rows of the current batch."""
        dev = self.device
        if self._B is None or B > self._B:
            cap = B
            self._Zb = [torch.empty(cap * self.dims[l + 1], dtype=torch.float32, device=dev)
                        for l in range(self.L)]
            self._Hb = [torch.empty(cap * self.dims[l + 1], dtype=torch.float32, device=dev)
                        for l in range(self.L - 1)]
            mx = max(self.dims[1:])
            self._dZb = [torch.empty(cap * mx, dtype=torch.float32, device=dev) for _ in range(2)]
            need = 0
            w = C.c_int64()
            for b in range(1, cap + 1):          # smaller batches may split K further
                for l in range(self.L):
                    K, N = self.dims[l], self.dims[l + 1]
                    for (M_, N_, K_) in ((b, N, K), (N, K, b), (b, K, N)):
                        check(self.lib.rtx_mlp_gemm_workspace(M_, N_, K_, C.byref(w)),
                              "gemm workspace")
                        need = max(need, int(w.value))
            self.work = torch.empty(max(need, 1), dtype=torch.float32, device=dev)
            if not hasattr(self, "loss"):
                self.loss = torch.zeros(1, dtype=torch.float64, device=dev)
            self._B = cap
        self.Z = [self._Zb[l][:B * self.dims[l + 1]].view(B, self.dims[l + 1])
                  for l in range(self.L)]
        self.H = [self._Hb[l][:B * self.dims[l + 1]].view(B, self.dims[l + 1])
                  for l in range(self.L - 1)]
        mx = max(self.dims[1:])
        self.dZ = [b[:B * mx].view(B, mx) for b in self._dZb]

    def _gemm(self, M, N, K, A, sam, sak, Bm, sbk, sbn, Cm, ldc, epi=0, bias=None, H=None,
              Z=None, ldz=0):
        check(self.lib.rtx_mlp_gemm(M, N, K, _vp(A), sam, sak, _vp(Bm), sbk, sbn, _vp(Cm), ldc,
                                    epi, _vp(bias), _vp(H), _vp(Z), ldz, _vp(self.work),
                                    self.work.numel(), stream_handle()), "rtx_mlp_gemm")

    def forward(self, X):
        """X: (B, dims[0]) FP32 contiguous; returns the (B, dims[-1]) output
        (a view of the last pre-activation buffer)."""
        B = X.shape[0]
        self._alloc(B)
        self.X = X
        h = X
        for l in range(self.L):
            K, N = self.dims[l], self.dims[l + 1]
            Hout = self.H[l] if l < self.L - 1 else None
            self._gemm(B, N, K, h, K, 1, self.W(l), 1, K, self.Z[l], N, epi=1, bias=self.b(l),
                       H=Hout)
            h = Hout
        return self.Z[-1]

    def backward(self, dout):
        """Gradients of sum(dout * out) into self.g (after ``forward``)."""
        B = dout.shape[0]
        dz = dout
        for l in range(self.L - 1, -1, -1):
            K, N = self.dims[l], self.dims[l + 1]
            hin = self.X if l == 0 else self.H[l - 1]
            gW, gb = self.W(l, self.g), self.b(l, self.g)
            # dW = dZ^T H_in: (N x B) (B x K)
            self._gemm(N, K, B, dz, 1, N, hin, K, 1, gW, K)
            check(self.lib.rtx_mlp_colsum(B, N, _vp(dz), _vp(gb), stream_handle()), "colsum")
            if l > 0:
                nxt = self._dZb[l % 2][:B * K].view(B, K)
                # dZ_{l-1} = (dZ W) * (Z_{l-1} > 0): (B x N) (N x K)
                self._gemm(B, K, N, dz, N, 1, self.W(l), K, 1, nxt, K, epi=2,
                           Z=self.Z[l - 1], ldz=K)
                dz = nxt
        return self.g

    def adam(self, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8, weight_decay=0.0):
        self.step += 1
        check(self.lib.rtx_adam_step(self.nparam, _vp(self.p), _vp(self.g), _vp(self.m),
                                     _vp(self.v), lr, beta1, beta2, eps, weight_decay, self.step,
                                     stream_handle()), "rtx_adam_step")

    def mse_grad(self, out, Y, dout):
        """dout = 2 (out - Y) / B; accumulates sum (out - Y)^2 into self.loss."""
        B = out.shape[0]
        check(self.lib.rtx_mse_grad(out.numel(), 2.0 / B, _vp(out), _vp(Y), _vp(dout),
                                    _vp(self.loss), stream_handle()), "rtx_mse_grad")
        return dout


class CoarseModel:
    """Inference form of a trained network for one problem: BF16 weights,
    the input and output scalings folded in, and the coefficient channels
    folded into the first layer's bias."""

    def __init__(self, mlp, n_moments, x_scale, y_scale):
        self.lib = mlp.lib
        self.mlp = mlp
        self.nm = int(n_moments)
        dev = mlp.device
        p = mlp.p.clone()
        W0 = mlp.W(0, p)
        W0[:, :self.nm] *= float(x_scale)
        mlp.W(mlp.L - 1, p).mul_(float(y_scale))
        mlp.b(mlp.L - 1, p).mul_(float(y_scale))
        self.p32 = p
        self.pbf = torch.empty(p.numel(), dtype=torch.int16, device=dev)
        check(self.lib.rtx_mlp_to_bf16(p.numel(), _vp(p), _vp(self.pbf), stream_handle()),
              "rtx_mlp_to_bf16")
        self.b0 = torch.empty(mlp.dims[1], dtype=torch.float32, device=dev)
        self.h = [torch.empty(d, dtype=torch.float32, device=dev) for d in mlp.dims[1:]]
        self.coef = None
        # relative error on held-out pairs (set by the trainer's caller);
        # Session gates the correction off when it is too large
        self.validation_error = None

    def set_coefficients(self, channels):
        """b0' = b0 + W0[:, moments:] c for this problem's coefficient channels."""
        mlp = self.mlp
        K0, N0 = mlp.dims[0], mlp.dims[1]
        nc = K0 - self.nm
        assert channels.size == nc
        self.coef = torch.from_numpy(np.ascontiguousarray(channels, np.float32)).to(mlp.device)
        W0 = mlp.W(0, self.p32)
        # (1 x nc) (nc x N0), B[k, n] = W0[n, nm + k]
        w0 = mlp.offs[0][0]
        Bptr = self.p32[w0 + self.nm:]
        check(self.lib.rtx_mlp_gemm(1, N0, nc, _vp(self.coef), nc, 1, _vp(Bptr), 1, K0,
                                    _vp(self.b0), N0, 1, _vp(mlp.b(0, self.p32)), None, None, 0,
                                    None, 0, stream_handle()), "rtx_mlp_gemm (bias fold)")
        del W0

    def infer(self, x, out=None):
        """x: normalised moments (n_moments FP32) -> coarse correction."""
        mlp = self.mlp
        st = stream_handle()
        h = x
        for l in range(mlp.L):
            K, N = mlp.dims[l], mlp.dims[l + 1]
            w0, b0 = mlp.offs[l]
            Wp = self.pbf[w0:]
            bias = self.b0 if l == 0 else self.p32[b0:b0 + N]
            y = self.h[l] if (out is None or l < mlp.L - 1) else out
            check(self.lib.rtx_mlp_gemv_bf16(N, self.nm if l == 0 else K, _vp(Wp), K, _vp(h),
                                             _vp(bias), 1 if l < mlp.L - 1 else 0, _vp(y), st),
                  "rtx_mlp_gemv_bf16")
            h = y
        return h


# ------------------------------------------------------ two-level precondition
class TwoLevelPC:
    """z = alpha z_c + BJ(r - alpha A z_c), z_c = P(N(R r)), alpha >= 0 chosen
    on the device: by default (step "coarse") to minimise the zeroth coarse
    moments of r - alpha A z_c, which is alpha = 1 for an exact Galerkin
    coarse solve and scales an imperfect network's correction in the coarse
    norm; step "full" minimises |r - alpha A z_c| instead (much weaker in
    the multiplicative two-level step)."""

    def __init__(self, op, bj, grid, model, apply_precision=PREC_F64, step="coarse"):
        if step not in ("coarse", "full"):
            raise ValueError("step must be 'coarse' or 'full'")
        self.op, self.bj, self.grid, self.model = op, bj, grid, model
        self.step = step
        self.lib = op.lib
        self.apply_precision = apply_precision
        dev = op.device
        nc = grid.nc
        self.mom = torch.empty(3 * nc, dtype=torch.float32, device=dev)
        self.momq = torch.empty(3 * nc, dtype=torch.float32, device=dev)
        self.x = torch.empty(3 * nc, dtype=torch.float32, device=dev)
        self.scale = torch.empty(1, dtype=torch.float32, device=dev)
        self.phi = torch.empty(nc, dtype=torch.float32, device=dev)
        self.alpha = torch.zeros(4, dtype=torch.float64, device=dev)
        w = C.c_int64()
        check(self.lib.rtx_line_search_workspace(C.byref(w)), "line search workspace")
        self.lswork = torch.empty(int(w.value), dtype=torch.float64, device=dev)
        self._bufs = {}
        self.enabled = True
        self.calls = 0

    def _buf(self, dtype):
        if dtype not in self._bufs:
            n = self.op.ndof
            self._bufs[dtype] = [torch.empty(n, dtype=dtype, device=self.op.device)
                                 for _ in range(3)]
        return self._bufs[dtype]

    def coarse_only(self, r, zc):
        """z_c = P(N(R r)) (no line search)."""
        self.grid.restrict(r, self.mom, mode=0)
        check(self.lib.rtx_scale_features(self.mom.numel(), _vp(self.mom), _vp(self.x),
                                          _vp(self.scale), stream_handle()), "scale features")
        self.model.infer(self.x, self.phi)
        self.grid.prolong(self.phi, self.scale, zc)
        return zc

    def apply(self, r, z, precision=PREC_F64, stream=None):
        if not self.enabled:
            return self.bj.apply(r, z, precision)
        self.calls += 1
        vprec = PREC_F64 if r.dtype == torch.float64 else PREC_F32
        zc, q, rr = self._buf(r.dtype)
        self.coarse_only(r, zc)
        if vprec == PREC_F64:
            aprec = PREC_F64
        else:
            aprec = self.apply_precision if self.apply_precision != PREC_F64 else PREC_MIXED
        self.op.apply(zc, q, ops=OP_SYSTEM | OP_RESTRICT, precision=aprec)
        st = stream_handle()
        if self.step == "coarse":
            # alpha minimises |R_0 (r - alpha q)| over the zeroth moments (the
            # Galerkin condition R_0 (r - A z_c) = 0 gives alpha = 1)
            self.grid.restrict(q, self.momq, mode=0)
            check(self.lib.rtx_coarse_line_search(self.grid.nc, _vp(self.mom), _vp(self.momq),
                                                  _vp(self.alpha), st), "rtx_coarse_line_search")
        else:
            check(self.lib.rtx_line_search(vprec, r.numel(), _vp(r), _vp(q), _vp(self.alpha),
                                           _vp(self.lswork), st), "rtx_line_search")
        rr.copy_(r)
        check(self.lib.rtx_axpy_dev(vprec, r.numel(), _vp(self.alpha), -1.0, _vp(q), _vp(rr), st),
              "axpy_dev")
        self.bj.apply(rr, z, precision)
        check(self.lib.rtx_axpy_dev(vprec, r.numel(), _vp(self.alpha), 1.0, _vp(zc), _vp(z), st),
              "axpy_dev")
        return z

    def last_removed_fraction(self):
        """Fraction of |r|^2 the last coarse step removed ((r.q)^2 / (q.q r.r))."""
        a = self.alpha.cpu().numpy()
        if a[2] <= 0 or a[3] <= 0 or a[0] <= 0:
            return 0.0
        return float(a[1] * a[1] / (a[2] * a[3]))


# ------------------------------------------------------------ training data
def generate_pairs(op, grid, seeds, cycles=26, restart=8, skip=2, field=None, device=None,
                   log=None, direction_stride=2):
    """Solver-generated (residual, correction) pairs.  For each seed: set the
    field's coefficients, factor block-Jacobi, run GMRES(restart) on A x = 0
    from a random interior start, and record exact pairs of two kinds:
    * at every restart after the first ``skip``: the residual r = -A x and
      the error e = -x of the iterate;
    * at every ``direction_stride``-th Arnoldi step: the preconditioned
      direction z and w = A z (residual w, correction z) - inputs like the
      ones the preconditioner sees inside GMRES.
    Each pair is stored as the residual's moments (R r) and the coarse mean
    of the correction.  Returns X (pairs x (3 nc + 2 nc)), Y (pairs x nc)
    FP32 device tensors, with X's moment block and Y divided by |R r|."""
    field = field or (lambda s: problems.grf_scatterer(s, n=grid.n))
    dev = op.device
    nc = grid.nc
    Xs, Ys = [], []
    mom = torch.empty(3 * nc, dtype=torch.float32, device=dev)
    tgt = torch.empty(nc, dtype=torch.float32, device=dev)
    for s in seeds:
        t0 = time.perf_counter()
        wl = field(int(s))
        op.set_coeffs(wl.kappa, wl.sigma)
        chans = torch.from_numpy(grid.coefficient_channels(wl.kappa, wl.sigma)).to(dev)
        bj = BlockJacobi(op, keep_f32=False, keep_f64=True)
        system = InteriorSystem(op, PREC_F64, bj, PREC_F64)
        gen = torch.Generator(device=dev)
        gen.manual_seed(int(s) + 12345)
        x0 = torch.rand(op.ndof, dtype=torch.float64, device=dev, generator=gen) * 2.0 - 1.0
        op.mask(x0)
        b = torch.zeros_like(x0)
        count = [0]

        def add(res, corr, sign):
            grid.restrict(res, mom, mode=0)
            grid.restrict(corr, tgt, mode=1)
            sc = torch.linalg.vector_norm(mom).clamp_min(1e-30)
            Xs.append(torch.cat([mom / sc, chans]))
            Ys.append(sign * tgt / sc)

        def hook(x, r):
            count[0] += 1
            if count[0] > skip:
                add(r, x, -1.0)            # r = -A x, e = -x

        step = [0]

        def dhook(z, w):
            step[0] += 1
            if direction_stride and step[0] % direction_stride == 0:
                add(w, z, 1.0)             # w = A z

        gmres(system, b, x0, rtol=0.0, atol=0.0, restart=restart, max_it=restart * cycles,
              on_restart=hook, on_direction=dhook if direction_stride else None)
        del bj, system
        if log:
            log("pairs: seed {} ratio {:.4f} -> {} pairs ({:.2f} s)".format(
                s, wl.ratio, len(Xs), time.perf_counter() - t0))
    return torch.stack(Xs).contiguous(), torch.stack(Ys).contiguous()


def train(mlp, X, Y, epochs=40, batch=64, lr=1e-3, weight_decay=1e-5, seed=0, val_frac=0.1,
          x_scale=None, log=None):
    """AdamW on mean squared error with the device kernels.  The moment block
    of X is multiplied by x_scale (default sqrt(n_moments)) and Y divided by
    its RMS, so both are O(1); the returned (x_scale, y_scale) undo that at
    inference.  Returns (x_scale, y_scale, history)."""
    n = X.shape[0]
    if X.shape[1] != mlp.dims[0] or Y.shape[1] != mlp.dims[-1]:
        raise ValueError("pairs of width {} -> {} do not fit a {} network".format(
            X.shape[1], Y.shape[1], mlp.dims))
    nmom = getattr(mlp, "n_moments", None)
    if nmom is None:
        raise ValueError("set mlp.n_moments (size of the moment block of X) before training")
    dev = mlp.device
    n_val = max(1, int(round(n * val_frac))) if n > 10 else 0
    g = torch.Generator()
    g.manual_seed(seed)
    perm = torch.randperm(n, generator=g).to(dev)
    val_idx, tr_idx = perm[:n_val], perm[n_val:]
    if tr_idx.numel() < batch:
        raise ValueError("{} training pairs are fewer than one batch of {}".format(
            tr_idx.numel(), batch))
    y_scale = float(torch.sqrt((Y.index_select(0, tr_idx) ** 2).mean()).item()) or 1.0
    xs = float(np.sqrt(nmom)) if x_scale is None else float(x_scale)
    colscale = torch.ones(X.shape[1], dtype=torch.float32, device=dev)
    colscale[:nmom] = xs
    g = torch.Generator()
    g.manual_seed(seed + 1)
    hist = []
    ntr = tr_idx.numel()
    dout = torch.empty(batch, mlp.dims[-1], dtype=torch.float32, device=dev)
    mlp._alloc(batch)
    for ep in range(epochs):
        order = tr_idx[torch.randperm(ntr, generator=g).to(dev)]
        mlp.loss.zero_()
        nb = 0
        for i in range(0, ntr - batch + 1, batch):
            idx = order[i:i + batch]
            Xb = (X.index_select(0, idx) * colscale).contiguous()
            Yb = (Y.index_select(0, idx) / y_scale).contiguous()
            out = mlp.forward(Xb)
            mlp.mse_grad(out, Yb, dout)
            mlp.backward(dout)
            mlp.adam(lr=lr, weight_decay=weight_decay)
            nb += 1
        tr_loss = float(mlp.loss.item()) / max(1, nb * batch)
        mlp.loss.zero_()
        va = validation_error(mlp, X, Y, val_idx, colscale, y_scale) if val_idx.numel() else None
        hist.append((tr_loss, va))
        if log and (ep % 10 == 0 or ep == epochs - 1):
            log("epoch {:3d}: train mse {:.4e}, validation rel. error {}".format(
                ep, tr_loss, "n/a" if va is None else "{:.4f}".format(va)))
    return xs, y_scale, hist


def validation_error(mlp, X, Y, idx, colscale, y_scale, batch=64):
    """Relative 2-norm error of the network on the pairs ``idx``."""
    num, den = 0.0, 0.0
    for i in range(0, idx.numel(), batch):
        j = idx[i:i + batch]
        Xb = (X.index_select(0, j) * colscale).contiguous()
        out = mlp.forward(Xb) * y_scale
        Yb = Y.index_select(0, j)
        num += float(((out - Yb) ** 2).sum().item())
        den += float((Yb ** 2).sum().item())
    return (num / den) ** 0.5 if den > 0 else None


# ------------------------------------------------------------------- solve
class Session:
    """A reusable interior-system solver for one operator: the systems, the
    GMRES objects and their captured graphs persist across ``solve`` calls
    (for example the repeated solves of an adaptive loop).

    precision "mixed": FP64 iterative refinement around flexible GMRES with
    FP32 / BF16 applies and block-Jacobi inverses stored in ``bj_precision``
    (PREC_FP4: e2m1 with column scales, the default; PREC_FP8: e4m3 with
    column scales; PREC_MIXED: BF16; PREC_F32: FP32) on an FP16, FP32 or
    FP64 Arnoldi basis (``basis`` "auto": FP16 with FP4 or FP8 inverses and
    no coarse correction, else FP32);
    "fp64": GMRES in FP64.  ``pc`` is a TwoLevelPC or
    None (plain block-Jacobi).  A coarse model whose held-out relative error
    (``model.validation_error``) exceeds ``gate`` is not used at all: the
    session runs plain block-Jacobi and reports ``gated``."""

    def __init__(self, op, pc=None, bj=None, precision="mixed", restart=None,
                 bj_precision=None, basis="auto", gate=0.5):
        self.gated = False
        if pc is not None:
            ve = getattr(getattr(pc, "model", None), "validation_error", None)
            if ve is not None and gate is not None and ve > gate:
                self.gated = True
                bj = pc.bj if bj is None else bj
                pc = None
        if precision not in ("mixed", "fp64"):
            raise ValueError("precision must be 'mixed' or 'fp64'")
        # mixed: FP4 inverses by default (on the uniform absorber, the
        # two-scatterer hp mesh and the random field: 1.5 %, 2.5 % and 12 %
        # more inner iterations than FP8, each at about 60 % of the FP8
        # apply's time: faster solves); a given block-Jacobi object is used
        # with the lowest-precision copy it holds
        if bj_precision is None and precision != "fp64":
            if bj is None or getattr(bj, "inv4", None) is not None:
                bj_precision = PREC_FP4
            elif getattr(bj, "inv8", None) is not None:
                bj_precision = PREC_FP8
            elif getattr(bj, "inv16", None) is not None:
                bj_precision = PREC_MIXED
            else:
                bj_precision = PREC_F32
        if bj is None:
            bj = BlockJacobi(op, keep_f32=precision != "fp64" and bj_precision == PREC_F32,
                             keep_bf16=precision != "fp64" and bj_precision == PREC_MIXED,
                             keep_fp8=precision != "fp64" and bj_precision == PREC_FP8,
                             keep_fp4=precision != "fp64" and bj_precision == PREC_FP4,
                             keep_f64=precision == "fp64")
        self.op, self.pc, self.bj, self.precision = op, pc, bj, precision
        the_pc = pc if pc is not None else bj
        flexible = pc is not None
        if precision == "fp64":
            # GMRES(50) as the reference's KSP (src/rt/solve.py:87)
            restart = 50 if restart is None else restart
            self.system = InteriorSystem(op, PREC_F64, the_pc, PREC_F64)
            self.gm = GMRES(self.system, op.ndof, op.device, restart=restart, flexible=flexible)
        else:
            self.hi = InteriorSystem(op, PREC_F64, None, PREC_F64)
            self.lo = InteriorSystem(op, PREC_MIXED, the_pc, bj_precision)
            if basis == "auto":
                basis = pick_basis(self.lo, flexible)
            # inner solves: GMRES(INNER_RESTART) on the FP16 basis, else
            # GMRES(20) (the flexible method of the learned correction)
            if restart is None:
                restart = INNER_RESTART if basis == "f16" else 20
            chunk = INNER_CHUNK if basis == "f16" else 10
            self.gm = GMRES(self.lo, op.ndof, op.device, restart=restart, flexible=flexible,
                            basis=basis, chunk=chunk)
        self.restart = restart

    def solve(self, rhs, x0=None, rtol=1e-8, max_it=5000, inner_rtol=1e-2, max_outer=60,
              judge=True):
        """Solve to an FP64 relative residual ``rtol``.  With ``judge`` the
        FP64 true residual decides whether the coarse correction stays on
        (``_Judge``: one trial step with and one without, then the faster
        per second; later, a step with the correction that does not cut the
        residual by 10 % switches it off)."""
        pc = self.pc
        if pc is not None:
            pc.enabled = True
        jd = _Judge(pc if judge else None)
        if self.precision == "fp64":
            x, info = self.gm.solve(rhs, x0, rtol=rtol, max_it=max_it,
                                    on_restart=jd.restart_hook(self.system, rhs))
        else:
            x, info = refine(self.hi, self.lo, rhs, x0, rtol=rtol, inner_rtol=inner_rtol,
                             restart=self.restart, max_outer=max_outer, max_inner=max_it,
                             flexible=pc is not None, after_outer=jd.outer_hook, inner=self.gm)
        info["fallback"] = jd.fell_back or self.gated
        info["gated"] = self.gated
        return x, info


def solve(op, rhs, x0=None, pc=None, bj=None, precision="mixed", rtol=1e-8, restart=None,
          max_it=5000, inner_rtol=1e-2, max_outer=60, bj_precision=None, judge=True,
          gate=0.5):
    """One solve with a fresh ``Session`` (see there)."""
    return Session(op, pc=pc, bj=bj, precision=precision, restart=restart,
                   bj_precision=bj_precision, gate=gate).solve(rhs, x0, rtol=rtol, max_it=max_it,
                                                               inner_rtol=inner_rtol,
                                                               max_outer=max_outer, judge=judge)


class _Judge:
    """The FP64 true residual decides whether the coarse correction stays on.

    Over the first two judged steps (outer refinement steps, or restart
    cycles in FP64) it runs one step with the correction and one without and
    keeps whichever reduced the FP64 residual faster per second; afterwards a
    step with the correction that does not cut the residual by 10 % switches
    it off for good."""

    def __init__(self, pc):
        self.pc = pc
        self.prev = None
        self.t_prev = None
        self.rates = []
        self.fell_back = False

    def _judge(self, rn):
        now = time.perf_counter()
        pc = self.pc
        if pc is not None and self.prev is not None and np.isfinite(rn) and rn > 0:
            dt = max(now - self.t_prev, 1e-9)
            rate = np.log(self.prev / rn) / dt
            if len(self.rates) == 0:
                self.rates.append(("learned", rate))
                pc.enabled = False                # trial step without the correction
            elif len(self.rates) == 1:
                self.rates.append(("plain", rate))
                pc.enabled = self.rates[0][1] > rate
                self.fell_back = not pc.enabled
            elif pc.enabled and not (rn < 0.9 * self.prev):
                pc.enabled = False
                self.fell_back = True
        self.prev, self.t_prev = rn, now

    def outer_hook(self, outer, rn):
        self._judge(rn)

    def restart_hook(self, system, b):
        def hook(x, r):
            self._judge(float(torch.linalg.vector_norm(r).item()))
        return hook
