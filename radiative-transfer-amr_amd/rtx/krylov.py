"""Device-resident GMRES, per-cell block-Jacobi and FP64 iterative refinement.

Counterpart of the PETSc KSP solve of ``src/rt/solve.py:81-117`` (GMRES with
restart 50 and a block-Jacobi preconditioner).  Differences that are design
choices, not accidents:

* right preconditioning, so the GMRES residual is the true residual of the
  interior system (PETSc's default is left preconditioning);
* the Krylov basis, the Arnoldi orthogonalisation (classical Gram-Schmidt,
  one pass as PETSc's default, optionally re-orthogonalised), the Givens QR
  of the Hessenberg matrix with its back substitution, the solution update
  and the true residual all stay on the device (FP64, or an FP32 or FP16
  basis for the inner solves of mixed refinement); the host only reads the residual
  estimates once per chunk of iterations;
* convergence is checked on the FP64 true residual at every restart; when
  the recurrence estimate has met the target but the true residual sits
  at its attainable floor (it did not halve over the last cycle and is
  below 1e-10 |b|), the solve stops as converged, as PETSc's
  estimate-based test does with the reference's rtol = atol = 1e-15;
* the block-Jacobi blocks are the angular cells (``BlockJacobi``);
* ``refine`` wraps an inexact inner GMRES that applies the operator and the
  preconditioner in FP32 / BF16 inside an FP64 residual loop.

Vectors are full-length (every DOF); inflow-boundary DOFs are kept at zero,
which makes the iteration identical to one on the compressed interior
system A_II of ``split_matrix`` (``src/dg/matrix/split_matrix.py:46-47``).
"""
import ctypes as C
import time

import numpy as np
import torch

from . import _lib
from ._lib import (check, PREC_F64, PREC_F32, PREC_MIXED, PREC_FP8, PREC_FP4, OP_SYSTEM,
                   OP_RESTRICT)
from .operator import stream_handle


# A true residual below ATTAINABLE |b| that stops decreasing counts as the
# rounding floor of the FP64 apply (about eps |A| |x|).
ATTAINABLE = 1e-10

# The inner solves of mixed refinement on an FP16 basis: GMRES(15), the host
# reading the residual estimates every 10 iterations.  Measured to 1e-8
# (warm solves, FP4 block-Jacobi) against GMRES(20): the uniform absorber of
# 2,048,000 DOF 0.0156 against 0.0189 s (176 against 206 inner iterations);
# the two-scatterer hp mesh of 8,263,762 DOF 0.0829 against 0.0998 s (173
# against 206); the random-field scatterer of 8,192,000 DOF 0.1228 against
# 0.1233 s at kappa scale 1, but 0.1189 against 0.1090 s (423 against 360)
# at kappa scale 30.  Longer cycles needed more iterations on the FP16 basis
# on the first two meshes; GMRES(10) stagnated on the random field.
INNER_RESTART, INNER_CHUNK = 15, 10


def _vp(t):
    return C.c_void_p(t.data_ptr())


class BlockJacobi:
    """Dense inverses of every cell's interior diagonal block, on the device."""

    def __init__(self, op, restrict=True, keep_f32=False, keep_f64=True, keep_bf16=False,
                 keep_fp8=False, keep_fp4=False):
        self.op, self.lib = op, op.lib
        n = C.c_int64()
        check(self.lib.rtx_bjacobi_size(op.plan, C.byref(n)), "rtx_bjacobi_size")
        self.entries = int(n.value)
        dev = op.device
        # BF16 / FP8 / FP4 are rounded from FP32
        need32 = keep_f32 or keep_bf16 or keep_fp8 or keep_fp4
        self.inv32 = torch.empty(self.entries, dtype=torch.float32, device=dev) if need32 else None
        st = stream_handle()
        rc = -1
        if need32 and not keep_f64:
            # no FP64 copy wanted: eliminate in FP32 straight into inv32
            # (register build, blocks of up to 128 DOFs; else the FP64 path)
            self.inv = None
            rc = self.lib.rtx_bjacobi_factor(op.plan, 1 if restrict else 0, None,
                                             _vp(self.inv32), None, st)
            if rc != 0 and not (rc == -1 and b"FP32-only" in self.lib.rtx_last_error()):
                check(rc, "rtx_bjacobi_factor")
        if rc != 0:
            self.inv = torch.empty(self.entries, dtype=torch.float64, device=dev)
            rc = self.lib.rtx_bjacobi_factor(op.plan, 1 if restrict else 0, _vp(self.inv),
                                             _vp(self.inv32) if need32 else None, None, st)
        if rc == -1 and b"LDS" in self.lib.rtx_last_error():
            scratch = torch.empty_like(self.inv)
            rc = self.lib.rtx_bjacobi_factor(op.plan, 1 if restrict else 0, _vp(self.inv),
                                             _vp(self.inv32) if need32 else None,
                                             _vp(scratch), st)
            del scratch
        check(rc, "rtx_bjacobi_factor")
        self.inv16 = None
        self.entries16 = 0
        if keep_bf16:
            n16 = C.c_int64()
            check(self.lib.rtx_bjacobi_size_bf16(op.plan, C.byref(n16)), "rtx_bjacobi_size_bf16")
            self.entries16 = int(n16.value)
            # packed: columns padded to a multiple of 8 rows for 16-byte loads
            self.inv16 = torch.empty(self.entries16, dtype=torch.int16, device=dev)
            check(self.lib.rtx_bjacobi_pack_bf16(op.plan, _vp(self.inv32), _vp(self.inv16), st),
                  "rtx_bjacobi_pack_bf16")
        self.inv8 = None
        self.bytes8 = 0
        if keep_fp8:
            n8 = C.c_int64()
            check(self.lib.rtx_bjacobi_size_fp8(op.plan, C.byref(n8)), "rtx_bjacobi_size_fp8")
            self.bytes8 = int(n8.value)
            # e4m3 codes (16-byte columns) followed by one FP32 scale per column
            self.inv8 = torch.empty(self.bytes8, dtype=torch.uint8, device=dev)
            check(self.lib.rtx_bjacobi_pack_fp8(op.plan, _vp(self.inv32), _vp(self.inv8), st),
                  "rtx_bjacobi_pack_fp8")
        self.inv4 = None
        self.bytes4 = 0
        if keep_fp4:
            n4 = C.c_int64()
            check(self.lib.rtx_bjacobi_size_fp4(op.plan, C.byref(n4)), "rtx_bjacobi_size_fp4")
            self.bytes4 = int(n4.value)
            # e2m1 codes (two per byte, 8-byte columns) followed by one FP32
            # scale per column
            self.inv4 = torch.empty(self.bytes4, dtype=torch.uint8, device=dev)
            check(self.lib.rtx_bjacobi_pack_fp4(op.plan, _vp(self.inv32), _vp(self.inv4), st),
                  "rtx_bjacobi_pack_fp4")
        if keep_bf16 or keep_fp8 or keep_fp4:
            if not keep_f32:
                torch.cuda.current_stream().synchronize()
                self.inv32 = None
        if not keep_f64:
            self.inv = None

    def bytes_per_apply(self, precision=PREC_F64):
        if precision == PREC_MIXED:
            return self.entries16 * 2
        if precision == PREC_FP8:
            return self.bytes8
        if precision == PREC_FP4:
            return self.bytes4
        return self.entries * {PREC_F64: 8, PREC_F32: 4}[precision]

    def apply_h(self, r, z, precision=PREC_FP8, stream=None):
        """z = M r with the FP8 (PREC_FP8) or FP4 (PREC_FP4) inverses, r in
        FP16 and z in FP32 (the preconditioner of an FP16 GMRES basis)."""
        inv = {PREC_FP8: self.inv8, PREC_FP4: self.inv4}.get(precision)
        if inv is None:
            raise ValueError("no FP8 / FP4 inverses for precision {} (keep_fp8 / keep_fp4)"
                             .format(precision))
        check(self.lib.rtx_bjacobi_apply_h(self.op.plan, precision, _vp(inv), _vp(r), _vp(z),
                                           stream_handle(stream)), "rtx_bjacobi_apply_h")
        return z

    def apply_fp8_h(self, r, z, stream=None):
        """apply_h with the FP8 inverses."""
        return self.apply_h(r, z, PREC_FP8, stream)

    def apply(self, r, z, precision=PREC_F64, stream=None):
        """precision PREC_F64: FP64 inverses and vectors; PREC_F32: FP32;
        PREC_MIXED: BF16 inverses, FP32 vectors; PREC_FP8: FP8 (e4m3)
        inverses with per-column scales, FP32 vectors; PREC_FP4: FP4 (e2m1)
        inverses with per-column scales, FP32 vectors."""
        inv = {PREC_F64: self.inv, PREC_F32: self.inv32, PREC_MIXED: self.inv16,
               PREC_FP8: self.inv8, PREC_FP4: self.inv4}[precision]
        if inv is None:
            raise ValueError("block-Jacobi inverses for this precision were not kept")
        check(self.lib.rtx_bjacobi_apply(self.op.plan, precision, _vp(inv), _vp(r), _vp(z),
                                         stream_handle(stream)), "rtx_bjacobi_apply")
        return z


class InteriorSystem:
    """y = A_II x in full-length FP64 vectors, the apply done in ``precision``."""

    def __init__(self, op, precision=PREC_F64, pc=None, pc_precision=None):
        self.op, self.lib, self.precision = op, op.lib, precision
        self.pc = pc
        self.pc_precision = precision if pc_precision is None else pc_precision
        n = op.ndof
        dev = op.device
        if precision != PREC_F64 or self.pc_precision != PREC_F64:
            self.x32 = torch.empty(n, dtype=torch.float32, device=dev)
            self.y32 = torch.empty(n, dtype=torch.float32, device=dev)

    def graph_key(self):
        """Identifies the preconditioner configuration captured in graphs."""
        return (id(self.pc), getattr(self.pc, "enabled", True), self.precision,
                self.pc_precision)

    def matvec(self, x, y):
        if self.precision == PREC_F64:
            self.op.apply(x, y, ops=OP_SYSTEM | OP_RESTRICT, precision=PREC_F64)
        else:
            st = stream_handle()
            check(self.lib.rtx_convert(0, x.numel(), _vp(x), _vp(self.x32), st), "convert")
            self.op.apply(self.x32, self.y32, ops=OP_SYSTEM | OP_RESTRICT, precision=self.precision)
            check(self.lib.rtx_convert(1, x.numel(), _vp(self.y32), _vp(y), st), "convert")
        return y

    def precondition(self, r, z):
        if self.pc is None:
            z.copy_(r)
            return z
        if self.pc_precision == PREC_F64:
            return self.pc.apply(r, z, PREC_F64)
        st = stream_handle()
        check(self.lib.rtx_convert(0, r.numel(), _vp(r), _vp(self.x32), st), "convert")
        self.pc.apply(self.x32, self.y32, self.pc_precision)
        check(self.lib.rtx_convert(1, r.numel(), _vp(self.y32), _vp(z), st), "convert")
        return z

    def matvec_lo(self, x, y):
        """y = A_II x on FP32 vectors, the apply in this system's reduced
        precision (FP32 basis of the inner solves)."""
        if self.precision == PREC_F64:
            raise ValueError("an FP32 basis needs a reduced-precision system")
        self.op.apply(x, y, ops=OP_SYSTEM | OP_RESTRICT, precision=self.precision)
        return y

    def precondition_lo(self, r, z):
        """z = M r on FP32 vectors (inverses in the preconditioner precision)."""
        if self.pc is None:
            z.copy_(r)
            return z
        if self.pc_precision == PREC_F64:
            raise ValueError("an FP32 basis needs a reduced-precision preconditioner")
        return self.pc.apply(r, z, self.pc_precision)

    def precondition_h(self, r, z):
        """z = M r from an FP16 r into an FP32 z (FP16 basis of the inner
        solves); the preconditioner must be the FP8 or FP4 block-Jacobi."""
        if self.pc is None:
            z.copy_(r)
            return z
        if not self.fp16_basis_ok():
            raise ValueError("an FP16 basis needs the FP8 or FP4 block-Jacobi preconditioner")
        return self.pc.apply_h(r, z, self.pc_precision)

    def fp16_basis_ok(self):
        """Whether this reduced-precision system can serve an FP16 basis."""
        return (self.precision != PREC_F64 and self.pc is not None
                and self.pc_precision in (PREC_FP8, PREC_FP4) and hasattr(self.pc, "apply_h"))

    def residual(self, b, x, r):
        """r = b - A_II x (always FP64)."""
        self.op.apply(x, r, ops=OP_SYSTEM | OP_RESTRICT, precision=PREC_F64)
        check(self.lib.rtx_axpby(PREC_F64, b.numel(), 1.0, _vp(b), -1.0, _vp(r), stream_handle()),
              "axpby")
        return r


class GMRESResult(dict):
    pass


def _norm(lib, v, work, out):
    check(lib.rtx_mdot(_vp(v), v.numel(), 1, v.numel(), _vp(v), _vp(out), _vp(work),
                       stream_handle()), "rtx_mdot")
    return float(out[0].item()) ** 0.5


class _ChunkGraphs:
    """HIP graphs of chunks of GMRES iterations, keyed by (j0, j1).  A chunk
    runs eagerly the first time it is needed (that also loads every kernel
    it uses and makes every lazily allocated buffer) and is captured the
    next time through the library's own capture calls (rtx_graph_*), on a
    side stream ordered after the current one; replays run on the current
    stream and reuse the same device buffers, so a graph stays valid for the
    whole solve."""

    def __init__(self, lib):
        self.lib = lib
        self.seen = set()
        self.graphs = {}
        self.stream = None

    def run(self, key, body, enabled):
        if not enabled:
            body()
            return
        g = self.graphs.get(key)
        if g is None:
            if key not in self.seen:
                self.seen.add(key)
                body()
                return
            cur = torch.cuda.current_stream()
            if self.stream is None:
                self.stream = torch.cuda.Stream(device=cur.device)
            self.stream.wait_stream(cur)
            h = C.c_void_p()
            with torch.cuda.stream(self.stream):
                st = stream_handle(self.stream)
                check(self.lib.rtx_graph_capture_begin(st), "rtx_graph_capture_begin")
                try:
                    body()
                finally:
                    rc = self.lib.rtx_graph_capture_end(st, C.byref(h))
                check(rc, "rtx_graph_capture_end")
            cur.wait_stream(self.stream)
            g = self.graphs[key] = h
        check(self.lib.rtx_graph_launch(g, stream_handle()), "rtx_graph_launch")

    def __del__(self):
        try:
            for g in self.graphs.values():
                self.lib.rtx_graph_destroy(g)
            self.graphs = {}
        except Exception:
            pass


class GMRES:
    """Right-preconditioned restarted GMRES on ``system`` (InteriorSystem)
    with persistent device buffers, so repeated solves with the same system
    (the inner solves of iterative refinement) reuse the captured graphs.

    The Arnoldi basis, the orthogonalisation, the Givens QR of the
    Hessenberg matrix and the small triangular solve all stay on the device;
    the host reads the residual estimates once per ``chunk`` iterations, and
    (``graphs``) each chunk of iterations - preconditioner, operator, Arnoldi
    step, Givens update - is replayed as one HIP graph.  Graphs are keyed by
    the chunk and by ``system.graph_key()``, so a change of preconditioner
    configuration never replays a stale graph.  ``refine`` selects the
    Gram-Schmidt refinement of ``rtx_arnoldi_step`` (0: none, PETSc's
    default; 1: when needed; 2: always).  ``basis`` "f32" keeps the basis
    (V, Z and the work vectors) in FP32 for a reduced-precision system: the
    operator and preconditioner then act on it directly and the
    orthogonalisation moves half the bytes; dot products, the Hessenberg
    matrix, the iterate and the true residual stay FP64."""

    def __init__(self, system, n, device, restart=50, flexible=False, chunk=10, graphs=True,
                 refine=0, basis="f64"):
        if basis not in ("f64", "f32", "f16"):
            raise ValueError("basis must be 'f64', 'f32' or 'f16'")
        if basis == "f16" and flexible:
            raise ValueError("an FP16 basis is not available for flexible GMRES")
        self.system, self.lib = system, system.lib
        self.n, self.m = int(n), int(restart)
        self.flexible, self.chunk, self.graphs, self.refine = flexible, int(chunk), graphs, refine
        self.basis = basis
        m, dev = self.m, device
        # basis storage; w and z are FP32 for either reduced-precision basis
        vt = {"f64": torch.float64, "f32": torch.float32, "f16": torch.float16}[basis]
        bt = torch.float64 if basis == "f64" else torch.float32
        # basis rows padded to a multiple of 8 entries, so every vector starts
        # on 16 bytes and the Gram-Schmidt kernels take their vector loads
        # for any n (V and Z are views of the first n entries of each row)
        self.ld = ld = (int(n) + 7) // 8 * 8
        self.V = torch.empty((m + 1, ld), dtype=vt, device=dev)[:, :n]
        self.Z = torch.empty((m, ld), dtype=bt, device=dev)[:, :n] if flexible else None
        self.w = torch.empty(n, dtype=bt, device=dev)
        self.z = torch.empty(n, dtype=bt, device=dev)
        self.r = torch.empty(n, dtype=torch.float64, device=dev)
        ws = C.c_int64()
        check(self.lib.rtx_arnoldi_workspace(m + 1, C.byref(ws)), "workspace")
        self.work = torch.zeros(int(ws.value), dtype=torch.float64, device=dev)
        self.hcol = torch.zeros(m + 2, dtype=torch.float64, device=dev)
        check(self.lib.rtx_gmres_state_size(m, C.byref(ws)), "gmres state size")
        self.state = torch.zeros(int(ws.value), dtype=torch.float64, device=dev)
        self.est_off = 3 * m + 1 + m * m + (m + 1) * m
        self.H_off = 3 * m + 1 + m * m
        self.est_h = torch.zeros(m, dtype=torch.float64).pin_memory()
        self.ydev = torch.empty(m, dtype=torch.float64, device=dev)
        self.scal = torch.empty(2, dtype=torch.float64, device=dev)
        self.chunks = _ChunkGraphs(self.lib)
        self._on_direction = None

    def _steps(self, j0, j1):
        lib, st = self.lib, stream_handle()
        V, Z, w, z, n, m = self.V, self.Z, self.w, self.z, self.n, self.m
        lo = self.basis != "f64"
        if self.basis == "f16":
            pre = self.system.precondition_h
        else:
            pre = self.system.precondition_lo if lo else self.system.precondition
        mv = self.system.matvec_lo if lo else self.system.matvec
        bp = {"f64": 0, "f32": 1, "f16": 2}[self.basis]
        for j in range(j0, j1):
            zj = Z[j] if self.flexible else z
            pre(V[j], zj)
            mv(zj, w)
            if self._on_direction is not None:
                self._on_direction(zj, w)
            # Arnoldi step with the Givens update of column j folded in
            check(lib.rtx_gmres_arnoldi(bp, _vp(V), self.ld, n, _vp(w), _vp(self.hcol),
                                        _vp(self.work), self.refine, j, m, _vp(self.state), st),
                  "rtx_gmres_arnoldi")

    def solve(self, b, x=None, rtol=1e-8, atol=0.0, max_it=5000, callback=None,
              on_restart=None, on_direction=None, bnorm=None, x_zero=False,
              trust_estimate=False):
        """Converged when |b - A x| <= max(rtol |b|, atol) for the FP64 true
        residual.  ``on_restart(x, r)`` sees the iterate and its true residual
        at the start of every cycle; ``on_direction(z, w)`` sees every
        preconditioned direction z and w = A z before orthogonalisation (an
        exact residual / correction pair); either hook (or ``callback``)
        runs the iterations eagerly.  Returns (x, info) with PETSc-like
        convergence information.

        For the inner solves of iterative refinement: ``bnorm`` passes |b|
        when the caller has it; ``x_zero`` says x is zero, so the first
        cycle starts from r = b without an operator apply; with
        ``trust_estimate`` the solve ends after the cycle whose residual
        estimate reaches the target, without forming the true residual
        (``res_f`` is then that estimate; the outer FP64 residual checks the
        step).  Each saves an FP64 operator apply and a host read per inner
        solve."""
        system, lib = self.system, self.lib
        n, m = self.n, self.m
        V, Z, w, z, r = self.V, self.Z, self.w, self.z, self.r
        x = torch.zeros_like(b) if x is None else x
        use_graphs = self.graphs and callback is None and on_direction is None
        self._on_direction = on_direction
        bnorm = _norm(lib, b, self.work, self.scal) if bnorm is None else float(bnorm)
        target = max(rtol * bnorm, atol)
        history, true_hist = [], []
        its, reason, k_last, stop = 0, 0, 0, None
        est_met = False          # the last cycle's estimate reached the target
        floor = False
        first = True
        t0 = time.perf_counter()
        while True:
            if first and x_zero:
                rv, beta = b, bnorm               # r = b - A 0
            else:
                system.residual(b, x, r)
                rv, beta = r, _norm(lib, r, self.work, self.scal)
            first = False
            true_hist.append(beta)
            if on_restart is not None:
                on_restart(x, rv)
            if its == 0:
                history.append(beta)
            if beta <= target:
                reason = 2 if beta <= rtol * bnorm else 3     # KSP_CONVERGED_RTOL / ATOL
                break
            if (est_met and len(true_hist) >= 2 and beta > 0.5 * true_hist[-2]
                    and beta <= ATTAINABLE * bnorm):
                # rounding floor: the recurrence has converged, the true
                # residual cannot follow (PETSc stops on the estimate here)
                reason, floor = 2, True
                break
            if its >= max_it:
                reason = -3                                   # KSP_DIVERGED_ITS
                break
            if not np.isfinite(beta):
                reason = -9                                   # KSP_DIVERGED_NANORINF
                break
            st = stream_handle()
            if self.basis == "f32":
                check(lib.rtx_axpby_mixed(1, n, 1.0 / beta, _vp(rv), 0.0, _vp(V[0]), st), "axpby")
            elif self.basis == "f16":
                check(lib.rtx_axpby_mixed(2, n, 1.0 / beta, _vp(rv), 0.0, _vp(V[0]), st), "axpby")
            else:
                check(lib.rtx_axpby(PREC_F64, n, 1.0 / beta, _vp(rv), 0.0, _vp(V[0]), st),
                      "axpby")
            check(lib.rtx_gmres_begin(m, beta, _vp(self.state), st), "rtx_gmres_begin")
            gkey = system.graph_key() if hasattr(system, "graph_key") else None
            j, k, stop = 0, 0, None
            while j < m and its < max_it:
                j1 = min(j + self.chunk, m, j + (max_it - its))
                self.chunks.run((j, j1, gkey), lambda a=j, bb=j1: self._steps(a, bb), use_graphs)
                self.est_h[j:j1].copy_(self.state[self.est_off + j:self.est_off + j1],
                                       non_blocking=True)
                torch.cuda.current_stream().synchronize()
                e = self.est_h[j:j1].numpy()
                for i in range(j1 - j):
                    if not np.isfinite(e[i]):
                        stop = i               # breakdown: keep the directions before it
                        break
                    history.append(float(e[i]))
                    if callback is not None:
                        callback(its + i + 1, float(e[i]))
                    if e[i] <= target:
                        stop = i + 1
                        est_met = True
                        break
                if stop is not None:
                    its += stop
                    k = j + stop
                    break
                its += j1 - j
                j = k = j1
            k_last = k
            if stop is None:
                est_met = False
            if k > 0:
                check(lib.rtx_gmres_solve(k, m, _vp(self.state), _vp(self.ydev), st),
                      "rtx_gmres_solve")
                if self.basis == "f16":
                    check(lib.rtx_combine_f16(_vp(V), self.ld, k, _vp(self.ydev), n, _vp(w), 1, 0,
                                              st), "combine")
                    system.precondition_lo(w, z)
                    check(lib.rtx_axpby_mixed(0, n, 1.0, _vp(z), 1.0, _vp(x), st), "axpby")
                elif self.basis == "f32":
                    if self.flexible:
                        check(lib.rtx_combine_f32(_vp(Z), self.ld, k, _vp(self.ydev), n, _vp(x),
                                                  0, 1, st), "combine")
                    else:
                        check(lib.rtx_combine_f32(_vp(V), self.ld, k, _vp(self.ydev), n, _vp(w),
                                                  1, 0, st), "combine")
                        system.precondition_lo(w, z)
                        check(lib.rtx_axpby_mixed(0, n, 1.0, _vp(z), 1.0, _vp(x), st), "axpby")
                elif self.flexible:
                    check(lib.rtx_combine(_vp(Z), self.ld, k, _vp(self.ydev), n, _vp(x), 1, st),
                          "combine")
                else:
                    check(lib.rtx_combine(_vp(V), self.ld, k, _vp(self.ydev), n, _vp(w), 0, st),
                          "combine")
                    system.precondition(w, z)
                    check(lib.rtx_axpby(PREC_F64, n, 1.0, _vp(z), 1.0, _vp(x), st), "axpby")
            elif stop is not None:
                reason = -9
                break
            if trust_estimate and est_met:
                reason = 2 if history[-1] <= rtol * bnorm else 3
                break
        torch.cuda.synchronize()
        elapsed = time.perf_counter() - t0
        if k_last > 0:
            Hd = self.state[self.H_off:self.H_off + (m + 1) * m].view(m, m + 1).cpu().numpy().T
            sv = hessenberg_singular_values(Hd, k_last)
        else:
            sv = [float("nan")] * 2
        res_f = history[-1] if (trust_estimate and est_met) else true_hist[-1]
        info = GMRESResult(converged_reason=reason, convergence_history=history,
                           iteration_number=its, res_f=res_f, res_best=min(true_hist + [res_f]),
                           true_residual_history=true_hist, rtol=rtol, atol=atol,
                           divtol=float(np.sqrt(np.finfo(np.float64).max)), max_it=max_it,
                           bnorm=bnorm,
                           rel_residual=res_f / bnorm if bnorm > 0 else 0.0,
                           seconds=elapsed, restart=m, extreme_singular_values=sv,
                           graphs=len(self.chunks.graphs), attainable_floor=floor)
        return x, info


def gmres(system, b, x=None, rtol=1e-8, atol=0.0, restart=50, max_it=5000,
          flexible=False, callback=None, true_residual_every_restart=True, on_restart=None,
          on_direction=None, chunk=10, graphs=True, refine=0, basis="f64"):
    """One solve with a fresh ``GMRES`` object (see there)."""
    solver = GMRES(system, b.numel(), b.device, restart=restart, flexible=flexible, chunk=chunk,
                   graphs=graphs, refine=refine, basis=basis)
    return solver.solve(b, x, rtol=rtol, atol=atol, max_it=max_it, callback=callback,
                        on_restart=on_restart, on_direction=on_direction)


def hessenberg_singular_values(H, k):
    """Extreme singular values of the (k+1) x k Arnoldi matrix (PETSc's
    KSPComputeExtremeSingularValues estimates them the same way)."""
    if k == 0:
        return [float("nan"), float("nan")]
    s = np.linalg.svd(H[:k + 1, :k], compute_uv=False)
    return [float(s.max()), float(s.min())]


def pick_basis(system_lo, flexible):
    """The inner solves' Arnoldi basis: FP16 when the reduced-precision
    system's preconditioner is the FP8 or FP4 block-Jacobi and the method is
    not flexible, else FP32.  The inner solves stop at a relative 1e-2, well
    above FP16's 2^-11 rounding; on the uniform absorber (2,048,000 DOF) and
    the two-scatterer hp mesh (8,263,762 DOF) the inner iteration counts
    were those of the FP32 basis (200 and 202), with half the Gram-Schmidt
    bytes."""
    return "f16" if (not flexible and system_lo.fp16_basis_ok()) else "f32"


def refine(system_hi, system_lo, b, x=None, rtol=1e-8, inner_rtol=1e-2, restart=50,
           max_outer=50, max_inner=200, flexible=False, after_outer=None, chunk=10,
           graphs=True, inner=None, basis="auto", max_total=None):
    """FP64 iterative refinement: r = b - A x in FP64 (``system_hi``), inner
    correction solve A d = r by GMRES whose operator and preconditioner run in
    reduced precision (``system_lo``) on a ``basis`` "f16", "f32" or "f64"
    Arnoldi basis ("auto": FP16 when the preconditioner is the FP8 or FP4
    block-Jacobi and the method is not flexible, else FP32).  One GMRES
    object serves every inner solve, so its graphs are captured once.
    ``flexible`` selects flexible inner GMRES (for a nonlinear
    preconditioner); ``after_outer(outer, |r|)`` sees every FP64 residual;
    ``inner`` may pass a GMRES object on ``system_lo`` to reuse (with its
    graphs) across calls; ``max_total`` caps the inner iterations summed
    over all outer steps (reason -3 when it is spent, as PETSc's max_it)."""
    lib = system_hi.lib
    n = b.numel()
    x = torch.zeros_like(b) if x is None else x
    r = torch.empty_like(b)
    work_n = C.c_int64()
    check(lib.rtx_arnoldi_workspace(2, C.byref(work_n)), "workspace")
    work = torch.zeros(int(work_n.value), dtype=torch.float64, device=b.device)
    scal = torch.empty(2, dtype=torch.float64, device=b.device)
    bnorm = _norm(lib, b, work, scal)
    hist, inner_its = [], []
    t0 = time.perf_counter()
    reason, floor = -3, False
    if inner is None:
        if basis == "auto":
            basis = pick_basis(system_lo, flexible)
        inner = GMRES(system_lo, n, b.device, restart=restart, flexible=flexible, chunk=chunk,
                      graphs=graphs, basis=basis)
    d = torch.empty_like(b)
    for outer in range(max_outer + 1):
        system_hi.residual(b, x, r)
        rn = _norm(lib, r, work, scal)
        hist.append(rn)
        if after_outer is not None:
            after_outer(outer, rn)
        if rn <= rtol * bnorm:
            reason = 2
            break
        if (len(hist) >= 3 and rn > 0.5 * hist[-2] and hist[-2] > 0.5 * hist[-3]
                and rn <= ATTAINABLE * bnorm):
            # two outer steps without halving the FP64 residual: its rounding floor
            reason, floor = 2, True
            break
        if outer == max_outer or not np.isfinite(rn):
            reason = -3 if np.isfinite(rn) else -9
            break
        budget = max_inner
        if max_total is not None:
            budget = min(max_inner, max_total - sum(inner_its))
            if budget <= 0:
                reason = -3
                break
        d.zero_()
        d, info = inner.solve(r, d, rtol=inner_rtol, max_it=budget, bnorm=rn, x_zero=True,
                              trust_estimate=True)
        inner_its.append(info["iteration_number"])
        check(lib.rtx_axpby(PREC_F64, n, 1.0, _vp(d), 1.0, _vp(x), stream_handle()), "axpby")
    torch.cuda.synchronize()
    return x, dict(converged_reason=reason, residual_history=hist, inner_iterations=inner_its,
                   iteration_number=int(sum(inner_its)), res_f=hist[-1], bnorm=bnorm,
                   rel_residual=hist[-1] / bnorm if bnorm > 0 else 0.0,
                   seconds=time.perf_counter() - t0, attainable_floor=floor)
