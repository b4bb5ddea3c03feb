"""rt.Problem: the reference's operator/solve interface on the device path."""
import time
import warnings
from typing import Callable

import numpy as np
import torch

import consts
from dg.projection import Projection, push_forward, pull_back
from dg.quadrature import quad_xyth, lag_eval
from rtx import _lib
from rtx.operator import TransportOperator, OP_QUADW, OP_SYSTEM, PREC_F64, PREC_F32, PREC_MIXED
from rtx._lib import PREC_FP4
from rtx import krylov

_PRECISIONS = {"fp64": PREC_F64, "fp32": PREC_F32, "mixed": PREC_MIXED}


class KaczmarzSubstituted(UserWarning):
    """``pc_type="kaczmarz"`` was run as per-cell block-Jacobi."""


def mesh_signature(mesh) -> tuple:
    """Leaf columns and cells of ``mesh``: (key, position, orders) of each,
    in key order, with the domain size and periodicity."""
    cols = []
    for ck in sorted(mesh.cols):
        col = mesh.cols[ck]
        if not col.is_lf:
            continue
        cells = tuple((kk, cell.ndofs[0], cell.pos[0], cell.pos[1])
                      for kk, cell in sorted(col.cells.items()) if cell.is_lf)
        cols.append((ck, col.ndofs[0], col.ndofs[1], tuple(col.pos), cells))
    return (tuple(mesh.Ls), tuple(mesh.pbcs), tuple(cols))


class Problem:
    def __init__(self, kappa: Callable, sigma: Callable, Phi: Callable, f: Callable,
                 bcs_dirac: list):
        self.kappa = kappa
        self.sigma = sigma
        self.Phi = Phi
        self.f = f
        self.bcs_dirac = bcs_dirac
        self._cache = None

    # ------------------------------------------------------------ operator
    def operator(self, mesh, device=None) -> TransportOperator:
        """Matrix-free device operator for ``mesh``.  The last plan is reused
        while the mesh is the same object with the same leaf columns and
        cells (keys, positions and orders); a mesh refined in place
        (``ref_mesh``, ``ref_col``, ``ref_cell``) or with changed orders gets
        a new plan, as the reference rebuilds every matrix per call."""
        sig = mesh_signature(mesh)
        if self._cache is not None and self._cache[0] is mesh and self._cache[1] == sig:
            return self._cache[2]
        self._cache = None
        op = TransportOperator(mesh, self.kappa, self.sigma, self.Phi, device=device)
        self._cache = (mesh, sig, op)
        return op

    # ------------------------------------------------------------- vectors
    def forcing_vector(self, mesh, **kwargs) -> np.ndarray:
        """F_beta = (dx dy dth / 8) w_i w_j w_a f(x_i, y_j, th_a)
        (``src/rt/forcing_vector.py:103-156``); nodal values from the callable,
        quadrature weighting on the device."""
        op = self.operator(mesh)
        fn = torch.from_numpy(op.nodal(self.f)).to(op.device)
        return op.apply(fn, ops=OP_QUADW).cpu().numpy()

    def boundary_conditions_vector(self, mesh, **kwargs) -> np.ndarray:
        """Nodal boundary data at every DOF
        (``src/rt/boundary_conditions_vector.py:101-207``)."""
        bcs, dirac = self.bcs_dirac
        # the reference's trigger is ``any(dirac)`` (boundary_conditions_vector.py:142):
        # a delta at coordinate 0.0 alone takes the nodal branch
        if dirac is None or not any(dirac):
            return self.operator(mesh).nodal(bcs)
        return dirac_bcs_vector(mesh, bcs, dirac)

    # ------------------------------------------------------------ matrices
    def mass_matrix(self, mesh, **kwargs):
        from rtx import assembly
        return assembly.assemble(self.operator(mesh), "M")

    def scattering_matrix(self, mesh, **kwargs):
        """The reference's returned matrix, i.e. -S."""
        from rtx import assembly
        return assembly.assemble(self.operator(mesh), "S")

    def interior_convection_matrix(self, mesh, **kwargs):
        from rtx import assembly
        return assembly.assemble(self.operator(mesh), "IC")

    def boundary_convection_matrix(self, mesh, **kwargs):
        from rtx import assembly
        return assembly.assemble(self.operator(mesh), "BC")

    def preconditioner_matrix(self, mesh, return_preconditioner: bool = False, **kwargs):
        """[M + S_ret, None] (``src/rt/preconditioner_matrix.py:22-114``); with
        ``return_preconditioner`` the second entry is the device block-Jacobi
        preconditioner object."""
        from rtx import assembly
        op = self.operator(mesh)
        P = assembly.assemble(op, "P")
        return [P, krylov.BlockJacobi(op) if return_preconditioner else None]

    # --------------------------------------------------------------- solve
    def solve(self, mesh, **kwargs) -> list:
        """Solve A_II x_I = f_I - A_IB b_B with x0 = f_I (``src/rt/solve.py:21-149``).

        kwargs (reference defaults first): verbose=False, ksp_type="gmres"
        ("fgmres" accepted), pc_type="bjacobi" ("kaczmarz" and "none"
        accepted; "kaczmarz" falls back to block-Jacobi and is reported in
        convergence_info["pc_type_used"]), rtol=EPS, atol=EPS, max_it=5000,
        restart=50, precision="fp64" | "fp32" | "mixed" (FP64 iterative
        refinement around reduced-precision applies: "mixed" runs FP32
        convection, BF16 scattering and FP4 block-Jacobi inverses with
        inner GMRES(15) on an FP16 basis unless ``restart`` is given; "fp32"
        runs FP32 throughout), inner_rtol=1e-2 (refinement only)."""
        kw = {"verbose": False, "ksp_type": "gmres", "pc_type": "bjacobi",
              "rtol": consts.EPS, "atol": consts.EPS, "max_it": 5000,
              "restart": krylov.INNER_RESTART if kwargs.get("precision") == "mixed" else 50,
              "precision": "fp64", "inner_rtol": 1e-2, **kwargs}
        t_start = time.perf_counter()
        op = self.operator(mesh)
        dev = op.device
        f_vec = torch.from_numpy(self.forcing_vector(mesh)).to(dev)
        b_vec = torch.from_numpy(np.ascontiguousarray(self.boundary_conditions_vector(mesh))).to(dev)
        # rhs = f_I - A_IB b_B, in full-length form with zeros on the boundary
        b_bdry = op.mask(b_vec.clone(), keep_interior=False)
        rhs = f_vec - op.apply(b_bdry, ops=OP_SYSTEM)
        op.mask(rhs, keep_interior=True)
        x0 = op.mask(f_vec.clone(), keep_interior=True)
        pc_type = kw["pc_type"]
        pc_used = {"bjacobi": "bjacobi", "kaczmarz": "bjacobi", "none": "none"}.get(pc_type)
        if pc_used is None:
            raise ValueError("unsupported pc_type {!r}".format(pc_type))
        if pc_type != pc_used:
            # PETSc's Kaczmarz sweep is sequential over rows; the device path
            # runs its per-cell block-Jacobi in its place
            warnings.warn("pc_type {!r} runs as per-cell block-Jacobi on the device "
                          "(convergence_info['pc_type_used'])".format(pc_type),
                          KaczmarzSubstituted, stacklevel=2)
        prec = _PRECISIONS[kw["precision"]]
        t_setup = time.perf_counter()
        pc = (krylov.BlockJacobi(op, keep_f64=prec == PREC_F64, keep_f32=prec == PREC_F32,
                                 keep_fp4=prec == PREC_MIXED)
              if pc_used == "bjacobi" else None)
        t_pc = time.perf_counter()
        # the refinement's FP64 system only forms residuals: no preconditioner
        hi = krylov.InteriorSystem(op, PREC_F64, pc if prec == PREC_F64 else None, PREC_F64)
        pc_prec = PREC_F64
        if prec == PREC_F64:
            x, info = krylov.gmres(hi, rhs, x0, rtol=kw["rtol"], atol=kw["atol"],
                                   restart=kw["restart"], max_it=kw["max_it"],
                                   flexible=kw["ksp_type"] == "fgmres")
        else:
            pc_prec = {PREC_F32: PREC_F32, PREC_MIXED: PREC_FP4}[prec] if pc is not None else prec
            lo = krylov.InteriorSystem(op, prec, pc, pc_prec)
            x, info = krylov.refine(hi, lo, rhs, x0, rtol=max(kw["rtol"], 1e-14),
                                    inner_rtol=kw["inner_rtol"], restart=kw["restart"],
                                    max_inner=kw["max_it"], max_total=kw["max_it"],
                                    chunk=krylov.INNER_CHUNK if prec == PREC_MIXED else 10)
        u = (x + b_bdry).cpu().numpy()
        t_end = time.perf_counter()
        uh = Projection(mesh)
        uh.from_vector(u)
        hist = info.get("convergence_history", info.get("residual_history", []))
        convergence_info = {
            "converged_reason": info["converged_reason"],
            "convergence_history": [float(v) for v in hist],
            "iteration_number": int(info["iteration_number"]),
            "res_f": float(info["res_f"]),
            "res_best": float(min(hist)) if hist else float(info["res_f"]),
            "rtol": kw["rtol"], "atol": kw["atol"],
            "divtol": float(np.sqrt(consts.INF)), "max_it": kw["max_it"],
            "pc_type_used": pc_used, "precision": kw["precision"],
            "attainable_floor": bool(info.get("attainable_floor", False)),
            "seconds_total": t_end - t_start, "seconds_pc_setup": t_pc - t_setup,
            "seconds_krylov": info["seconds"],
        }
        matrix_info = {
            "ndof": int(op.ndof), "ndof_interior": int(op.mask(torch.ones(op.ndof, dtype=torch.float64, device=dev)).sum().item()),
            "block_size": 1,
            "extreme_singular_values": info.get("extreme_singular_values", [float("nan")] * 2),
            "pc_bytes": int(pc.bytes_per_apply(pc_prec if prec != PREC_F64 else PREC_F64))
            if pc is not None else 0,
        }
        if kw["verbose"]:
            print("Solve {} - {} finished: reason {}, {} iterations, residual {:.4e}, "
                  "{:.3f} s".format(pc_used, kw["ksp_type"], convergence_info["converged_reason"],
                                    convergence_info["iteration_number"],
                                    convergence_info["res_f"], t_end - t_start))
        return [uh, convergence_info, matrix_info]


def _dirac_factors(nodes_ref, nodes_phys, lo, hi, delta):
    """Per-node factor and coordinate along one direction of a cell: with a
    delta inside [lo, hi], max(0, l_i(xi_delta)) and the delta's coordinate;
    otherwise 1 and the node's own coordinate.  Returns (inside, factor,
    coordinate); a direction without a delta counts as inside, as in the
    reference."""
    n = len(nodes_ref)
    if delta is None or delta < lo or hi < delta:
        return delta is None, np.ones(n), np.asarray(nodes_phys, dtype=float)
    xi = float(pull_back(lo, hi, delta))
    phi = np.array([max(0.0, lag_eval(nodes_ref, i, xi)) for i in range(n)])
    return True, phi, np.full(n, float(delta))


def dirac_bcs_vector(mesh, bcs, dirac):
    """Boundary data with Dirac deltas (``boundary_conditions_vector.py:142-196``).

    For each direction whose delta coordinate lies in the cell's interval the
    nodal value is replaced by bcs at the delta's coordinate times
    max(0, l_i(xi_delta)); other directions keep their nodes.  A cell in which
    every given delta coordinate lies outside the cell (and all three are
    given) is skipped without advancing the column's write position, as in
    the reference, so the column's later cells land earlier and its last
    entries stay zero.  Vectorised per cell: one bcs call on the cell's
    broadcast (x, y, theta) node arrays."""
    ndof = mesh.get_ndof()
    out = np.zeros(ndof)
    col_start = 0
    for ck in sorted(mesh.cols.keys()):
        col = mesh.cols[ck]
        if not col.is_lf:
            continue
        x0, y0, x1, y1 = col.pos
        nx, ny = col.ndofs
        xb, _, yb, _, _, _ = quad_xyth(nnodes_x=nx, nnodes_y=ny)
        in_x, phi, xs = _dirac_factors(xb, push_forward(x0, x1, xb), x0, x1, dirac[0])
        in_y, psi, ys = _dirac_factors(yb, push_forward(y0, y1, yb), y0, y1, dirac[1])
        w = col_start
        col_ndof = 0
        for kk in sorted(col.cells.keys()):
            cell = col.cells[kk]
            if not cell.is_lf:
                continue
            th0, th1 = cell.pos
            nth = cell.ndofs[0]
            col_ndof += nx * ny * nth
            thb = quad_xyth(nnodes_th=nth)[4]
            in_t, xsi, ts = _dirac_factors(thb, push_forward(th0, th1, thb), th0, th1, dirac[2])
            if not (in_x or in_y or in_t):
                continue
            X, Y, T = np.meshgrid(xs, ys, ts, indexing="ij")
            vals = np.broadcast_to(np.asarray(bcs(X, Y, T), dtype=float), X.shape)
            # products in the reference's order: ((bcs phi) psi) xsi
            blk = (((vals * phi[:, None, None]) * psi[None, :, None])
                   * xsi[None, None, :]).ravel()
            out[w:w + blk.size] = blk
            w += blk.size
        col_start += col_ndof
    return out
