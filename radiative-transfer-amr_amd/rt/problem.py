"""rt.Problem: the reference's operator/solve interface on the device path."""
import time
from typing import Callable

import numpy as np
import torch

import consts
from dg.projection import Projection, push_forward, pull_back
from dg.quadrature import quad_xyth, lag_eval
from rtx import _lib
from rtx.operator import TransportOperator, OP_QUADW, OP_SYSTEM, PREC_F64, PREC_F32, PREC_MIXED
from rtx._lib import PREC_FP8
from rtx import krylov

_PRECISIONS = {"fp64": PREC_F64, "fp32": PREC_F32, "mixed": PREC_MIXED}


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
        """Matrix-free device operator for ``mesh`` (cached for the last mesh
        object; a refined mesh is a new object and gets a new plan)."""
        if self._cache is not None and self._cache[0] is mesh:
            return self._cache[1]
        op = TransportOperator(mesh, self.kappa, self.sigma, self.Phi, device=device)
        self._cache = (mesh, op)
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
        op = self.operator(mesh)
        if dirac is None or not any(d is not None for d in dirac):
            return op.nodal(bcs)
        return _dirac_bcs_vector(mesh, bcs, dirac)

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
        convection, BF16 scattering and FP8 block-Jacobi inverses with
        inner GMRES(20) unless ``restart`` is given; "fp32" runs FP32
        throughout), inner_rtol=1e-2 (refinement only)."""
        kw = {"verbose": False, "ksp_type": "gmres", "pc_type": "bjacobi",
              "rtol": consts.EPS, "atol": consts.EPS, "max_it": 5000,
              "restart": 20 if kwargs.get("precision") == "mixed" else 50,
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
        prec = _PRECISIONS[kw["precision"]]
        t_setup = time.perf_counter()
        pc = (krylov.BlockJacobi(op, keep_f64=prec == PREC_F64, keep_f32=prec == PREC_F32,
                                 keep_fp8=prec == PREC_MIXED)
              if pc_used == "bjacobi" else None)
        t_pc = time.perf_counter()
        # the refinement's FP64 system only forms residuals: no preconditioner
        hi = krylov.InteriorSystem(op, PREC_F64, pc if prec == PREC_F64 else None, PREC_F64)
        if prec == PREC_F64:
            x, info = krylov.gmres(hi, rhs, x0, rtol=kw["rtol"], atol=kw["atol"],
                                   restart=kw["restart"], max_it=kw["max_it"],
                                   flexible=kw["ksp_type"] == "fgmres")
        else:
            pc_prec = {PREC_F32: PREC_F32, PREC_MIXED: PREC_FP8}[prec] if pc is not None else prec
            lo = krylov.InteriorSystem(op, prec, pc, pc_prec)
            x, info = krylov.refine(hi, lo, rhs, x0, rtol=max(kw["rtol"], 1e-14),
                                    inner_rtol=kw["inner_rtol"], restart=kw["restart"],
                                    max_inner=kw["max_it"])
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
            "seconds_total": t_end - t_start, "seconds_pc_setup": t_pc - t_setup,
            "seconds_krylov": info["seconds"],
        }
        matrix_info = {
            "ndof": int(op.ndof), "ndof_interior": int(op.mask(torch.ones(op.ndof, dtype=torch.float64, device=dev)).sum().item()),
            "block_size": 1,
            "extreme_singular_values": info.get("extreme_singular_values", [float("nan")] * 2),
            "pc_bytes": int(pc.bytes_per_apply(PREC_F64 if prec == PREC_F64 else PREC_F32))
            if pc is not None else 0,
        }
        if kw["verbose"]:
            print("Solve {} - {} finished: reason {}, {} iterations, residual {:.4e}, "
                  "{:.3f} s".format(pc_used, kw["ksp_type"], convergence_info["converged_reason"],
                                    convergence_info["iteration_number"],
                                    convergence_info["res_f"], t_end - t_start))
        return [uh, convergence_info, matrix_info]


def _dirac_bcs_vector(mesh, bcs, dirac):
    """Dirac-delta boundary data (``boundary_conditions_vector.py:142-196``),
    including its indexing behaviour: a cell none of whose coordinates hosts
    a delta leaves its entries unwritten and does not advance the index."""
    ndof = mesh.get_ndof()
    out = np.zeros(ndof)
    idx_g = 0
    for ck in sorted(mesh.cols.keys()):
        col = mesh.cols[ck]
        x0, y0, x1, y1 = col.pos
        nx, ny = col.ndofs
        xb, _, yb, _, _, _ = quad_xyth(nnodes_x=nx, nnodes_y=ny)
        xf, yf = push_forward(x0, x1, xb), push_forward(y0, y1, yb)
        col_start = idx_g
        idx = 0
        for kk in sorted(col.cells.keys()):
            cell = col.cells[kk]
            th0, th1 = cell.pos
            nth = cell.ndofs[0]
            thb = quad_xyth(nnodes_th=nth)[4]
            thf = push_forward(th0, th1, thb)
            in_cell = [True, True, True]
            pb = [None, None, None]
            for d, (lo, hi) in enumerate(((x0, x1), (y0, y1), (th0, th1))):
                if dirac[d] is not None:
                    if dirac[d] < lo or hi < dirac[d]:
                        in_cell[d] = False
                    else:
                        pb[d] = pull_back(lo, hi, dirac[d])
            if any(in_cell):
                for i in range(nx):
                    if dirac[0] is not None and in_cell[0]:
                        phi, xi = max(0.0, lag_eval(xb, i, pb[0])), dirac[0]
                    else:
                        phi, xi = 1.0, xf[i]
                    for j in range(ny):
                        if dirac[1] is not None and in_cell[1]:
                            psi, yj = max(0.0, lag_eval(yb, j, pb[1])), dirac[1]
                        else:
                            psi, yj = 1.0, yf[j]
                        for a in range(nth):
                            if dirac[2] is not None and in_cell[2]:
                                xsi, ta = max(0.0, lag_eval(thb, a, pb[2])), dirac[2]
                            else:
                                xsi, ta = 1.0, thf[a]
                            out[col_start + idx] = bcs(xi, yj, ta) * phi * psi * xsi
                            idx += 1
        nth_tot = sum(col.cells[k].ndofs[0] for k in col.cells)
        idx_g = col_start + nx * ny * nth_tot
    return out
