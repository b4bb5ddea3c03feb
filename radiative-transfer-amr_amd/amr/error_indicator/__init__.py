"""``Error_Indicator``: error measures, refinement flags, hp steering and
refinement, with every per-DOF sum on the GPU.

Mirrors ``src/amr/Error_Indicator`` of the reference, imported as its
experiments do (``from amr.error_indicator import Error_Indicator``):

* ``class_Error_Indicator.py:11-29``: the constructor and attributes;
* ``error/error_analytic.py:22-190``: analytic L2 error (``rtx_error_l2``);
* ``error/error_high_resolution.py:27-271``: error against the solution of a
  p-refined mesh (``rtx_error_l2_nodal``; the solve is ``rt.Problem.solve``);
* ``error/error_angular_jump.py:14-137``, ``error/error_spatial_jump.py:
  17-289``: jump indicators (``rtx_error_angular_jump``,
  ``rtx_error_spatial_jump``);
* ``error/error_random.py:14-91``: random errors from a numpy generator;
* ``cell_hp_steer.py:15-48``, ``col_hp_steer.py:14-64``: h or p by the decay
  of the last Legendre coefficient of the angle- or space-integrated
  solution (integrals from ``rtx_indicator_moments``);
* ``ref_by_ind.py:13-46``: the refined mesh;
* ``to_file.py:11-60``, ``from_file.py:18-69``: the JSON layout.

The threshold and flag rules repeated at the end of each reference error
function (e.g. ``error_analytic.py:151-190``) are ``_flag`` here.
"""
import copy
import json

import numpy as np
from scipy.integrate import nquad
from scipy.special import eval_legendre

import consts
from dg.quadrature import quad_xyth

__all__ = ["Error_Indicator", "Error_Indicator_Column", "Error_Indicator_Cell", "from_file"]

# int u^2 over the domain, computed once per process like the reference's
# module-level cache (error_analytic.py:17, 28-32)
intg_u2 = None


class Error_Indicator_Cell:
    def __init__(self, error, ref_form="", do_ref=False):
        self.error = error
        self.ref_form = ref_form
        self.do_ref = do_ref


class Error_Indicator_Column:
    def __init__(self, error, cells, ref_form="", do_ref=False):
        self.error = error
        self.cells = cells
        self.ref_form = ref_form
        self.do_ref = do_ref


class Error_Indicator:
    def __init__(self, proj, ref_kind="all", ref_form="hp", ref_tol=(consts.INF, consts.INF)):
        self.proj = proj
        self.ref_kind = ref_kind
        self.ref_form = ref_form
        self.ang_ref_tol, self.spt_ref_tol = ref_tol
        self._evaluator = None
        self._moments = None
        self._t = None

    # ------------------------------------------------------------ device
    def _tables(self):
        """Host tables of the mesh (column / cell order and offsets)."""
        if self._t is None:
            from rtx.plan import MeshTables
            self._t = MeshTables(self.proj.mesh)
        return self._t

    def evaluator(self):
        if self._evaluator is None:
            from rtx.error import ErrorEvaluator
            self._evaluator = ErrorEvaluator(self.proj.mesh, tables=self._tables())
        return self._evaluator

    def _integrals(self):
        """Device integrals of the projection for hp steering, computed once."""
        if self._moments is None:
            xy, th, _ = self.evaluator().moments(self.proj.to_vector())
            t = self._tables()
            col_i = {ck: i for i, ck in enumerate(t.col_keys)}
            cell_i = {key: k for k, key in enumerate(t.cell_keys)}
            self._moments = (xy, th, col_i, cell_i)
        return self._moments

    def _cell_intg_xy(self, col_key, cell_key):
        xy, _, col_i, cell_i = self._integrals()
        t = self._tables()
        k, ci = cell_i[(col_key, cell_key)], col_i[col_key]
        a0 = int(t.col_ang_start[ci]) + int(t.cell_ang_off[k])
        return xy[a0:a0 + int(t.cell_nth[k])]

    def _col_intg_th(self, col_key):
        _, th, col_i, _ = self._integrals()
        t = self._tables()
        ci = col_i[col_key]
        nx, ny = int(t.col_nx[ci]), int(t.col_ny[ci])
        s0 = int(t.col_sp_off[ci])
        return th[s0:s0 + nx * ny].reshape(nx, ny)

    # -------------------------------------------------------- hp steering
    def cell_hp_steer(self, col_key, cell_key):
        """"p" when the last Legendre coefficient of the space-integrated
        solution decays fast enough for the cell's order, else "h"
        (``cell_hp_steer.py:15-48``)."""
        nth = self.proj.mesh.cols[col_key].cells[cell_key].ndofs[0]
        _, _, _, _, thb, wth = quad_xyth(nnodes_th=nth)
        thb = thb.reshape(1, nth)
        wth = wth.reshape(1, nth)
        uh_hat = self._cell_intg_xy(col_key, cell_key).reshape(1, nth)
        with np.errstate(all="ignore"):
            L = eval_legendre(nth - 1, thb)
            a_nth = (2.0 * nth - 1.0) / 2.0 * (wth * uh_hat) @ L.transpose()
            ath_sq = a_nth[0, 0] ** 2 * (2.0 / (2.0 * nth - 1.0))
            lp = np.log((2.0 * nth - 1.0) / (2 * ath_sq)) / (2.0 * np.log(nth - 1))
        return "p" if lp - 0.5 >= float(nth) else "h"

    def col_hp_steer(self, col_key):
        """"p" when the last Legendre coefficients of the angle-integrated
        solution in x and y decay fast enough for the column's orders, else
        "h" (``col_hp_steer.py:14-64``)."""
        nx, ny = self.proj.mesh.cols[col_key].ndofs
        xxb, wx, yyb, wy, _, _ = quad_xyth(nnodes_x=nx, nnodes_y=ny)
        xxb = xxb.reshape(1, nx)
        wx = wx.reshape(nx, 1)
        yyb = yyb.reshape(1, ny)
        wy = wy.reshape(1, ny)
        w_uh = wx * wy * self._col_intg_th(col_key)
        with np.errstate(all="ignore"):
            jy = np.arange(0, ny).reshape(1, ny)
            a_nxj = (2.0 * nx - 1.0) * (2.0 * jy + 1.0) / 4.0 * (
                eval_legendre(nx - 1, xxb) @ w_uh @ eval_legendre(jy.transpose(), yyb).transpose())
            ax_sq = 0
            for jj in range(ny):
                ax_sq += a_nxj[0, jj] ** 2 * (2.0 / (2.0 * jj + 1.0))
            ix = np.arange(0, nx).reshape(nx, 1)
            a_iny = (2.0 * ix + 1.0) * (2.0 * ny - 1.0) / 4.0 * (
                eval_legendre(ix, xxb) @ w_uh @ eval_legendre(ny - 1, yyb).transpose())
            ay_sq = 0
            for ii in range(nx):
                ay_sq += a_iny[ii, 0] ** 2 * (2.0 / (2.0 * ii + 1.0))
            t0 = np.log((2.0 * nx - 1.0) / (2 * ax_sq)) / (2.0 * np.log(nx - 1.0))
            t1 = np.log((2.0 * ny - 1.0) / (2 * ay_sq)) / (2.0 * np.log(ny - 1.0))
            lp = 0.5 * (t0 + t1)
        return "p" if lp - 0.5 >= 0.5 * (nx + ny) else "h"

    # ------------------------------------------------------------ flags
    def _set_errors(self, cols, col_max, cell_max, error):
        self.cols = cols
        self.col_max_error = col_max
        self.cell_max_error = cell_max
        self.error = error
        self.error_to_resolve = 0.0

    def _flag(self):
        """Refinement flags against ang_ref_tol * cell_max_error and
        spt_ref_tol * col_max_error, with the form (steered when "hp"), in the
        reference's order (``error_analytic.py:151-190``)."""
        with np.errstate(all="ignore"):
            ang_thr = self.ang_ref_tol * self.cell_max_error
            spt_thr = self.spt_ref_tol * self.col_max_error
        for ck in sorted(self.cols):
            col = self.cols[ck]
            if self.ref_kind in ("spt", "all"):
                if col.error >= spt_thr:
                    col.do_ref = True
                    self.error_to_resolve += col.error
                    col.ref_form = (self.col_hp_steer(ck) if self.ref_form == "hp"
                                    else self.ref_form)
                else:
                    col.do_ref = False
            if self.ref_kind in ("ang", "all"):
                for kk in sorted(col.cells):
                    cell = col.cells[kk]
                    if cell.error >= ang_thr:
                        cell.do_ref = True
                        self.error_to_resolve += cell.error
                        cell.ref_form = (self.cell_hp_steer(ck, kk) if self.ref_form == "hp"
                                         else self.ref_form)
                    else:
                        cell.do_ref = False

    def _from_cell_errors(self, cell_err, norm, root=True):
        """Columns of cell errors (global cell order): cell error
        sqrt(e / norm), column sqrt(sum e / norm), mesh sqrt(sum / norm);
        without ``root`` the plain sums (error_random.py:37-55)."""
        t = self._tables()
        f = (lambda e: np.sqrt(e / norm)) if root else (lambda e: e)
        cols = {}
        col_max, cell_max, mesh_err = -consts.INF, -consts.INF, 0.0
        for ci, ck in enumerate(t.col_keys):
            k0, k1 = int(t.col_cell_start[ci]), int(t.col_cell_start[ci + 1])
            cells = {}
            col_err = 0.0
            for k in range(k0, k1):
                e = float(cell_err[k])
                col_err += e
                mesh_err += e
                ce = f(e)
                cells[t.cell_keys[k][1]] = Error_Indicator_Cell(ce)
                cell_max = max(cell_max, ce)
            cols[ck] = Error_Indicator_Column(f(col_err), cells)
            col_max = max(col_max, f(col_err))
        self._set_errors(cols, col_max, cell_max, f(mesh_err))

    # ------------------------------------------------------------ errors
    def error_analytic(self, anl_sol, intg=None):
        """Relative L2 errors sqrt(err / int u^2) by cell, column and mesh
        (``error_analytic.py:22-149``) and the refinement flags
        (``:151-190``).  ``intg`` overrides the cached int u^2."""
        global intg_u2
        if intg is None:
            if intg_u2 is None:
                Lx, Ly = self.proj.mesh.Ls[:]
                intg_u2, _ = nquad(lambda x, y, th: anl_sol(x, y, th) ** 2,
                                   [[0, Lx], [0, Ly], [0, 2.0 * consts.PI]])
            intg = intg_u2
        ev = self.evaluator()
        cell_err = ev.cell_errors(self.proj.to_vector(), ev.hr_values(anl_sol))
        self._from_cell_errors(cell_err, intg)
        self._flag()
        return self

    def error_high_resolution(self, problem, uh_hr=None, **kwargs):
        """Relative L2 errors against the solution on the mesh p-refined
        ``ang_ref_offset`` times in angle and ``spt_ref_offset`` times in
        space (or against ``uh_hr``), normalised by int uh_hr^2
        (``error_high_resolution.py:27-271``).  Returns [uh_hr, info,
        mat_info] like the reference (info and mat_info None when ``uh_hr``
        is given)."""
        kw = {"ang_ref_offset": 0, "spt_ref_offset": 0, "ksp_type": "gmres",
              "pc_type": "bjacobi", "verbose": False, "blocking": True, **kwargs}
        info = mat_info = None
        if uh_hr is None:
            mesh_hr = copy.deepcopy(self.proj.mesh)
            for _ in range(kw["ang_ref_offset"]):
                mesh_hr.ref_mesh(kind="ang", form="p")
            for _ in range(kw["spt_ref_offset"]):
                mesh_hr.ref_mesh(kind="spt", form="p")
            # the reference passes the solver and preconditioner only; solve
            # tolerances and precision may be given here as well
            extra = {k: kw[k] for k in ("rtol", "atol", "max_it", "precision") if k in kw}
            uh_hr, info, mat_info = problem.solve(mesh_hr, ksp_type=kw["ksp_type"],
                                                  pc_type=kw["pc_type"], **extra)
        from rtx.plan import MeshTables
        ev = self.evaluator()
        t_hr = MeshTables(uh_hr.mesh)
        cell_err, hr_norm2 = ev.nodal_errors(self.proj.to_vector(), uh_hr.to_vector(), t_hr)
        self._from_cell_errors(cell_err, float(np.sum(hr_norm2)))
        self._flag()
        return [uh_hr, info, mat_info]

    def error_angular_jump(self):
        """Angular-jump indicator (``error_angular_jump.py:14-137``): cell
        error sqrt of its jump sum, column and mesh errors sqrt of the sums."""
        ev = self.evaluator()
        self._from_cell_errors(ev.angular_jump(self.proj.to_vector()), 1.0)
        self._flag()
        return self

    def error_spatial_jump(self):
        """Spatial-jump indicator (``error_spatial_jump.py:17-289``): column
        error sqrt(col_jump / (perimeter 2 pi)); cells carry -INF, as in the
        reference."""
        t = self._tables()
        col_err = self.evaluator().spatial_jump(self.proj.to_vector())
        cols = {}
        col_max, mesh_err = -consts.INF, 0.0
        for ci, ck in enumerate(t.col_keys):
            k0, k1 = int(t.col_cell_start[ci]), int(t.col_cell_start[ci + 1])
            cells = {t.cell_keys[k][1]: Error_Indicator_Cell(-consts.INF) for k in range(k0, k1)}
            e = float(col_err[ci])
            mesh_err += e
            cols[ck] = Error_Indicator_Column(np.sqrt(e), cells)
            col_max = max(col_max, np.sqrt(e))
        self._set_errors(cols, col_max, -consts.INF, np.sqrt(mesh_err))
        self._flag()
        return self

    def error_random(self, rng=None):
        """Uniform random cell errors from ``rng`` (a fresh default generator
        when None), in column then cell key order (``error_random.py:14-91``)."""
        if rng is None:
            rng = np.random.default_rng()
        t = self._tables()
        cell_err = np.array([rng.random() for _ in range(t.ncell)])
        self._from_cell_errors(cell_err, 1.0, root=False)
        self._flag()
        return self

    # -------------------------------------------------------- refinement
    def ref_by_ind(self):
        """A copy of the mesh refined where flagged: first the flagged cells of
        a column, then the column, skipping what an earlier 2:1 balance has
        already refined (``ref_by_ind.py:13-46``)."""
        mesh = copy.deepcopy(self.proj.mesh)
        for ck in sorted(self.proj.mesh.cols):
            if ck not in mesh.cols:
                continue
            if self.ref_kind in ("ang", "all"):
                for kk in sorted(self.proj.mesh.cols[ck].cells):
                    if kk in mesh.cols[ck].cells:
                        cell = self.cols[ck].cells[kk]
                        if cell.do_ref:
                            mesh.ref_cell(ck, kk, form=cell.ref_form)
            if self.ref_kind in ("spt", "all"):
                col = self.cols[ck]
                if col.do_ref:
                    mesh.ref_col(ck, kind="spt", form=col.ref_form)
        return mesh

    # ---------------------------------------------------------------- io
    def to_file(self, file_path="err_ind.json", **kwargs):
        """JSON of the indicator (``to_file.py:11-60``)."""
        kw = {"write_mesh": True, "mesh_file_path": "mesh.json", "write_projection": True,
              "projection_file_path": "projection.npy", **kwargs}
        if kw["write_projection"]:
            self.proj.to_file(kw["projection_file_path"], **kw)
        d = {"ref_kind": self.ref_kind, "ref_form": self.ref_form,
             "ang_ref_tol": self.ang_ref_tol, "spt_ref_tol": self.spt_ref_tol,
             "error": float(self.error), "error_to_resolve": float(self.error_to_resolve),
             "col_max_error": float(self.col_max_error),
             "cell_max_error": float(self.cell_max_error)}
        cols = {}
        for ck in sorted(self.cols):
            col = self.cols[ck]
            cells = {}
            for kk in sorted(col.cells):
                c = col.cells[kk]
                cells[kk] = {"do_ref": c.do_ref, "ref_form": c.ref_form, "error": float(c.error)}
            cols[ck] = {"do_ref": col.do_ref, "ref_form": col.ref_form, "error": float(col.error),
                        "cells": cells}
        d["cols"] = cols
        with open(file_path, "w") as f:
            json.dump(d, f)


def from_file(mesh_file_path="mesh.json", projection_file_path="projection.npy",
              err_ind_file_path="err_ind.json"):
    """Read an indicator written by ``Error_Indicator.to_file`` with its
    projection (``from_file.py:18-69``)."""
    from dg.projection import from_file as projection_from_file
    proj = projection_from_file(mesh_file_path, projection_file_path)
    with open(err_ind_file_path, "r") as f:
        d = json.load(f)
    ind = Error_Indicator(proj, ref_kind=d["ref_kind"], ref_form=d["ref_form"],
                          ref_tol=[d["ang_ref_tol"], d["spt_ref_tol"]])
    ind.error = d["error"]
    ind.error_to_resolve = d["error_to_resolve"]
    ind.col_max_error = d["col_max_error"]
    ind.cell_max_error = d["cell_max_error"]
    cols = {}
    for ck in sorted(d["cols"].keys()):
        cd = d["cols"][ck]
        cells = {int(kk): Error_Indicator_Cell(cd["cells"][kk]["error"], cd["cells"][kk]["ref_form"],
                                               cd["cells"][kk]["do_ref"])
                 for kk in sorted(cd["cells"].keys())}
        cols[int(ck)] = Error_Indicator_Column(cd["error"], cells, cd["ref_form"], cd["do_ref"])
    ind.cols = cols
    return ind
