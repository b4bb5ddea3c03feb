"""``Error_Indicator`` with the analytic L2 error on the GPU.

Mirrors ``src/amr/Error_Indicator/class_Error_Indicator.py:11-29`` (the
constructor and attributes), ``error/error_analytic.py:22-190`` (the errors
and refinement flags) and ``to_file.py:11-62`` (the JSON layout), imported
as the reference's experiments do (``from amr.error_indicator import
Error_Indicator``).  The per-cell interpolation to n + 3 nodes and the
weighted squared error run in ``rtx_error_l2``; the analytic solution is
evaluated on the host at the same nodes.  hp steering (``col_hp_steer``,
``cell_hp_steer``) is outside this build and raises when a refinement
would need it.
"""
import json

import numpy as np
from scipy.integrate import nquad

import consts

__all__ = ["Error_Indicator", "Error_Indicator_Column", "Error_Indicator_Cell"]

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

    # ------------------------------------------------------------------
    def col_hp_steer(self, col_key):
        raise NotImplementedError("hp steering (src/amr/Error_Indicator/col_hp_steer.py) is "
                                  "not part of this build; use ref_form 'h' or 'p'")

    def cell_hp_steer(self, col_key, cell_key):
        raise NotImplementedError("hp steering (src/amr/Error_Indicator/cell_hp_steer.py) is "
                                  "not part of this build; use ref_form 'h' or 'p'")

    def evaluator(self):
        if self._evaluator is None:
            from rtx.error import ErrorEvaluator
            self._evaluator = ErrorEvaluator(self.proj.mesh)
        return self._evaluator

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
        t = ev.t
        cell_err = ev.cell_errors(self.proj.to_vector(), ev.hr_values(anl_sol))
        cols = {}
        col_max, cell_max, mesh_err = -consts.INF, -consts.INF, 0.0
        for ci, ck in enumerate(t.col_keys):
            k0, k1 = int(t.col_cell_start[ci]), int(t.col_cell_start[ci + 1])
            cells = {}
            col_err = 0.0
            for k in range(k0, k1):
                e = float(cell_err[k])
                col_err += e
                ce = np.sqrt(e / intg)
                cells[t.cell_keys[k][1]] = Error_Indicator_Cell(ce)
                cell_max = max(cell_max, ce)
            mesh_err += col_err
            cols[ck] = Error_Indicator_Column(np.sqrt(col_err / intg), cells)
            col_max = max(col_max, np.sqrt(col_err / intg))
        self.cols = cols
        self.col_max_error = col_max
        self.cell_max_error = cell_max
        self.error = np.sqrt(mesh_err / intg)
        self.error_to_resolve = 0.0
        ang_thr = self.ang_ref_tol * self.cell_max_error
        spt_thr = self.spt_ref_tol * self.col_max_error
        for ck, col in cols.items():
            if self.ref_kind in ("spt", "all"):
                if col.error >= spt_thr:
                    col.do_ref = True
                    self.error_to_resolve += col.error
                    col.ref_form = (self.col_hp_steer(ck) if self.ref_form == "hp"
                                    else self.ref_form)
                else:
                    col.do_ref = False
            if self.ref_kind in ("ang", "all"):
                for kk, cell in col.cells.items():
                    if cell.error >= ang_thr:
                        cell.do_ref = True
                        self.error_to_resolve += cell.error
                        cell.ref_form = (self.cell_hp_steer(ck, kk) if self.ref_form == "hp"
                                         else self.ref_form)
                    else:
                        cell.do_ref = False
        return self

    def to_file(self, file_path="err_ind.json", **kwargs):
        """JSON of the indicator (``to_file.py:11-62``)."""
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
