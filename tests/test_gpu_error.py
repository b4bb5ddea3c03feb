"""Device analytic L2 error (``rtx_error_l2``) against the oracle's
restatement of ``src/amr/Error_Indicator/error/error_analytic.py:22-149``,
and the ``Error_Indicator`` mirror built on it."""
import json
import math

import numpy as np
import pytest

from oracle import error as oerr, meshq
from rtx import problems

TOL = 1e-12


def _mesh():
    m = problems.uniform_mesh([1.0, 1.0], [3, 3, 3], 1, 2)
    keys = sorted(m.cols)
    m.ref_col(keys[0], "spt", "p")
    m.ref_col(keys[5], "spt", "h")
    m.ref_cell(keys[-1], sorted(m.cols[keys[-1]].cells)[0], "p")
    return m


def _u(x, y, th):
    return np.exp(-((x - 0.4) ** 2 + (y - 0.6) ** 2) / 0.05) * (1.0 + 0.5 * np.cos(th))


@pytest.mark.gpu
def test_cell_errors_match_oracle(cuda):
    from rtx.error import ErrorEvaluator
    from dg.projection import Projection
    mesh = _mesh()
    uh = Projection(mesh, _u).to_vector()
    uh = uh + 1e-3 * np.random.default_rng(0).normal(size=uh.size)
    ev = ErrorEvaluator(mesh)
    dev = ev.cell_errors(uh, ev.hr_values(_u))
    ref = np.array([e for _, _, e in oerr.cell_errors(mesh, uh, _u)])
    print("cells", ref.size, "max err", ref.max(), flush=True)
    assert dev.shape == ref.shape
    assert np.max(np.abs(dev - ref)) <= TOL * np.max(ref)


@pytest.mark.gpu
def test_error_indicator_matches_oracle_and_flags(cuda, tmp_path):
    from amr.error_indicator import Error_Indicator
    from dg.projection import Projection
    mesh = _mesh()
    proj = Projection(mesh, _u)
    intg = 1.2345
    ind = Error_Indicator(proj, ref_kind="all", ref_form="h", ref_tol=[0.5, 0.5])
    ind.error_analytic(_u, intg=intg)
    err, cols, cells = oerr.error_analytic(mesh, proj.to_vector(), _u, intg)
    assert abs(ind.error - err) <= TOL * err
    for ck, col in ind.cols.items():
        assert abs(col.error - cols[ck]) <= TOL * max(cols[ck], 1e-300)
        assert col.do_ref == (col.error >= 0.5 * ind.col_max_error)
        assert col.ref_form == ("h" if col.do_ref else "")
        for kk, cell in col.cells.items():
            assert abs(cell.error - cells[(ck, kk)]) <= TOL * max(cells[(ck, kk)], 1e-300)
            assert cell.do_ref == (cell.error >= 0.5 * ind.cell_max_error)
    assert abs(ind.col_max_error - max(cols.values())) <= TOL * max(cols.values())
    want = sum(c.error for c in ind.cols.values() if c.do_ref) + \
        sum(k.error for c in ind.cols.values() for k in c.cells.values() if k.do_ref)
    assert abs(ind.error_to_resolve - want) < 1e-12 * want
    f = tmp_path / "err_ind.json"
    ind.to_file(str(f), write_projection=False, write_mesh=False)
    d = json.loads(f.read_text())
    assert abs(d["error"] - ind.error) < 1e-15 and len(d["cols"]) == len(ind.cols)
    # ref_form "hp": every flagged column and cell is steered as the oracle's
    # steering rules decide on the oracle's integrals
    from oracle import indicator as oind
    ind = Error_Indicator(proj, ref_tol=[0.5, 0.5]).error_analytic(_u, intg=intg)
    u = proj.to_vector()
    steered = 0
    for ck, col in ind.cols.items():
        if col.do_ref:
            nx, ny = mesh.cols[ck].ndofs
            _, _, lp = oind.col_steer_lp(nx, ny, oind.col_intg_th(mesh, u, ck))
            assert col.ref_form == ("p" if lp - 0.5 >= 0.5 * (nx + ny) else "h")
            steered += 1
        for kk, cell in col.cells.items():
            if cell.do_ref:
                nth = mesh.cols[ck].cells[kk].ndofs[0]
                _, lp = oind.cell_steer_lp(nth, oind.cell_intg_xy(mesh, u, ck, kk))
                assert cell.ref_form == ("p" if lp - 0.5 >= nth else "h")
                steered += 1
    assert steered > 0
    # default tolerances flag nothing
    ind = Error_Indicator(proj).error_analytic(_u, intg=intg)
    assert not any(c.do_ref for c in ind.cols.values())
