"""Host logic of Error_Indicator that needs no GPU: random errors
(error_random.py:14-91), refinement by the flags (ref_by_ind.py:13-46) and
the JSON round trip (to_file.py, from_file.py)."""
import numpy as np
import pytest

from amr.error_indicator import Error_Indicator, from_file
from dg.projection import Projection
from oracle.meshq import leaf_order
from rtx import problems
from test_oracle_indicator import hp_mesh


def _proj(mesh):
    p = Projection(mesh)
    p.from_vector(np.random.default_rng(0).uniform(-1, 1, mesh.get_ndof()))
    return p


def test_error_random_sequence_and_flags():
    mesh = hp_mesh()
    ind = Error_Indicator(_proj(mesh), ref_kind="all", ref_form="h", ref_tol=[0.8, 0.7])
    ind.error_random(rng=np.random.default_rng(11))
    draws = np.random.default_rng(11).random(sum(len(k) for _, k in leaf_order(mesh)))
    i = 0
    tot = 0.0
    for ck, kks in leaf_order(mesh):
        col_sum = 0.0
        for kk in kks:
            assert ind.cols[ck].cells[kk].error == draws[i]
            col_sum += draws[i]
            i += 1
        assert ind.cols[ck].error == pytest.approx(col_sum, rel=1e-15)
        tot += col_sum
    assert ind.error == pytest.approx(tot, rel=1e-15)
    assert ind.cell_max_error == draws.max()
    for ck, col in ind.cols.items():
        assert col.do_ref == (col.error >= 0.7 * ind.col_max_error)
        assert col.ref_form == ("h" if col.do_ref else "")
        for kk, cell in col.cells.items():
            assert cell.do_ref == (cell.error >= 0.8 * ind.cell_max_error)


def test_ref_by_ind_refines_the_flagged_entities():
    mesh = problems.uniform_mesh([1.0, 1.0], [3, 3, 3], 2, 2)
    ind = Error_Indicator(_proj(mesh), ref_kind="all", ref_form="h", ref_tol=[0.9, 0.9])
    ind.error_random(rng=np.random.default_rng(2))
    new = ind.ref_by_ind()
    assert new is not mesh and sorted(mesh.cols) == sorted(ind.proj.mesh.cols)
    flagged_cols = [ck for ck, c in ind.cols.items() if c.do_ref]
    flagged_cells = [(ck, kk) for ck, c in ind.cols.items() for kk, k in c.cells.items()
                     if k.do_ref]
    assert flagged_cols and flagged_cells
    for ck in flagged_cols:
        # h-refined: the column is replaced by its four children
        assert ck not in new.cols
        from dg.mesh.column import calc_key
        lv = mesh.cols[ck].lv
        i0, j0 = mesh.cols[ck].idx
        kids = {calc_key([2 * i0 + di, 2 * j0 + dj], lv + 1) for di in (0, 1) for dj in (0, 1)}
        assert kids <= set(new.cols)
    for ck, kk in flagged_cells:
        if ck in new.cols:
            assert kk not in new.cols[ck].cells
    assert new.get_ndof() > mesh.get_ndof()
    # ref_kind "ang" leaves the columns alone
    ind.ref_kind = "ang"
    new = ind.ref_by_ind()
    assert sorted(new.cols) == sorted(mesh.cols)


def test_to_file_from_file_round_trip(tmp_path):
    mesh = hp_mesh()
    proj = _proj(mesh)
    ind = Error_Indicator(proj, ref_kind="all", ref_form="h", ref_tol=[0.6, 0.6])
    ind.error_random(rng=np.random.default_rng(4))
    m, p, e = (str(tmp_path / n) for n in ("mesh.json", "projection.npy", "err_ind.json"))
    ind.to_file(e, mesh_file_path=m, projection_file_path=p)
    back = from_file(m, p, e)
    assert back.ref_kind == "all" and back.ref_form == "h"
    assert back.error == ind.error and back.col_max_error == ind.col_max_error
    assert np.array_equal(back.proj.to_vector(), proj.to_vector())
    for ck, col in ind.cols.items():
        b = back.cols[ck]
        assert (b.error, b.do_ref, b.ref_form) == (col.error, col.do_ref, col.ref_form)
        for kk, cell in col.cells.items():
            bc = b.cells[kk]
            assert (bc.error, bc.do_ref, bc.ref_form) == (cell.error, cell.do_ref, cell.ref_form)
