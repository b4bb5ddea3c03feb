"""Host logic of ``rt.Problem`` that needs no GPU: the mesh signature that
keys the cached device plan (a ``Problem`` of the reference rebuilds every
matrix per call, ``src/rt/class_Problem.py:11-32``, so any change of the
leaf columns or cells must invalidate the plan)."""
import copy

from rtx import problems


def test_mesh_signature_tracks_in_place_changes():
    from rt.problem import mesh_signature
    mesh = problems.uniform_mesh([1.0, 1.0], [3, 3, 3], 2, 1)
    s0 = mesh_signature(mesh)
    assert mesh_signature(mesh) == s0
    assert mesh_signature(copy.deepcopy(mesh)) == s0
    # spatial refinement of one column
    m1 = copy.deepcopy(mesh)
    m1.ref_col(sorted(m1.cols)[0], kind="spt", form="h")
    assert mesh_signature(m1) != s0
    # angular refinement of one cell
    m2 = copy.deepcopy(mesh)
    ck = sorted(m2.cols)[0]
    m2.ref_cell(ck, sorted(m2.cols[ck].cells)[0], form="h")
    assert mesh_signature(m2) != s0
    # p-refinement: same keys, other orders
    m3 = copy.deepcopy(mesh)
    m3.ref_mesh(kind="ang", form="p")
    assert mesh_signature(m3) != s0
    m4 = copy.deepcopy(mesh)
    col = m4.cols[sorted(m4.cols)[-1]]
    col.ndofs = [col.ndofs[0] + 1, col.ndofs[1]]
    assert mesh_signature(m4) != s0
