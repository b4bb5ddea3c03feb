"""Host mesh, projection and device-table logic (CPU).

Mirrors the reference's structural tests: the refinement sequence of
``tests/test_dg/test_mesh/test_Mesh.py:40-62`` (a known-answer test for
``calc_key`` and refinement), the mesh JSON round trip (``test_Mesh.py:27-32``)
and the projection vector/file round trip (``test_Projection.py:44-47``).
Adds 2:1 balance, neighbour symmetry, geometric adjacency, and the device
tables (DOF offsets, masked faces, upwind links, face-matrix pools) against
the oracle's independent restatements.
"""
import math

import numpy as np
import pytest

from oracle import emat, meshq, assemble as asm
from rtx import problems


def reference_sequence_mesh():
    """``test_Mesh.py:13-62``: Mesh([1, 2], [F, F], [2, 3, 4], True), then
    columns 0, 0, 0, 1, 10, 48 refined (ang h, ang h, spt h, spt h, spt hp,
    spt hp) and cells 3, 7, 15 of column 204 (h, hp, hp)."""
    from dg.mesh import Mesh
    mesh = Mesh([1.0, 2.0], [False, False], [2, 3, 4], True)
    for key, kind, form in zip([0, 0, 0, 1, 10, 48], ["ang", "ang", "spt", "spt", "spt", "spt"],
                               ["h", "h", "h", "h", "hp", "hp"]):
        mesh.ref_col(col_key=key, kind=kind, form=form)
    for ck, form in zip([3, 7, 15], ["h", "hp", "hp"]):
        mesh.ref_cell(col_key=204, cell_key=ck, form=form)
    return mesh


def test_calc_keys_known_answers():
    from dg.mesh.column import calc_key as ckey
    from dg.mesh.column.cell import calc_key as kkey
    assert [ckey([0, 0], 0), ckey([0, 0], 1), ckey([1, 0], 1), ckey([0, 1], 1),
            ckey([1, 1], 1)] == [0, 1, 2, 3, 4]
    # level 2 starts after 1 + 4 keys; idx[1] strides by 2^lv
    assert ckey([0, 0], 2) == 5 and ckey([1, 0], 2) == 6 and ckey([0, 1], 2) == 9
    assert [kkey(0, 0), kkey(0, 1), kkey(1, 1), kkey(0, 2), kkey(3, 2)] == [0, 1, 2, 3, 6]


def test_reference_refinement_sequence():
    mesh = reference_sequence_mesh()
    keys = set(mesh.cols)
    # the sequence only runs if each refined key exists when it is reached
    # (1..4 after the first split, 10 under 1, 48 under 10, 204 under 48);
    # the 2:1 rule splits the level-1 neighbours 2 and 3 before 10 splits
    for k in (187, 188, 203, 204, 39, 40, 47, 5, 6, 9):
        assert k in keys, k
    for k in (0, 1, 2, 3, 4, 10, 48):
        assert k not in keys, k          # refined columns are replaced
    assert all(c.is_lf for c in mesh.cols.values())
    # the whole mesh (keys, orders, cells, quadrants) equals the geometric
    # model of the refinement rules run on the same sequence
    from refine_model import RefineModel, mesh_summary
    model = RefineModel(2, 3, 4)
    for key, kind, form in zip([0, 0, 0, 1, 10, 48], ["ang", "ang", "spt", "spt", "spt", "spt"],
                               ["h", "h", "h", "h", "hp", "hp"]):
        model.ref_col(key, kind, form)
    for ck, form in zip([3, 7, 15], ["h", "hp", "hp"]):
        model.ref_cell(204, ck, form)
    assert mesh_summary(mesh) == model.summary()
    cells = mesh.cols[204].cells
    # cell 3 -> 7, 8; 7 -> 15, 16; 15 -> 31, 32
    assert {8, 16, 31, 32} <= set(cells) and not ({3, 7, 15} & set(cells))
    # hp refinement raised the angular order of 7's and 15's children
    assert cells[31].ndofs[0] == cells[8].ndofs[0] + 2
    # spatial hp refinement of 10 and 48 raised the orders of their children
    assert mesh.cols[204].ndofs == [4, 5]
    assert mesh.cols[5].ndofs == [2, 3]
    # angular quadrants are set at the second angular refinement
    for col in mesh.cols.values():
        for cell in col.cells.values():
            q = cell.quadrant
            assert q is not None
            assert q * math.pi / 2 - 1e-12 <= cell.pos[0] < cell.pos[1] <= (q + 1) * math.pi / 2 + 1e-12


@pytest.mark.parametrize("seed", range(6))
def test_random_refinement_matches_model(seed):
    """Seeded random sequences of spatial and angular h / p / hp
    refinements: the product mesh equals the geometric rule model after
    every step, and stays 2:1 balanced with consistent neighbour keys."""
    from dg.mesh import Mesh
    from refine_model import RefineModel, mesh_summary
    rng = np.random.default_rng(seed)
    mesh = Mesh([1.0, 1.0], [False, False], [3, 3, 3], True)
    model = RefineModel(3, 3, 3)
    for _ in range(2):
        mesh.ref_mesh(kind="ang", form="h")
        for key in sorted(model.summary()):
            model.ref_col(key, "ang", "h")
    mesh.ref_mesh(kind="spt", form="h")
    for key in sorted(model.summary()):
        model.ref_col(key, "spt", "h")
    for step in range(14):
        keys = sorted(mesh.cols)
        ck = int(keys[rng.integers(len(keys))])
        form = ["h", "p", "hp"][rng.integers(3)]
        if rng.random() < 0.5 and mesh.cols[ck].lv < 4 and max(mesh.cols[ck].ndofs) < 8:
            mesh.ref_col(ck, kind="spt", form=form)
            model.ref_col(ck, "spt", form)
        else:
            kks = sorted(mesh.cols[ck].cells)
            kk = int(kks[rng.integers(len(kks))])
            if mesh.cols[ck].cells[kk].lv >= 5 or mesh.cols[ck].cells[kk].ndofs[0] >= 8:
                continue
            mesh.ref_cell(ck, kk, form=form)
            model.ref_cell(ck, kk, form)
        assert mesh_summary(mesh) == model.summary(), (seed, step)
    for ck, col in mesh.cols.items():
        geo = meshq.geometric_face_neighbours(mesh, ck)
        assert [sorted({k for k in col.nhbr_keys[F] if k is not None}) for F in range(4)] == geo
        for F in range(4):
            for nk in {k for k in col.nhbr_keys[F] if k is not None}:
                assert abs(mesh.cols[nk].lv - col.lv) <= 1


def test_two_to_one_balance_and_neighbour_symmetry():
    mesh = reference_sequence_mesh()
    for ck, col in mesh.cols.items():
        for F in range(4):
            for nk in {k for k in col.nhbr_keys[F] if k is not None}:
                nb = mesh.cols[nk]
                assert abs(nb.lv - col.lv) <= 1, (ck, nk)
                assert ck in nb.nhbr_keys[(F + 2) % 4], (ck, F, nk)
        geo = meshq.geometric_face_neighbours(mesh, ck)
        stored = [sorted({k for k in col.nhbr_keys[F] if k is not None}) for F in range(4)]
        assert stored == geo, ck
        # angular neighbours within one level inside a column
        for kk, cell in col.cells.items():
            for nk in cell.nhbr_keys:
                if nk is not None and nk in col.cells:
                    assert abs(col.cells[nk].lv - cell.lv) <= 1


def test_ndof_and_offsets():
    from dg.matrix import get_col_dof_offsets
    mesh = reference_sequence_mesh()
    offs, n = meshq.dof_offsets(mesh)
    assert mesh.get_ndof() == n
    coffs = get_col_dof_offsets(mesh)
    for ck, kks in meshq.leaf_order(mesh):
        assert coffs[ck] == offs[(ck, kks[0])]


def test_mesh_json_round_trip(tmp_path):
    from dg.mesh import Mesh, from_file
    mesh = Mesh([1.0, 2.0], [False, False], [2, 3, 4], True)
    p = str(tmp_path / "mesh.json")
    mesh.to_file(p)
    assert from_file(p) == mesh
    mesh = reference_sequence_mesh()
    mesh.to_file(p)
    same = from_file(p)
    assert same == mesh
    assert str(same) == str(mesh)


def test_projection_round_trip(tmp_path):
    from dg.mesh import Mesh
    from dg.projection import Projection, from_file
    mesh = Mesh([1.0, 2.0], [False, False], [3, 3, 3], True)
    for _ in range(2):
        mesh.ref_mesh(kind="all", form="hp")

    def f(x, y, th):
        return np.cos(th) * (x + y)

    proj = Projection(mesh, f)
    pp, mp = str(tmp_path / "proj.npy"), str(tmp_path / "mesh.json")
    proj.to_file(pp, write_mesh=True, mesh_file_path=mp)
    same = from_file(mp, pp)
    assert same == proj
    v = proj.to_vector()
    assert v.size == mesh.get_ndof()
    # vector order: columns, cells, beta = nth ny i + nth j + a
    offs, _ = meshq.dof_offsets(mesh)
    ck, kks = meshq.leaf_order(mesh)[3]
    col = mesh.cols[ck]
    cell = col.cells[kks[1]]
    nx, ny = col.ndofs
    nth = cell.ndofs[0]
    from dg.quadrature import quad_xyth
    xb, _, yb, _, tb, _ = quad_xyth(nx, ny, nth)
    i, j, a = nx - 1, 1, nth - 1
    x = col.pos[0] + (col.pos[2] - col.pos[0]) * (xb[i] + 1) / 2
    y = col.pos[1] + (col.pos[3] - col.pos[1]) * (yb[j] + 1) / 2
    th = cell.pos[0] + (cell.pos[1] - cell.pos[0]) * (tb[a] + 1) / 2
    assert v[offs[(ck, kks[1])] + nth * ny * i + nth * j + a] == pytest.approx(f(x, y, th), abs=1e-15)
    w = Projection(mesh, None).from_vector(2.0 * v)
    assert np.array_equal(w.to_vector(), 2.0 * v)


def test_projection_integrals():
    """col_intg_th of 1 is 2 pi; cell_intg_xy of 1 is the column area."""
    from dg.projection import Projection
    mesh = problems.uniform_mesh([1.0, 2.0], [3, 3, 3], 2, 1)
    proj = Projection(mesh, lambda x, y, th: 1.0 + 0 * x * y * th)
    ck = sorted(mesh.cols)[0]
    assert np.allclose(proj.col_intg_th(ck), 2 * math.pi, atol=1e-13)
    kk = sorted(mesh.cols[ck].cells)[0]
    col = mesh.cols[ck]
    area = (col.pos[2] - col.pos[0]) * (col.pos[3] - col.pos[1])
    assert np.allclose(proj.cell_intg_xy(ck, kk), area, atol=1e-14)


# ------------------------------------------------------------ device tables

def hp_mesh():
    m = problems.uniform_mesh([1.0, 1.0], [3, 3, 3], 2, 1)
    keys = sorted(m.cols)
    m.ref_col(keys[0], "spt", "hp")
    last = sorted(m.cols)[-1]
    m.ref_cell(last, sorted(m.cols[last].cells)[0], "hp")
    first = sorted(m.cols)[0]
    m.ref_cell(first, sorted(m.cols[first].cells)[2], "h")
    return m


@pytest.mark.parametrize("which", ["hp", "ref"])
def test_tables_integer_parity(which):
    """DOF offsets, masked faces and upwind links are bit-exact with the
    oracle's enumeration (outflow rule, neighbour columns, neighbour cells)."""
    from rtx.plan import MeshTables
    mesh = hp_mesh() if which == "hp" else reference_sequence_mesh()
    t = MeshTables(mesh)
    offs, n = meshq.dof_offsets(mesh)
    assert t.ndof == n
    order = [(ck, kk) for ck, kks in meshq.leaf_order(mesh) for kk in kks]
    assert t.cell_keys == order
    assert [int(v) for v in t.cell_dof_off] == [offs[c] for c in order]
    # masked faces reproduce the interior mask
    mask = asm.intr_mask(mesh)
    m2 = np.ones(n, dtype=bool)
    for k, (ck, kk) in enumerate(order):
        col = mesh.cols[ck]
        nx, ny = col.ndofs
        nth = col.cells[kk].ndofs[0]
        blk = np.ones((nx, ny, nth), dtype=bool)
        b = int(t.cell_bdry[k])
        if b & 1: blk[nx - 1] = False
        if b & 2: blk[:, ny - 1] = False
        if b & 4: blk[0] = False
        if b & 8: blk[:, 0] = False
        m2[offs[(ck, kk)]:offs[(ck, kk)] + blk.size] = blk.ravel()
    assert np.array_equal(mask, m2)
    # upwind links
    idx = {c: i for i, c in enumerate(order)}
    expect = []
    for i, (ck, kk) in enumerate(order):
        col = mesh.cols[ck]
        q = col.cells[kk].quadrant
        for F in range(4):
            if q is not None and F in (q, (q + 1) % 4):
                continue
            nbrs = []
            for nk in col.nhbr_keys[F]:
                if nk is not None and nk not in nbrs:
                    nbrs.append(nk)
            for nk in nbrs:
                for k1 in meshq.nhbr_cells(mesh, ck, kk, nk):
                    expect.append((i, idx[(nk, k1)], F))
    got = sorted(zip(t.link_target.tolist(), t.link_src.tolist(), t.link_face.tolist()))
    assert got == sorted(expect)
    assert np.all(np.diff(t.link_target) >= 0)
    assert t.link_start[-1] == t.nlink


def test_tables_face_matrix_pools_match_oracle():
    """Every pooled E^s / E^theta block the device reads equals the oracle's
    branch-by-branch loop evaluation for that link."""
    from rtx.plan import MeshTables
    mesh = hp_mesh()
    t = MeshTables(mesh)
    for l in range(t.nlink):
        k, s, F = int(t.link_target[l]), int(t.link_src[l]), int(t.link_face[l])
        (ck0, kk0), (ck1, kk1) = t.cell_keys[k], t.cell_keys[s]
        c0, c1 = mesh.cols[ck0], mesh.cols[ck1]
        d = 1 if F % 2 == 0 else 0
        Es = emat.E_space(c0.ndofs[d], c0.pos[d], c0.pos[d + 2], c0.lv,
                          c1.ndofs[d], c1.pos[d], c1.pos[d + 2], c1.lv)
        got = t.Es_pool[t.link_Es_off[l]:t.link_Es_off[l] + Es.size].reshape(Es.shape)
        assert np.max(np.abs(got - Es)) < 1e-15, l
        e0, e1 = c0.cells[kk0], c1.cells[kk1]
        Et = emat.E_theta(e0.ndofs[0], e0.pos[0], e0.pos[1], e0.lv,
                          e1.ndofs[0], e1.pos[0], e1.pos[1], e1.lv, F)
        got = t.Eth_pool[t.link_Eth_off[l]:t.link_Eth_off[l] + Et.size].reshape(Et.shape)
        assert np.max(np.abs(got - Et)) < 1e-15, l


def test_tables_reject_unsupported_orders():
    from rtx.plan import MeshTables
    from dg.mesh import Mesh
    with pytest.raises(ValueError):
        MeshTables(Mesh([1.0, 1.0], [False, False], [11, 3, 3], True))
    with pytest.raises(ValueError):
        MeshTables(Mesh([1.0, 1.0], [False, False], [3, 3, 3], False))
