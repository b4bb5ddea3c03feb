"""The C-ABI library: loads, exports every declared symbol, and its ctypes
mirror of rtx_mesh_desc matches the C layout.  Table validation of
rtx_plan_create runs before any HIP call, so it is exercised on the CPU."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from rtx import _lib, problems
from rtx.plan import MeshTables
from rtx.operator import build_desc

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_symbols():
    src = open(os.path.join(ROOT, "include", "rtx.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(rtx_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    names = _declared_symbols()
    assert len(names) >= 15
    for name in names:
        assert hasattr(lib, name), name
    bound = {n for n, _, _ in _lib.SIGNATURES}
    assert set(names) == bound, set(names) ^ bound


def test_mesh_desc_layout_matches_c():
    lib = _lib.load()
    n = lib.rtx_mesh_desc_layout(None, 0)
    offs = (C.c_int64 * n)()
    lib.rtx_mesh_desc_layout(offs, n)
    fields = [f[0] for f in _lib.MeshDesc._fields_]
    assert n == len(fields) + 1
    for i, name in enumerate(fields):
        assert getattr(_lib.MeshDesc, name).offset == offs[i], name
    assert C.sizeof(_lib.MeshDesc) == offs[n - 1]


def _create(T, G_off):
    lib = _lib.load()
    d, keep = build_desc(T, G_off)
    h = C.c_void_p()
    rc = lib.rtx_plan_create(C.byref(d), C.byref(h))
    msg = lib.rtx_last_error().decode()
    if rc == 0:
        lib.rtx_plan_destroy(h)
    return rc, msg, d, keep


def _tables(case):
    if case == "c2":
        mesh, wl = problems.uniform_mesh([1, 1], [5, 5, 5], 4, 5), problems.absorber_c2()
    else:
        from test_gpu_apply import hp_mesh
        mesh, wl = hp_mesh(), problems.two_scatterers()
    T = MeshTables(mesh)
    return T, T.G_pool(wl.Phi)[1]


@pytest.mark.parametrize("case", ["c2", "hp"])
def test_plan_tables_pass_validation(case):
    T, G_off = _tables(case)
    rc, msg, _, _ = _create(T, G_off)
    # no device here: validation passes and the first HIP call reports it
    assert rc == 0 or (rc == -2 and "hipGetDevice" in msg), (rc, msg)


def test_plan_rejects_corrupt_tables():
    T, G_off = _tables("hp")
    lm = T.link_meta.copy()
    lm[4 * 3 + 2] += 1           # wrong trace offset of link 3
    T.link_meta = lm
    rc, msg, _, _ = _create(T, G_off)
    assert rc == -1 and "link_meta" in msg, (rc, msg)
    T, G_off = _tables("hp")
    T.link_src = T.link_src.copy()
    T.link_src[0] = T.ncell      # out-of-range source cell
    rc, msg, _, _ = _create(T, G_off)
    assert rc == -1, (rc, msg)
    T, G_off = _tables("hp")
    T.cell_nth = T.cell_nth.copy()
    T.cell_nth[0] = _lib.NMAX + 1
    rc, msg, _, _ = _create(T, G_off)
    assert rc == -1 and "nth" in msg, (rc, msg)
