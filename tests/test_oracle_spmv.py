"""The oracle's C CSR product (the multi-core CPU baseline) against scipy.

The C library restates y = A x row by row, left to right, without FMA
contraction, so on the same CSR matrix it must give scipy's result bit for
bit; checked on a random matrix with empty
rows and on the oracle's assembled system of a small hp mesh.
"""
import numpy as np
import pytest
import scipy.sparse as sp

from oracle import build_c, assemble as asm
from rtx import problems


@pytest.fixture(scope="module")
def lib():
    build_c.build()
    return build_c.load()


@pytest.mark.parametrize("nthreads", [1, 3])
def test_random_csr_matches_scipy(lib, nthreads):
    rng = np.random.default_rng(5)
    A = sp.random(500, 300, density=0.05, random_state=rng, format="csr")
    A = sp.vstack([A, sp.csr_matrix((7, 300))]).tocsr()      # empty trailing rows
    x = rng.standard_normal(300)
    y = build_c.csr_spmv(lib, A, x, nthreads)
    ref = A @ x
    assert np.array_equal(y, ref)
    assert (y[-7:] == 0).all()


def test_assembled_system_matches_scipy(lib):
    mesh = problems.uniform_mesh([1.0, 1.0], [3, 3, 3], 2, 2)
    wl = problems.two_scatterers()
    V = asm.assemble(mesh, wl.kappa, wl.sigma, wl.Phi)
    A = asm.system_matrix(V).tocsr()
    x = np.random.default_rng(1).uniform(-1, 1, A.shape[1])
    y = build_c.csr_spmv(lib, A, x, 2)
    ref = A @ x
    assert np.array_equal(y, ref)
    assert lib.oracle_spmv_threads() >= 1


def test_bj_apply_matches_block_jacobi_oracle(lib):
    """The C block-Jacobi apply of the CPU solve against the oracle's
    per-block numpy product, on interior positions of an hp mesh's system
    (blocks of different sizes, grouped by size)."""
    from oracle import system as osys
    from test_gpu_apply import hp_mesh
    mesh = hp_mesh()
    wl = problems.two_scatterers()
    V = asm.assemble(mesh, wl.kappa, wl.sigma, wl.Phi)
    A = asm.system_matrix(V)
    mask = asm.intr_mask(mesh)
    invs = osys.block_jacobi_inverses(A, mesh, mask)
    blocks = osys.bj_block_dofs(mesh)
    r = np.random.default_rng(2).uniform(-1, 1, mask.size)
    ref = osys.block_jacobi_apply(invs, mesh, r)
    I = np.flatnonzero(mask)
    pos = np.full(mask.size, -1, dtype=np.int64)
    pos[I] = np.arange(I.size)
    z = np.zeros(I.size)
    sizes = sorted({len(b) for b in blocks})
    assert len(sizes) > 1 and (~mask).any()
    for bs in sizes:
        sel = [k for k, b in enumerate(blocks) if len(b) == bs]
        idx = np.stack([pos[blocks[k]] for k in sel])
        inv = np.stack([invs[k] for k in sel])
        build_c.bj_apply(lib, idx, inv, r[I], 3, z)
    assert np.linalg.norm(z - ref[I]) / np.linalg.norm(ref[I]) < 1e-14
