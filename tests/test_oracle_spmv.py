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
