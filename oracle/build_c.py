"""Build recipe for the oracle's C restatement (``csr_spmv.c``).

Compiles ``oracle/csr_spmv.c`` with gcc and OpenMP into
``oracle/liboracle_spmv.so`` (git-ignored, like every built library).  Test
infrastructure only: the library is the multi-core CPU baseline that
``bench.py`` times beside the GPU apply (CSR product) and the GPU solve (CSR
product and block-Jacobi apply inside scipy's GMRES).
"""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "csr_spmv.c")
OUT = os.path.join(HERE, "liboracle_spmv.so")


def build(verbose=False):
    # no FMA contraction: row sums round like scipy's csr_matvec
    # a per-process temporary name: parallel test workers may build at once
    tmp = "{}.{}.tmp".format(OUT, os.getpid())
    cmd = ["gcc", "-O3", "-ffp-contract=off", "-fopenmp", "-fPIC", "-shared", SRC, "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("gcc failed for {}:\n{}\n{}".format(SRC, r.stdout, r.stderr))
    os.replace(tmp, OUT)
    return OUT


def load():
    """ctypes handle of the built library (raises if it was not built)."""
    import ctypes as C
    if not os.path.exists(OUT):
        raise OSError("{} is not built (python oracle/build_c.py)".format(OUT))
    lib = C.CDLL(OUT)
    lib.oracle_spmv_threads.restype = C.c_int
    lib.oracle_csr_spmv.argtypes = [C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                    C.c_void_p, C.c_int]
    lib.oracle_csr_spmv.restype = None
    lib.oracle_bj_apply.argtypes = [C.c_int64, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p,
                                    C.c_void_p, C.c_int]
    lib.oracle_bj_apply.restype = None
    return lib


def csr_spmv(lib, A, x, nthreads):
    """y = A x by the C library (A: scipy CSR, int32 indices; x: float64)."""
    import numpy as np
    indptr = np.ascontiguousarray(A.indptr, dtype=np.int64)
    indices = np.ascontiguousarray(A.indices, dtype=np.int32)
    data = np.ascontiguousarray(A.data, dtype=np.float64)
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.empty(A.shape[0])
    lib.oracle_csr_spmv(A.shape[0], indptr.ctypes.data, indices.ctypes.data, data.ctypes.data,
                        x.ctypes.data, y.ctypes.data, int(nthreads))
    return y


def bj_apply(lib, idx, inv, r, nthreads, z=None):
    """z = blockdiag(inv) r by the C library.  idx: int64 (nblocks, bs)
    positions in r (-1: a boundary DOF, left out); inv: float64 (nblocks,
    bs, bs).  Entries of z that no block names keep their value (zero when
    z is not given)."""
    import numpy as np
    idx = np.ascontiguousarray(idx, dtype=np.int64)
    inv = np.ascontiguousarray(inv, dtype=np.float64)
    r = np.ascontiguousarray(r, dtype=np.float64)
    if z is None:
        z = np.zeros_like(r)
    nb, bs = idx.shape
    assert inv.shape == (nb, bs, bs) and idx.max(initial=-1) < r.size and z.size == r.size
    lib.oracle_bj_apply(nb, bs, idx.ctypes.data, inv.ctypes.data, r.ctypes.data,
                        z.ctypes.data, int(nthreads))
    return z


if __name__ == "__main__":
    print(build(verbose=True))
