/* CPU baseline: y = A x for the oracle's assembled FP64 CSR matrix, one row per
 * OpenMP iteration, and the per-cell block-Jacobi apply of the CPU solve.  Test
 * infrastructure only (timed beside the GPU apply and solve); the product never
 * links it.
 *
 * The matrix is the reference's assembled system A = (BC - IC) + M + (-S)
 * (src/rt/solve.py:49-52) in scipy CSR form; PETSc's MatMult on an AIJ matrix
 * is the same row-by-row product (src/rt/solve.py:81-111 drives it through
 * KSP).  Row sums are accumulated left to right in FP64, as scipy's csr_matvec
 * does, so the result is bit-identical to scipy's for the same matrix.
 */
#include <stdint.h>
#include <stdlib.h>
#include <omp.h>

int oracle_spmv_threads(void) { return omp_get_max_threads(); }

void oracle_csr_spmv(int64_t nrows, const int64_t *indptr, const int32_t *indices,
                     const double *data, const double *x, double *y, int nthreads) {
#pragma omp parallel for schedule(static) num_threads(nthreads)
    for (int64_t i = 0; i < nrows; ++i) {
        double s = 0.0;
        for (int64_t k = indptr[i]; k < indptr[i + 1]; ++k) s += data[k] * x[indices[k]];
        y[i] = s;
    }
}

/* z = blockdiag(inv) r over interior DOFs only: block b has bs entries
 * idx[b bs + i] (the DOF's position in r and z, or -1 for a boundary DOF, which
 * the block's inverse leaves uncoupled) and the dense row-major inverse
 * inv[b bs bs ...].  The oracle's block_jacobi_apply on the interior DOFs (its
 * inverses have identity rows and columns at boundary DOFs), summed left to
 * right, one block per OpenMP iteration. */
void oracle_bj_apply(int64_t nblocks, int bs, const int64_t *idx, const double *inv,
                     const double *r, double *z, int nthreads) {
#pragma omp parallel num_threads(nthreads)
    {
        double *rb = (double *)malloc(sizeof(double) * (size_t)bs);
#pragma omp for schedule(static)
        for (int64_t b = 0; b < nblocks; ++b) {
            const int64_t *ix = idx + b * bs;
            const double *M = inv + b * (int64_t)bs * bs;
            /* the block vector, zero at boundary DOFs (their columns of the
             * inverse multiply nothing) */
            for (int j = 0; j < bs; ++j) rb[j] = ix[j] >= 0 ? r[ix[j]] : 0.0;
            /* four rows at a time: four independent left-to-right sums */
            for (int i0 = 0; i0 < bs; i0 += 4) {
                const int nr = bs - i0 < 4 ? bs - i0 : 4;
                double s[4] = {0.0, 0.0, 0.0, 0.0};
                const double *m0 = M + (int64_t)i0 * bs;
                for (int j = 0; j < bs; ++j)
                    for (int k = 0; k < nr; ++k) s[k] += m0[(int64_t)k * bs + j] * rb[j];
                for (int k = 0; k < nr; ++k)
                    if (ix[i0 + k] >= 0) z[ix[i0 + k]] = s[k];
            }
        }
        free(rb);
    }
}
