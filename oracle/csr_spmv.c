/* CPU baseline: y = A x for the oracle's assembled FP64 CSR matrix, one row per
 * OpenMP iteration.  Test infrastructure only (timed beside the GPU apply); the
 * product never links it.
 *
 * The matrix is the reference's assembled system A = (BC - IC) + M + (-S)
 * (src/rt/solve.py:49-52) in scipy CSR form; PETSc's MatMult on an AIJ matrix
 * is the same row-by-row product (src/rt/solve.py:81-111 drives it through
 * KSP).  Row sums are accumulated left to right in FP64, as scipy's csr_matvec
 * does, so the result is bit-identical to scipy's for the same matrix.
 */
#include <stdint.h>
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
