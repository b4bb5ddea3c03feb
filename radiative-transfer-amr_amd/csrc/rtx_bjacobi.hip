// Per-cell block-Jacobi preconditioner.
//
// The reference solves with PETSc bjacobi (src/rt/solve.py:93-95; one block
// per MPI rank with an ILU(0) sub-solver) and computes, but discards, a dense
// inverse of each column block (src/rt/preconditioner_matrix.py:259-262).
// Here the blocks are the angular cells: the cell's diagonal block of the
// interior system A_II (mass, interior convection, outflow faces, in-cell
// scattering and any self-coupling link), identity on inflow-boundary DOFs.
// Each block is assembled on the device, inverted by Gauss-Jordan with
// partial pivoting (in LDS when it fits, else in place in HBM), and stored
// column-major so the apply is a coalesced batched GEMV.
#include "rtx_internal.h"

namespace rtx {

constexpr int NMq = RTX_NMAX;
constexpr int BJ_THREADS = 256;

__device__ inline bool bj_on_bdry(int bd, int i, int j, int nx, int ny) {
    return ((bd & 1) && i == nx - 1) || ((bd & 2) && j == ny - 1) || ((bd & 4) && i == 0) ||
           ((bd & 8) && j == 0);
}

// Entry (row (p, q, r), column (i, j, a)) of the cell's diagonal block.
__device__ double block_entry(const DevMesh &m, int k, int row, int col, const double *cth,
                              const double *sth) {
    const int c = m.cell_col[k];
    const int nx = m.col_nx[c], ny = m.col_ny[c], nth = m.cell_nth[k];
    const int nyt = ny * nth;
    const int p = row / nyt, q = (row / nth) % ny, r = row % nth;
    const int i = col / nyt, j = (col / nth) % ny, a = col % nth;
    const double x0 = m.col_geom[4 * c + 0], y0 = m.col_geom[4 * c + 1];
    const double dx = m.col_geom[4 * c + 2] - x0, dy = m.col_geom[4 * c + 3] - y0;
    const double th0 = m.cell_geom[2 * k + 0], dth = m.cell_geom[2 * k + 1] - th0;
    const double *wx = m.lgl_w + nx * NMq, *wy = m.lgl_w + ny * NMq, *wt = m.lg_w + nth * NMq;
    double v = 0.0;
    if (p == i && q == j) {
        // in-cell scattering: -dx dy w_p w_q sigma G'[(K, r), (K, a)]
        const int ao = m.cell_ang_off[k], na = m.col_nang[c];
        const double g = m.G[m.col_G_off[c] + size_t(ao + a) * na + (ao + r)];
        v -= dx * dy * wx[p] * wy[q] * m.sigma[m.col_sp_off[c] + p * ny + q] * g;
        if (r == a) {
            v += (dx * dy * dth * 0.125) * wx[p] * wy[q] * wt[r] *
                 m.kappa[m.col_sp_off[c] + p * ny + q];
            const int quad = m.cell_quad[k];
            if (quad >= 0) {
                for (int h = 0; h < 2; ++h) {
                    const int F = (quad + h) & 3;
                    bool on = (F == 0 && p == nx - 1) || (F == 2 && p == 0) ||
                              (F == 1 && q == ny - 1) || (F == 3 && q == 0);
                    if (on) {
                        const double tf =
                            F == 0 ? cth[r] : (F == 1 ? sth[r] : (F == 2 ? -cth[r] : -sth[r]));
                        v += ((F & 1) ? dx : dy) * dth * 0.25 * ((F & 1) ? wx[p] : wy[q]) *
                             wt[r] * tf;
                    }
                }
            }
        }
    }
    if (r == a) {
        // -IC: x-derivative couples (i, q) -> (p, q); y-derivative (p, j) -> (p, q)
        if (q == j)
            v -= (dy * dth * 0.25) * wy[q] * wt[r] * cth[r] * wx[i] *
                 m.lgl_D[(size_t(nx) * NMq + p) * NMq + i];
        if (p == i)
            v -= (dx * dth * 0.25) * wx[p] * wt[r] * sth[r] * wy[j] *
                 m.lgl_D[(size_t(ny) * NMq + q) * NMq + j];
    }
    // a cell coupled to itself through a periodic face
    for (int l = m.link_start[k]; l < m.link_start[k + 1]; ++l) {
        if (m.link_src[l] != k) continue;
        const int F = m.link_face[l];
        int m0, m1;
        bool row_on, col_on;
        if (F == 0) { row_on = p == nx - 1; col_on = i == 0; m0 = q; m1 = j; }
        else if (F == 2) { row_on = p == 0; col_on = i == nx - 1; m0 = q; m1 = j; }
        else if (F == 1) { row_on = q == ny - 1; col_on = j == 0; m0 = p; m1 = i; }
        else { row_on = q == 0; col_on = j == ny - 1; m0 = p; m1 = i; }
        if (!(row_on && col_on)) continue;
        const int n0 = (F & 1) ? nx : ny;
        v += ((F & 1) ? dx : dy) * dth * 0.25 * m.Es_pool[m.link_Es_off[l] + m1 * n0 + m0] *
             m.Eth_pool[m.link_Eth_off[l] + a * nth + r];
    }
    return v;
}

// Gauss-Jordan inversion with partial pivoting of the n x n row-major matrix
// A (LDS or global), in place.  Returns nonzero in *bad on a zero pivot.
__device__ void gauss_jordan(double *A, int n, int *perm, double *red_v, int *red_i,
                             int *bad) {
    const int tid = threadIdx.x, nt = blockDim.x;
    for (int k = 0; k < n; ++k) {
        // pivot search in column k, rows k..n-1
        double best = -1.0;
        int bi = k;
        for (int i = k + tid; i < n; i += nt) {
            double v = fabs(A[size_t(i) * n + k]);
            if (v > best) { best = v; bi = i; }
        }
        red_v[tid] = best;
        red_i[tid] = bi;
        __syncthreads();
        for (int s = nt / 2; s > 0; s >>= 1) {
            if (tid < s) {
                double o = red_v[tid + s];
                int oi = red_i[tid + s];
                if (o > red_v[tid] || (o == red_v[tid] && oi < red_i[tid])) {
                    red_v[tid] = o;
                    red_i[tid] = oi;
                }
            }
            __syncthreads();
        }
        const int pr = red_i[0];
        const bool zero = !(red_v[0] > 0.0);
        __syncthreads();
        if (zero) {
            if (tid == 0) *bad = 1;
            return;
        }
        if (tid == 0) perm[k] = pr;
        if (pr != k)
            for (int c = tid; c < n; c += nt) {
                double t = A[size_t(k) * n + c];
                A[size_t(k) * n + c] = A[size_t(pr) * n + c];
                A[size_t(pr) * n + c] = t;
            }
        __syncthreads();
        const double inv = 1.0 / A[size_t(k) * n + k];
        __syncthreads();
        for (int c = tid; c < n; c += nt)
            A[size_t(k) * n + c] = (c == k) ? inv : A[size_t(k) * n + c] * inv;
        __syncthreads();
        // eliminate column k from every other row
        for (int e = tid; e < n * n; e += nt) {
            const int i = e / n, c = e % n;
            if (i == k) continue;
            const double f = A[size_t(i) * n + k];
            if (c == k) continue;
            A[size_t(i) * n + c] -= f * A[size_t(k) * n + c];
        }
        __syncthreads();
        for (int i = tid; i < n; i += nt)
            if (i != k) A[size_t(i) * n + k] = -A[size_t(i) * n + k] * inv;
        __syncthreads();
    }
    // undo the row interchanges as column interchanges, in reverse order
    for (int k = n - 1; k >= 0; --k) {
        const int pr = perm[k];
        if (pr != k)
            for (int i = tid; i < n; i += nt) {
                double t = A[size_t(i) * n + k];
                A[size_t(i) * n + k] = A[size_t(i) * n + pr];
                A[size_t(i) * n + pr] = t;
            }
        __syncthreads();
    }
}

// One workgroup per cell.  USE_LDS: factor in dynamic LDS, then write the
// inverse transposed (column-major) to out; otherwise factor in place in a
// global scratch block and transpose into out.
template <bool USE_LDS>
__global__ __launch_bounds__(BJ_THREADS) void k_bj_build(DevMesh m, int restrict_, const int64_t *inv_off,
                                                        double *__restrict__ out,
                                                        double *__restrict__ scratch,
                                                        int *__restrict__ status) {
    extern __shared__ double lds[];
    __shared__ double cth[NMq], sth[NMq];
    __shared__ double red_v[BJ_THREADS];
    __shared__ int red_i[BJ_THREADS];
    __shared__ int perm[RTX_NMAX * RTX_NMAX * RTX_NMAX];
    __shared__ int bad;
    const int k = blockIdx.x, tid = threadIdx.x;
    const int c = m.cell_col[k];
    const int nx = m.col_nx[c], ny = m.col_ny[c], nth = m.cell_nth[k];
    const int n = nx * ny * nth, nyt = ny * nth;
    const int bd = restrict_ ? m.cell_bdry[k] : 0;
    if (tid == 0) bad = 0;
    if (tid < nth) {
        const double th0 = m.cell_geom[2 * k], dth = m.cell_geom[2 * k + 1] - th0;
        double s, cc;
        sincos(th0 + dth * 0.5 * (m.lg_x[nth * NMq + tid] + 1.0), &s, &cc);
        cth[tid] = cc;
        sth[tid] = s;
    }
    __syncthreads();
    double *A = USE_LDS ? lds : scratch + inv_off[k];
    for (int e = tid; e < n * n; e += BJ_THREADS) {
        const int row = e / n, col = e % n;
        const bool rb = bd && bj_on_bdry(bd, row / nyt, (row / nth) % ny, nx, ny);
        const bool cb = bd && bj_on_bdry(bd, col / nyt, (col / nth) % ny, nx, ny);
        double v;
        if (rb || cb) v = (row == col) ? 1.0 : 0.0;
        else v = block_entry(m, k, row, col, cth, sth);
        A[e] = v;
    }
    __syncthreads();
    gauss_jordan(A, n, perm, red_v, red_i, &bad);
    __syncthreads();
    if (bad) {
        if (tid == 0) atomicMax(status, k + 1);
        return;
    }
    double *o = out + inv_off[k];
    if (USE_LDS) {
        for (int e = tid; e < n * n; e += BJ_THREADS) {
            const int col = e / n, row = e % n;     // o is column-major
            o[e] = A[size_t(row) * n + col];
        }
    } else {
        // out and scratch are distinct buffers
        for (int e = tid; e < n * n; e += BJ_THREADS) {
            const int col = e / n, row = e % n;
            o[e] = A[size_t(row) * n + col];
        }
    }
}

__global__ void k_to_f32_bj(const double *__restrict__ s, float *__restrict__ d, int64_t n) {
    int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    for (; i < n; i += int64_t(gridDim.x) * blockDim.x) d[i] = float(s[i]);
}

// z_K = inv_K r_K, inv column-major.  One workgroup per cell.
template <typename T, typename M>
__global__ __launch_bounds__(BJ_THREADS) void k_bj_apply(DevMesh m, const int64_t *inv_off,
                                                        const M *__restrict__ inv,
                                                        const T *__restrict__ r,
                                                        T *__restrict__ z) {
    __shared__ double rs[RTX_NMAX * RTX_NMAX * RTX_NMAX];
    const int k = blockIdx.x, tid = threadIdx.x;
    const int c = m.cell_col[k];
    const int n = m.col_nx[c] * m.col_ny[c] * m.cell_nth[k];
    const int off = m.cell_dof_off[k];
    for (int i = tid; i < n; i += BJ_THREADS) rs[i] = double(r[off + i]);
    __syncthreads();
    const M *B = inv + inv_off[k];
    for (int i = tid; i < n; i += BJ_THREADS) {
        double acc = 0.0;
        for (int j = 0; j < n; ++j) acc += double(B[size_t(j) * n + i]) * rs[j];
        z[off + i] = T(acc);
    }
}

}  // namespace rtx

using namespace rtx;

extern "C" int rtx_bjacobi_size(rtx_plan *p, int64_t *n_entries) {
    RTX_CHECK(p && n_entries, "rtx_bjacobi_size: null argument");
    std::vector<int32_t> nx(p->m.ncol), ny(p->m.ncol), cc(p->m.ncell), nt(p->m.ncell);
    RTX_HIP(hipMemcpy(nx.data(), p->m.col_nx, nx.size() * 4, hipMemcpyDeviceToHost));
    RTX_HIP(hipMemcpy(ny.data(), p->m.col_ny, ny.size() * 4, hipMemcpyDeviceToHost));
    RTX_HIP(hipMemcpy(cc.data(), p->m.cell_col, cc.size() * 4, hipMemcpyDeviceToHost));
    RTX_HIP(hipMemcpy(nt.data(), p->m.cell_nth, nt.size() * 4, hipMemcpyDeviceToHost));
    std::vector<int64_t> off(p->m.ncell + 1, 0);
    for (int k = 0; k < p->m.ncell; ++k) {
        int64_t n = int64_t(nx[cc[k]]) * ny[cc[k]] * nt[k];
        off[k + 1] = off[k] + n * n;
    }
    if (!p->bj_off) {
        RTX_HIP(hipMalloc(&p->bj_off, off.size() * sizeof(int64_t)));
        RTX_HIP(hipMemcpy(p->bj_off, off.data(), off.size() * sizeof(int64_t),
                          hipMemcpyHostToDevice));
    }
    p->bj_entries = off.back();
    *n_entries = off.back();
    return RTX_OK;
}

extern "C" int rtx_bjacobi_factor(rtx_plan *p, int restrict_interior, double *d_inv,
                                  float *d_inv_f32, double *d_scratch, void *stream) {
    RTX_CHECK(p && d_inv, "rtx_bjacobi_factor: null argument");
    if (!p->has_coeffs) {
        set_error("rtx_bjacobi_factor: coefficients not set");
        return RTX_ERR_STATE;
    }
    if (!p->bj_off) {
        int64_t n;
        int rc = rtx_bjacobi_size(p, &n);
        if (rc) return rc;
    }
    hipStream_t st = static_cast<hipStream_t>(stream);
    const int64_t nmax = p->max_cell_ndof;
    const size_t lds = size_t(nmax) * nmax * sizeof(double);
    int dev = 0, max_lds = 0;
    RTX_HIP(hipGetDevice(&dev));
    RTX_HIP(hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
    const size_t static_lds = 2 * NMq * 8 + BJ_THREADS * 12 + RTX_NMAX * RTX_NMAX * RTX_NMAX * 4 + 16;
    int *d_status = nullptr;
    RTX_HIP(hipMallocAsync((void **)&d_status, sizeof(int), st));
    RTX_HIP(hipMemsetAsync(d_status, 0, sizeof(int), st));
    if (lds + static_lds <= size_t(max_lds)) {
        RTX_HIP(hipFuncSetAttribute((const void *)k_bj_build<true>,
                                    hipFuncAttributeMaxDynamicSharedMemorySize, int(lds)));
        k_bj_build<true><<<p->m.ncell, BJ_THREADS, lds, st>>>(p->m, restrict_interior, p->bj_off,
                                                            d_inv, nullptr, d_status);
    } else {
        RTX_CHECK(d_scratch != nullptr,
                  "rtx_bjacobi_factor: blocks of %lld DOFs do not fit LDS; a scratch buffer of "
                  "the same size as the inverses is required",
                  (long long)nmax);
        k_bj_build<false><<<p->m.ncell, BJ_THREADS, 0, st>>>(p->m, restrict_interior, p->bj_off,
                                                           d_inv, d_scratch, d_status);
    }
    RTX_HIP(hipGetLastError());
    if (d_inv_f32)
        k_to_f32_bj<<<2048, 256, 0, st>>>(d_inv, d_inv_f32, p->bj_entries);
    int h_status = 0;
    RTX_HIP(hipMemcpyAsync(&h_status, d_status, sizeof(int), hipMemcpyDeviceToHost, st));
    RTX_HIP(hipStreamSynchronize(st));
    RTX_HIP(hipFreeAsync(d_status, st));
    if (h_status) {
        set_error("rtx_bjacobi_factor: singular diagonal block in cell %d", h_status - 1);
        return RTX_ERR_NUMERIC;
    }
    p->bj_ready = true;
    return RTX_OK;
}

extern "C" int rtx_bjacobi_apply(rtx_plan *p, int precision, const void *d_inv, const void *d_r,
                                 void *d_z, void *stream) {
    RTX_CHECK(p && d_inv && d_r && d_z, "rtx_bjacobi_apply: null argument");
    RTX_CHECK(d_r != d_z, "rtx_bjacobi_apply: r and z must not alias");
    if (!p->bj_ready) {
        set_error("rtx_bjacobi_apply: call rtx_bjacobi_factor first");
        return RTX_ERR_STATE;
    }
    hipStream_t st = static_cast<hipStream_t>(stream);
    if (precision == RTX_PREC_F64)
        k_bj_apply<double, double><<<p->m.ncell, BJ_THREADS, 0, st>>>(
            p->m, p->bj_off, (const double *)d_inv, (const double *)d_r, (double *)d_z);
    else if (precision == RTX_PREC_F32)
        k_bj_apply<float, float><<<p->m.ncell, BJ_THREADS, 0, st>>>(
            p->m, p->bj_off, (const float *)d_inv, (const float *)d_r, (float *)d_z);
    else {
        set_error("rtx_bjacobi_apply: unsupported precision %d", precision);
        return RTX_ERR_ARG;
    }
    RTX_HIP(hipGetLastError());
    return RTX_OK;
}
