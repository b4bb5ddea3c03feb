// L2 error of a nodal solution against values at more nodes, per cell.
//
// src/amr/Error_Indicator/error/error_analytic.py:102-137 interpolates each
// cell's nodal solution (nx, ny, nth) to nx + 3, ny + 3, nth + 3 nodes (LGL
// in space, Gauss-Legendre in angle) with a quintuple loop and sums
// (dx dy dth / 8) w_p w_q w_r (u - u_hr)^2 against the analytic solution;
// error_high_resolution.py:103-203 does the same against the nodal solution
// of a p-refined mesh (mx, my, mth nodes per cell, read per cell).  Here one
// workgroup per cell does the interpolation as three 1-D contractions in LDS
// (angle, then y, then x fused with the weighted squared difference) and a
// block reduction.
#include "rtx_internal.h"

namespace rtx {

constexpr int ENM = RTX_NMAX;       // largest node count
constexpr int EMM = RTX_NMAX + 3;   // largest high-resolution node count
constexpr int ETHREADS = 256;

// NODAL: target node counts from hr_dims[3 k + 0..2] and the (n, m) table
// at tab + (n (EMM + 1) + m) RTX_ERR_TAB_STRIDE; also hr_norm2[k] =
// (dx dy dth / 8) sum w_p w_q w_r u_hr^2.  Otherwise m = n + 3 and the table
// of n at tab + n RTX_ERR_TAB_STRIDE.
template <bool NODAL>
__global__ __launch_bounds__(ETHREADS) void k_error_l2(DevMesh m, const double *__restrict__ uh,
                                                       const double *__restrict__ u_hr,
                                                       const int64_t *__restrict__ hr_off,
                                                       const int32_t *__restrict__ hr_dims,
                                                       const double *__restrict__ tab,
                                                       double *__restrict__ cell_err,
                                                       double *__restrict__ hr_norm2) {
    __shared__ double u[ENM * ENM * ENM];
    __shared__ double t1[ENM * ENM * EMM];      // (i, j, r)
    __shared__ double t2[ENM * EMM * EMM];      // (i, q, r)
    __shared__ double red[ETHREADS];
    const int k = blockIdx.x, tid = threadIdx.x;
    const int c = m.cell_col[k];
    const int nx = m.col_nx[c], ny = m.col_ny[c], nt = m.cell_nth[k];
    const int mx = NODAL ? hr_dims[3 * k + 0] : nx + 3;
    const int my = NODAL ? hr_dims[3 * k + 1] : ny + 3;
    const int mt = NODAL ? hr_dims[3 * k + 2] : nt + 3;
    auto tab_of = [&](int n, int mm) {
        return tab + (NODAL ? size_t(n) * (EMM + 1) + mm : size_t(n)) * RTX_ERR_TAB_STRIDE;
    };
    const double *Px = tab_of(nx, mx);                       // LGL nx -> mx
    const double *Py = tab_of(ny, my);                       // LGL ny -> my
    const double *Pt = tab_of(nt, mt) + ENM * EMM;           // LG nt -> mt
    const double *wx = tab_of(nx, mx) + 2 * ENM * EMM;       // LGL mx weights
    const double *wy = tab_of(ny, my) + 2 * ENM * EMM;
    const double *wt = tab_of(nt, mt) + 2 * ENM * EMM + EMM; // LG mt weights
    const int64_t off = m.cell_dof_off[k];
    const int n = nx * ny * nt;
    for (int t = tid; t < n; t += ETHREADS) u[t] = uh[off + t];
    __syncthreads();
    // angle: t1[i][j][r] = sum_a u[i][j][a] Pt[a][r]
    for (int e = tid; e < nx * ny * mt; e += ETHREADS) {
        const int ij = e / mt, r = e % mt;
        double s = 0.0;
        for (int a = 0; a < nt; ++a) s += u[ij * nt + a] * Pt[a * mt + r];
        t1[e] = s;
    }
    __syncthreads();
    // y: t2[i][q][r] = sum_j t1[i][j][r] Py[j][q]
    for (int e = tid; e < nx * my * mt; e += ETHREADS) {
        const int i = e / (my * mt), q = (e / mt) % my, r = e % mt;
        double s = 0.0;
        for (int j = 0; j < ny; ++j) s += t1[(i * ny + j) * mt + r] * Py[j * my + q];
        t2[e] = s;
    }
    __syncthreads();
    // x, fused with the weighted squared difference
    const double *uk = u_hr + hr_off[k];
    double acc = 0.0, acc2 = 0.0;
    for (int e = tid; e < mx * my * mt; e += ETHREADS) {
        const int p = e / (my * mt), q = (e / mt) % my, r = e % mt;
        double s = 0.0;
        for (int i = 0; i < nx; ++i) s += t2[(i * my + q) * mt + r] * Px[i * mx + p];
        const double d = uk[e] - s;
        acc += wx[p] * wy[q] * wt[r] * d * d;
        if (NODAL) acc2 += wx[p] * wy[q] * wt[r] * uk[e] * uk[e];
    }
    const double dx = m.col_geom[4 * c + 2] - m.col_geom[4 * c + 0];
    const double dy = m.col_geom[4 * c + 3] - m.col_geom[4 * c + 1];
    const double dth = m.cell_geom[2 * k + 1] - m.cell_geom[2 * k + 0];
    red[tid] = acc;
    __syncthreads();
    for (int s = ETHREADS / 2; s > 0; s >>= 1) {
        if (tid < s) red[tid] += red[tid + s];
        __syncthreads();
    }
    if (tid == 0) cell_err[k] = dx * dy * dth / 8.0 * red[0];
    if (NODAL) {
        __syncthreads();
        red[tid] = acc2;
        __syncthreads();
        for (int s = ETHREADS / 2; s > 0; s >>= 1) {
            if (tid < s) red[tid] += red[tid + s];
            __syncthreads();
        }
        if (tid == 0) hr_norm2[k] = dx * dy * dth / 8.0 * red[0];
    }
}

}  // namespace rtx

using namespace rtx;

extern "C" int rtx_error_l2(rtx_plan *p, const double *d_uh, const double *d_u_hr,
                            const int64_t *d_hr_off, const double *d_tab, double *d_cell_err,
                            void *stream) {
    RTX_CHECK(p && d_uh && d_u_hr && d_hr_off && d_tab && d_cell_err,
              "rtx_error_l2: null argument");
    RTX_CHECK(p->max_cell_ndof <= ENM * ENM * ENM, "rtx_error_l2: cell too large");
    if (p->m.ncell == 0) return RTX_OK;
    k_error_l2<false><<<p->m.ncell, ETHREADS, 0, static_cast<hipStream_t>(stream)>>>(
        p->m, d_uh, d_u_hr, d_hr_off, nullptr, d_tab, d_cell_err, nullptr);
    RTX_HIP(hipGetLastError());
    return RTX_OK;
}

extern "C" int rtx_error_l2_nodal(rtx_plan *p, const double *d_uh, const double *d_u_hr,
                                  const int64_t *d_hr_off, const int32_t *d_hr_dims,
                                  const double *d_tab, double *d_cell_err, double *d_hr_norm2,
                                  void *stream) {
    RTX_CHECK(p && d_uh && d_u_hr && d_hr_off && d_hr_dims && d_tab && d_cell_err && d_hr_norm2,
              "rtx_error_l2_nodal: null argument");
    RTX_CHECK(p->max_cell_ndof <= ENM * ENM * ENM, "rtx_error_l2_nodal: cell too large");
    if (p->m.ncell == 0) return RTX_OK;
    k_error_l2<true><<<p->m.ncell, ETHREADS, 0, static_cast<hipStream_t>(stream)>>>(
        p->m, d_uh, d_u_hr, d_hr_off, d_hr_dims, d_tab, d_cell_err, d_hr_norm2);
    RTX_HIP(hipGetLastError());
    return RTX_OK;
}
