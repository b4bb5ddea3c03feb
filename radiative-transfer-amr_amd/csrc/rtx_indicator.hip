// Error indicators of a nodal solution and the integrals hp steering reads.
//
// src/amr/Error_Indicator/error/error_angular_jump.py:14-99 (jumps of the
// angular traces between neighbouring cells of a column),
// error_spatial_jump.py:17-236 (jumps of the spatial face traces between
// neighbouring columns, projected onto a common node set) and the
// projection integrals of src/dg/projection/cell_intg_xy.py:13-36 and
// col_intg_th.py:13-37 that cell_hp_steer.py / col_hp_steer.py read.  The
// reference walks every cell in Python; here one workgroup owns a cell (or a
// column), threads own spatial nodes (or projected face nodes) and a block
// reduction forms each sum.
#include "rtx_internal.h"

namespace rtx {

constexpr int INM = RTX_NMAX;
constexpr int ITHREADS = 128;

__device__ inline double block_sum(double v, double *red) {
    const int tid = threadIdx.x;
    red[tid] = v;
    __syncthreads();
    for (int s = ITHREADS / 2; s > 0; s >>= 1) {
        if (tid < s) red[tid] += red[tid + s];
        __syncthreads();
    }
    const double r = red[0];
    __syncthreads();
    return r;
}

// Per cell: intg_xy[a] = (dx dy / 4) sum_ij wx_i wy_j u[i, j, a] at the
// cell's global angle index (cell_intg_xy.py:21-34), and, when norm2 is
// given, (dx dy / 4)(dth / 2) sum wx wy wth u^2.
__global__ __launch_bounds__(ITHREADS) void k_cell_moments(DevMesh m, const double *__restrict__ u,
                                                          double *__restrict__ intg_xy,
                                                          double *__restrict__ norm2) {
    __shared__ double red[ITHREADS];
    const int k = blockIdx.x, tid = threadIdx.x;
    const int c = m.cell_col[k];
    const int nx = m.col_nx[c], ny = m.col_ny[c], nt = m.cell_nth[k];
    const double dx = m.col_geom[4 * c + 2] - m.col_geom[4 * c + 0];
    const double dy = m.col_geom[4 * c + 3] - m.col_geom[4 * c + 1];
    const double dth = m.cell_geom[2 * k + 1] - m.cell_geom[2 * k + 0];
    const double *wx = m.lgl_w + nx * INM, *wy = m.lgl_w + ny * INM, *wt = m.lg_w + nt * INM;
    const double *uk = u + m.cell_dof_off[k];
    if (tid < nt) {
        double s = 0.0;
        for (int i = 0; i < nx; ++i)
            for (int j = 0; j < ny; ++j) s += wx[i] * wy[j] * uk[(i * ny + j) * nt + tid];
        intg_xy[m.col_ang_start[c] + m.cell_ang_off[k] + tid] = dx * dy / 4.0 * s;
    }
    if (norm2) {
        double acc = 0.0;
        for (int e = tid; e < nx * ny * nt; e += ITHREADS) {
            const int i = e / (ny * nt), j = (e / nt) % ny, a = e % nt;
            acc += wx[i] * wy[j] * wt[a] * uk[e] * uk[e];
        }
        acc = block_sum(acc, red);
        if (tid == 0) norm2[k] = dx * dy / 4.0 * (dth / 2.0) * acc;
    }
}

// Per column: intg_th[i, j] = sum over cells (dth / 2) sum_a wth_a u[i, j, a]
// (col_intg_th.py:22-35), at the column's spatial offset.
__global__ __launch_bounds__(ITHREADS) void k_col_intg_th(DevMesh m, const double *__restrict__ u,
                                                         double *__restrict__ intg_th) {
    const int c = blockIdx.x;
    const int nsp = m.col_nx[c] * m.col_ny[c];
    const int k0 = m.col_cell_start[c], k1 = m.col_cell_start[c + 1];
    for (int s = threadIdx.x; s < nsp; s += ITHREADS) {
        double acc = 0.0;
        for (int k = k0; k < k1; ++k) {
            const int nt = m.cell_nth[k];
            const double dth = m.cell_geom[2 * k + 1] - m.cell_geom[2 * k + 0];
            const double *wt = m.lg_w + nt * INM;
            const double *uk = u + m.cell_dof_off[k] + s * nt;
            double t = 0.0;
            for (int a = 0; a < nt; ++a) t += wt[a] * uk[a];
            acc += dth / 2.0 * t;
        }
        intg_th[m.col_sp_off[c] + s] = acc;
    }
}

// Per cell: traces at the bottom (theta pulled back to -1) and top (+1) of
// the cell and of its lower / upper neighbours in the column, and
// err[k] = (1/4) sum_ij wx_i wy_j (1/2)((bot_k - top_lo)^2 + (top_k - bot_up)^2)
// (error_angular_jump.py:36-81).  lend[(n 2 + s) INM + a] = l_a(2 s - 1) of
// the n-point Gauss-Legendre basis.
__global__ __launch_bounds__(ITHREADS) void k_angular_jump(DevMesh m, const double *__restrict__ u,
                                                          const int32_t *__restrict__ nbr,
                                                          const double *__restrict__ lend,
                                                          double *__restrict__ err) {
    __shared__ double red[ITHREADS];
    const int k = blockIdx.x, tid = threadIdx.x;
    const int c = m.cell_col[k];
    const int nx = m.col_nx[c], ny = m.col_ny[c];
    const int lo = nbr[2 * k], up = nbr[2 * k + 1];
    const int nt = m.cell_nth[k], ntl = m.cell_nth[lo], ntu = m.cell_nth[up];
    const double *uk = u + m.cell_dof_off[k], *ul = u + m.cell_dof_off[lo],
                 *uu = u + m.cell_dof_off[up];
    const double *bot_k = lend + (nt * 2 + 0) * INM, *top_k = lend + (nt * 2 + 1) * INM;
    const double *top_l = lend + (ntl * 2 + 1) * INM, *bot_u = lend + (ntu * 2 + 0) * INM;
    const double *wx = m.lgl_w + nx * INM, *wy = m.lgl_w + ny * INM;
    double acc = 0.0;
    for (int s = tid; s < nx * ny; s += ITHREADS) {
        double b0 = 0.0, t0 = 0.0, tl = 0.0, bu = 0.0;
        for (int a = 0; a < nt; ++a) {
            b0 += uk[s * nt + a] * bot_k[a];
            t0 += uk[s * nt + a] * top_k[a];
        }
        for (int a = 0; a < ntl; ++a) tl += ul[s * ntl + a] * top_l[a];
        for (int a = 0; a < ntu; ++a) bu += uu[s * ntu + a] * bot_u[a];
        const double jl = (b0 - tl) * (b0 - tl), ju = (t0 - bu) * (t0 - bu);
        acc += wx[s / ny] * wy[s % ny] * 0.5 * (jl + ju);
    }
    acc = block_sum(acc, red);
    if (tid == 0) err[k] = 0.25 * acc;
}

// Per column: the squared jumps of the face traces against every
// neighbouring cell across every face (error_spatial_jump.py:61-225).  Pair
// p (pairs of column c: col_pair_start[c] .. col_pair_start[c + 1] - 1) is
//   pair[RTX_JUMP_PAIR_INTS p + 0..8] = own cell, neighbour cell, face F,
//   nB, nH, and pool offsets of psi_0 (n0 x nB), psi_1 (n1 x nB),
//   xsi_0 (nth0 x nH), xsi_1 (nth1 x nH), row-major;
//   scale[p] = d_A dth_G / 4.
// The own trace is taken on face F, the neighbour's on the opposite face;
// both are projected onto nB LGL nodes x nH Gauss-Legendre nodes of the
// shared interval and col_err[c] = sum scale sum wB wH (U0 - U1)^2 /
// (perimeter 2 pi) (:227).
__global__ __launch_bounds__(ITHREADS) void k_spatial_jump(DevMesh m,
                                                          const double *__restrict__ u,
                                                          const int32_t *__restrict__ col_pair_start,
                                                          const int32_t *__restrict__ pair,
                                                          const double *__restrict__ scale,
                                                          const double *__restrict__ pool,
                                                          double *__restrict__ col_err) {
    __shared__ double h0[INM * INM], h1[INM * INM], T0[INM * INM], T1[INM * INM];
    __shared__ double red[ITHREADS];
    const int c = blockIdx.x, tid = threadIdx.x;
    double col_jump = 0.0;
    for (int p = col_pair_start[c]; p < col_pair_start[c + 1]; ++p) {
        const int32_t *q = pair + RTX_JUMP_PAIR_INTS * p;
        const int k0 = q[0], k1 = q[1], F = q[2], nB = q[3], nH = q[4];
        const double *psi0 = pool + q[5], *psi1 = pool + q[6];
        const double *xsi0 = pool + q[7], *xsi1 = pool + q[8];
        const int c0 = m.cell_col[k0], c1 = m.cell_col[k1];
        const int nx0 = m.col_nx[c0], ny0 = m.col_ny[c0], nt0 = m.cell_nth[k0];
        const int nx1 = m.col_nx[c1], ny1 = m.col_ny[c1], nt1 = m.cell_nth[k1];
        const bool even = (F % 2) == 0;
        const int n0 = even ? ny0 : nx0, n1 = even ? ny1 : nx1;
        const double *u0 = u + m.cell_dof_off[k0], *u1 = u + m.cell_dof_off[k1];
        // own face: F = 0 -> i = nx - 1, 1 -> j = ny - 1, 2 -> i = 0, 3 -> j = 0;
        // the neighbour's is the opposite face
        for (int e = tid; e < n0 * nt0; e += ITHREADS) {
            const int t = e / nt0, a = e % nt0;
            const int i = even ? (F == 0 ? nx0 - 1 : 0) : t;
            const int j = even ? t : (F == 1 ? ny0 - 1 : 0);
            h0[e] = u0[(i * ny0 + j) * nt0 + a];
        }
        for (int e = tid; e < n1 * nt1; e += ITHREADS) {
            const int t = e / nt1, a = e % nt1;
            const int i = even ? (F == 0 ? 0 : nx1 - 1) : t;
            const int j = even ? t : (F == 1 ? 0 : ny1 - 1);
            h1[e] = u1[(i * ny1 + j) * nt1 + a];
        }
        __syncthreads();
        // spatial projection: T[b][a] = sum_t psi[t][b] h[t][a]
        for (int e = tid; e < nB * nt0; e += ITHREADS) {
            const int b = e / nt0, a = e % nt0;
            double s = 0.0;
            for (int t = 0; t < n0; ++t) s += psi0[t * nB + b] * h0[t * nt0 + a];
            T0[e] = s;
        }
        for (int e = tid; e < nB * nt1; e += ITHREADS) {
            const int b = e / nt1, a = e % nt1;
            double s = 0.0;
            for (int t = 0; t < n1; ++t) s += psi1[t * nB + b] * h1[t * nt1 + a];
            T1[e] = s;
        }
        __syncthreads();
        // angular projection and the weighted squared difference
        const double *wB = m.lgl_w + nB * INM, *wH = m.lg_w + nH * INM;
        double acc = 0.0;
        for (int e = tid; e < nB * nH; e += ITHREADS) {
            const int b = e / nH, r = e % nH;
            double s0 = 0.0, s1 = 0.0;
            for (int a = 0; a < nt0; ++a) s0 += T0[b * nt0 + a] * xsi0[a * nH + r];
            for (int a = 0; a < nt1; ++a) s1 += T1[b * nt1 + a] * xsi1[a * nH + r];
            acc += wB[b] * wH[r] * (s0 - s1) * (s0 - s1);
        }
        acc = block_sum(acc, red);      // its barriers also guard h0/h1/T0/T1
        col_jump += scale[p] * acc;
    }
    if (tid == 0) {
        const double dx = m.col_geom[4 * c + 2] - m.col_geom[4 * c + 0];
        const double dy = m.col_geom[4 * c + 3] - m.col_geom[4 * c + 1];
        const double perim = 2.0 * dx + 2.0 * dy;
        col_err[c] = col_jump / (perim * 2.0 * 3.141592653589793);
    }
}

}  // namespace rtx

using namespace rtx;

namespace {
int check_cells(rtx_plan *p, const char *who) {
    RTX_CHECK(p->max_cell_ndof <= INM * INM * INM, "%s: cell too large", who);
    return RTX_OK;
}
}  // namespace

extern "C" int rtx_indicator_moments(rtx_plan *p, const double *d_u, double *d_intg_xy,
                                     double *d_intg_th, double *d_norm2, void *stream) {
    RTX_CHECK(p && d_u && d_intg_xy && d_intg_th, "rtx_indicator_moments: null argument");
    if (int rc = check_cells(p, "rtx_indicator_moments")) return rc;
    if (p->m.ncell == 0) return RTX_OK;
    const hipStream_t st = static_cast<hipStream_t>(stream);
    k_cell_moments<<<p->m.ncell, ITHREADS, 0, st>>>(p->m, d_u, d_intg_xy, d_norm2);
    RTX_HIP(hipGetLastError());
    k_col_intg_th<<<p->m.ncol, ITHREADS, 0, st>>>(p->m, d_u, d_intg_th);
    RTX_HIP(hipGetLastError());
    return RTX_OK;
}

extern "C" int rtx_error_angular_jump(rtx_plan *p, const double *d_u, const int32_t *d_nbr,
                                      const double *d_lend, double *d_cell_err, void *stream) {
    RTX_CHECK(p && d_u && d_nbr && d_lend && d_cell_err, "rtx_error_angular_jump: null argument");
    if (int rc = check_cells(p, "rtx_error_angular_jump")) return rc;
    if (p->m.ncell == 0) return RTX_OK;
    k_angular_jump<<<p->m.ncell, ITHREADS, 0, static_cast<hipStream_t>(stream)>>>(
        p->m, d_u, d_nbr, d_lend, d_cell_err);
    RTX_HIP(hipGetLastError());
    return RTX_OK;
}

extern "C" int rtx_error_spatial_jump(rtx_plan *p, const double *d_u, const int32_t *d_col_pair_start,
                                      const int32_t *d_pair, const double *d_scale,
                                      const double *d_pool, double *d_col_err, void *stream) {
    RTX_CHECK(p && d_u && d_col_pair_start && d_pair && d_scale && d_pool && d_col_err,
              "rtx_error_spatial_jump: null argument");
    if (int rc = check_cells(p, "rtx_error_spatial_jump")) return rc;
    if (p->m.ncol == 0) return RTX_OK;
    k_spatial_jump<<<p->m.ncol, ITHREADS, 0, static_cast<hipStream_t>(stream)>>>(
        p->m, d_u, d_col_pair_start, d_pair, d_scale, d_pool, d_col_err);
    RTX_HIP(hipGetLastError());
    return RTX_OK;
}
