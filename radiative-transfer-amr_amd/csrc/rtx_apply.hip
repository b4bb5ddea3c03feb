// Matrix-free DG transport operator: generic hp kernels.
//
//   k_transport : one wavefront per angular cell.  Mass (M), interior
//                 convection (-IC), outflow face terms and upwind inflow
//                 couplings through the link table (+BC).  The cell's nodal
//                 values are staged in LDS.
//   k_scatter_mma : one workgroup per (column, 16 spatial nodes, 128
//                 output angles) work item.  Adds the angular scattering
//                 contraction (-S) as a matrix-core GEMM against the
//                 column's G^T table.
//
// Formulas (matrix entries of the reference restated as actions):
//   M      : src/rt/mass_matrix.py:140-173
//   IC     : src/rt/interior_convection_matrix.py:161-243
//   BC     : src/rt/boundary_convection_matrix.py:258-537
//   -S     : src/rt/scattering_matrix.py:142-237
#include "rtx_internal.h"

namespace rtx {

constexpr int NM = RTX_NMAX;
constexpr int SCAT_THREADS = 256;
constexpr int SC_KC = 32;                // angles per K step of the tiled scattering
constexpr int TT = 64;                   // threads per cell of k_transport (one wavefront;
                                         // four per cell measured slower)
constexpr int MAX_NANG = 1024;           // angular nodes per column

__device__ inline bool on_bdry(int bd, int i, int j, int nx, int ny) {
    return ((bd & 1) && i == nx - 1) || ((bd & 2) && j == ny - 1) || ((bd & 4) && i == 0) ||
           ((bd & 8) && j == 0);
}

// Theta_F(th) = cos(th - F pi / 2) from the cos/sin of th.
__device__ inline double theta_f(int F, double c, double s) {
    return F == 0 ? c : (F == 1 ? s : (F == 2 ? -c : -s));
}

template <typename T>
__global__ __launch_bounds__(TT) void k_transport(DevMesh m, uint32_t ops, int accumulate,
                                                  const int32_t *__restrict__ cells, int ncells,
                                                  const T *__restrict__ x, T *__restrict__ y) {
    // x and y of the cell (dynamic: twice the plan's largest cell), the
    // small operands of the cell and of one upwind link
    extern __shared__ double t_lds[];
    __shared__ double Dxw[NM * NM], Dyw[NM * NM];   // w_i D[p][i]
    __shared__ double wxs[NM], wys[NM], wts[NM], cth[NM], sth[NM];
    __shared__ double kap[NM * NM];
    __shared__ double tr[NM * NM], inn[NM * NM], Ets[NM * NM], Ess[NM * NM];
    // workgroups go to the 8 XCDs round-robin; XCD j takes the j-th
    // contiguous eighth of the cell list (cells in column order), so the
    // neighbour columns whose traces a cell reads are mostly handled, and
    // cached, on the same XCD
    const int per = (ncells + 7) / 8;
    const int ci = (blockIdx.x % 8) * per + blockIdx.x / 8;
    if (ci >= ncells) return;
    const int k = cells ? cells[ci] : ci;
    const int lane = threadIdx.x;
    const int c = m.cell_col[k];
    const int nx = m.col_nx[c], ny = m.col_ny[c], nth = m.cell_nth[k];
    const int n = nx * ny * nth;
    double *xs = t_lds, *ys = t_lds + n;
    const int off = m.cell_dof_off[k];
    const bool restrict_ = (ops & RTX_OP_RESTRICT) != 0;
    const int bd = restrict_ ? m.cell_bdry[k] : 0;
    const double x0 = m.col_geom[4 * c + 0], y0 = m.col_geom[4 * c + 1];
    const double dx = m.col_geom[4 * c + 2] - x0, dy = m.col_geom[4 * c + 3] - y0;
    const double th0 = m.cell_geom[2 * k + 0], dth = m.cell_geom[2 * k + 1] - th0;
    const int nyt = ny * nth;
    const int sp0 = m.col_sp_off[c];

    for (int t = lane; t < n; t += TT) {
        const int i = t / nyt, j = (t / nth) % ny;
        double v = double(x[off + t]);
        if (bd && on_bdry(bd, i, j, nx, ny)) v = 0.0;
        xs[t] = v;
    }
    for (int t = lane; t < nx * nx; t += TT) {
        const int pp = t / nx, i = t % nx;
        Dxw[t] = m.lgl_w[nx * NM + i] * m.lgl_D[(size_t(nx) * NM + pp) * NM + i];
    }
    for (int t = lane; t < ny * ny; t += TT) {
        const int q = t / ny, j = t % ny;
        Dyw[t] = m.lgl_w[ny * NM + j] * m.lgl_D[(size_t(ny) * NM + q) * NM + j];
    }
    for (int t = lane; t < nx * ny; t += TT) kap[t] = m.kappa[sp0 + t];
    if (lane < nx) wxs[lane] = m.lgl_w[nx * NM + lane];
    if (lane < ny) wys[lane] = m.lgl_w[ny * NM + lane];
    if (lane < nth) {
        wts[lane] = m.lg_w[nth * NM + lane];
        const double th = th0 + dth * 0.5 * (m.lg_x[nth * NM + lane] + 1.0);
        double sn, cs;
        sincos(th, &sn, &cs);
        cth[lane] = cs;
        sth[lane] = sn;
    }
    __syncthreads();

    const int quad = m.cell_quad[k];
    const double vol = dx * dy * dth * 0.125;
    for (int t = lane; t < n; t += TT) {
        const int p = t / nyt, q = (t / nth) % ny, r = t % nth;
        double acc = 0.0;
        const double w3 = vol * wxs[p] * wys[q] * wts[r];
        if (ops & RTX_OP_MASS) acc += w3 * kap[p * ny + q] * xs[t];
        if (ops & RTX_OP_QUADW) acc += w3 * xs[t];
        if (ops & RTX_OP_ICONV) {
            double sx = 0.0, sy = 0.0;
            const double *xi = xs + q * nth + r, *xj = xs + p * nyt + r;
            for (int i = 0; i < nx; ++i) sx += Dxw[p * nx + i] * xi[i * nyt];
            for (int j = 0; j < ny; ++j) sy += Dyw[q * ny + j] * xj[j * nth];
            acc -= (dy * dth * 0.25) * wys[q] * wts[r] * cth[r] * sx +
                   (dx * dth * 0.25) * wxs[p] * wts[r] * sth[r] * sy;
        }
        if ((ops & RTX_OP_BCONV) && quad >= 0) {
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int F = (quad + h) & 3;
                bool on = (F == 0 && p == nx - 1) || (F == 2 && p == 0) ||
                          (F == 1 && q == ny - 1) || (F == 3 && q == 0);
                if (on) {
                    const double cf = ((F & 1) ? dx : dy) * dth * 0.25;
                    const double wm = (F & 1) ? wxs[p] : wys[q];
                    acc += cf * wm * wts[r] * theta_f(F, cth[r], sth[r]) * xs[t];
                }
            }
        }
        ys[t] = acc;
    }

    if (ops & RTX_OP_BCONV) {
        // Upwind links, up to 64 at a time: lane li holds link li's metadata
        // (one round of loads for all links), and the next link's trace and
        // E blocks are fetched into registers while the current one computes.
        static_assert(TT == 64, "link metadata is held one link per lane of the wavefront");
        const int l0 = m.link_start[k], nlk = m.link_start[k + 1] - l0;
        for (int c0 = 0; c0 < nlk; c0 += TT) {
            const int ncl = (nlk - c0) < TT ? (nlk - c0) : TT;
            int mF = 0, moff = 0, mnx = 1, mny = 1, mnt = 1, mbd = 0, mes = 0, met = 0;
            if (lane < ncl) {
                const int l = l0 + c0 + lane;
                const int s = m.link_src[l];
                mF = m.link_face[l];
                mes = m.link_Es_off[l];
                met = m.link_Eth_off[l];
                const int c1 = m.cell_col[s];
                mnx = m.col_nx[c1];
                mny = m.col_ny[c1];
                mnt = m.cell_nth[s];
                moff = m.cell_dof_off[s];
                mbd = restrict_ ? m.cell_bdry[s] : 0;
            }
            constexpr int PU = (NM * NM + TT - 1) / TT;    // register slots per lane
            double pt[PU], pe[PU], ps[PU];
            auto fetch = [&](int li) {
                const int F = __shfl(mF, li), nx1 = __shfl(mnx, li), ny1 = __shfl(mny, li);
                const int nt1 = __shfl(mnt, li), off1 = __shfl(moff, li), bd1 = __shfl(mbd, li);
                const int es = __shfl(mes, li), et = __shfl(met, li);
                const int n1 = (F & 1) ? nx1 : ny1, n0 = (F & 1) ? nx : ny;
#pragma unroll
                for (int u = 0; u < PU; ++u) {
                    const int t = lane + u * TT;
                    double v = 0.0;
                    if (t < n1 * nt1) {
                        const int m1 = t / nt1, a = t % nt1;
                        int i1, j1;
                        if (F == 0) { i1 = 0; j1 = m1; }
                        else if (F == 2) { i1 = nx1 - 1; j1 = m1; }
                        else if (F == 1) { i1 = m1; j1 = 0; }
                        else { i1 = m1; j1 = ny1 - 1; }
                        v = double(x[off1 + (i1 * ny1 + j1) * nt1 + a]);
                        if (bd1 && on_bdry(bd1, i1, j1, nx1, ny1)) v = 0.0;
                    }
                    pt[u] = v;
                    pe[u] = (t < nt1 * nth) ? m.Eth_pool[et + t] : 0.0;
                    ps[u] = (t < n1 * n0) ? m.Es_pool[es + t] : 0.0;
                }
            };
            if (ncl > 0) fetch(0);
            for (int li = 0; li < ncl; ++li) {
                const int F = __shfl(mF, li), nx1 = __shfl(mnx, li), ny1 = __shfl(mny, li);
                const int nt1 = __shfl(mnt, li);
                const int n1 = (F & 1) ? nx1 : ny1;   // neighbour trace length
                const int n0 = (F & 1) ? nx : ny;     // own trace length
                __syncthreads();                      // tr, inn, Ets, Ess reuse
#pragma unroll
                for (int u = 0; u < PU; ++u) {
                    const int t = lane + u * TT;
                    if (t < n1 * nt1) tr[t] = pt[u];
                    if (t < nt1 * nth) Ets[t] = pe[u];
                    if (t < n1 * n0) Ess[t] = ps[u];
                }
                __syncthreads();
                if (li + 1 < ncl) fetch(li + 1);      // in flight during this link's sums
                // sum-factorised: angle first (n1 x nth), then space (n0 x nth)
                for (int t = lane; t < n1 * nth; t += TT) {
                    const int m1 = t / nth, r = t % nth;
                    double v = 0.0;
                    for (int a = 0; a < nt1; ++a) v += Ets[a * nth + r] * tr[m1 * nt1 + a];
                    inn[t] = v;
                }
                __syncthreads();
                const double cf = ((F & 1) ? dx : dy) * dth * 0.25;
                for (int t = lane; t < n0 * nth; t += TT) {
                    const int m0 = t / nth, r = t % nth;
                    double acc = 0.0;
                    for (int m1 = 0; m1 < n1; ++m1) acc += Ess[m1 * n0 + m0] * inn[m1 * nth + r];
                    int node;
                    if (F == 0) node = (nx - 1) * nyt + m0 * nth + r;
                    else if (F == 2) node = m0 * nth + r;
                    else if (F == 1) node = m0 * nyt + (ny - 1) * nth + r;
                    else node = m0 * nyt + r;
                    ys[node] += cf * acc;   // one lane per node: no conflict
                }
            }
        }
    }
    __syncthreads();
    for (int t = lane; t < n; t += TT) {
        const int i = t / nyt, j = (t / nth) % ny;
        const bool masked = restrict_ && on_bdry(m.cell_bdry[k], i, j, nx, ny);
        double v = masked ? 0.0 : ys[t];
        if (accumulate) {
            if (!masked) y[off + t] = T(double(y[off + t]) + v);
        } else {
            y[off + t] = T(v);
        }
    }
}

template <typename T> struct ScatMfma;
template <> struct ScatMfma<double> {
    typedef double v4 __attribute__((ext_vector_type(4)));
    __device__ static v4 step(double a, double b, v4 c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    // row (spatial node of the tile) of accumulator v of this lane
    __device__ static int row(int lg, int v) { return lg + 4 * v; }
    __device__ static const double *G(const DevMesh &m) { return m.G; }
};
template <> struct ScatMfma<float> {
    typedef float v4 __attribute__((ext_vector_type(4)));
    __device__ static v4 step(float a, float b, v4 c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    __device__ static int row(int lg, int v) { return 4 * lg + v; }
    __device__ static const float *G(const DevMesh &m) { return m.G_f; }
};

// Scattering of one work item: 16 spatial nodes (s) x SC_RG output angles
// (R) of a column, Y[s][R] = sum_A X[s][A] G^T[A][R] on the matrix cores
// (FP64 MFMA for double, FP32 MFMA for float), then scaled by
// -dx dy w_i w_j sigma_s (the angular weights and constant factors of the
// reference's entry are folded into G^T).  The X tile (16 x SC_KC angles, inflow-masked) is gathered into
// LDS per K step; every wave owns SC_RG / 64 output tiles of 16 angles and
// reads its G^T operands straight from L2.  Each (s, R) has one owner, so
// accumulation into y is a plain read-modify-write.
template <typename T>
__global__ __launch_bounds__(SCAT_THREADS) void k_scatter_mma(DevMesh m, uint32_t ops,
                                                              int accumulate,
                                                              const int32_t *__restrict__ items,
                                                              const T *__restrict__ x,
                                                              T *__restrict__ y) {
    typedef ScatMfma<T> F;
    typedef typename F::v4 v4;
    constexpr int RT = SC_RG / 64;             // output tiles per wave
    __shared__ int a_dof[MAX_NANG];            // first DOF of the cell owning angle A
    __shared__ int a_meta[MAX_NANG];           // nth | local index << 8 | masked faces << 16
    __shared__ T xs[16][SC_KC + 1];
    const int c = items[2 * blockIdx.x];
    if (c < 0) return;                         // padding of the XCD interleave
    const int word = items[2 * blockIdx.x + 1];
    const int s0 = (word & 0xffff) * 16, rbase = (word >> 16) * SC_RG;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int l16 = lane & 15, lg = lane >> 4;
    const int nx = m.col_nx[c], ny = m.col_ny[c], nsp = nx * ny;
    const int na = m.col_nang[c];
    const int k0 = m.col_cell_start[c], k1 = m.col_cell_start[c + 1];
    const bool restrict_ = (ops & RTX_OP_RESTRICT) != 0;
    for (int k = k0 + tid; k < k1; k += SCAT_THREADS) {
        const int nt = m.cell_nth[k], ao = m.cell_ang_off[k], dof = m.cell_dof_off[k];
        const int bd = restrict_ ? m.cell_bdry[k] : 0;
        for (int a = 0; a < nt; ++a) {
            a_dof[ao + a] = dof;
            a_meta[ao + a] = nt | (a << 8) | (bd << 16);
        }
    }
    const T *G = F::G(m) + m.col_G_off[c];
    v4 acc[RT];
#pragma unroll
    for (int t = 0; t < RT; ++t) acc[t] = v4{0, 0, 0, 0};
    __syncthreads();
    for (int A0 = 0; A0 < na; A0 += SC_KC) {
        for (int t = tid; t < 16 * SC_KC; t += SCAT_THREADS) {
            const int sl = t / SC_KC, al = t % SC_KC, s = s0 + sl, A = A0 + al;
            T v = T(0);
            if (s < nsp && A < na) {
                const int meta = a_meta[A], nt = meta & 255, loc = (meta >> 8) & 255;
                const int bd = meta >> 16;
                v = x[a_dof[A] + s * nt + loc];
                if (bd && on_bdry(bd, s / ny, s % ny, nx, ny)) v = T(0);
            }
            xs[sl][al] = v;
        }
        __syncthreads();
#pragma unroll
        for (int t = 0; t < RT; ++t) {
            const int R = rbase + (wave * RT + t) * 16 + l16;
            if (rbase + (wave * RT + t) * 16 >= na) continue;   // uniform per wave
#pragma unroll
            for (int kk = 0; kk < SC_KC; kk += 4) {
                const int Ak = A0 + kk + lg;
                const T b = (Ak < na && R < na) ? G[size_t(Ak) * na + R] : T(0);
                acc[t] = F::step(xs[l16][kk + lg], b, acc[t]);
            }
        }
        __syncthreads();
    }
    const double x0 = m.col_geom[4 * c + 0], y0 = m.col_geom[4 * c + 1];
    const double dxdy = (m.col_geom[4 * c + 2] - x0) * (m.col_geom[4 * c + 3] - y0);
    const double *wx = m.lgl_w + nx * NM, *wy = m.lgl_w + ny * NM;
    const double *sig = m.sigma + m.col_sp_off[c];
#pragma unroll
    for (int t = 0; t < RT; ++t) {
        const int R = rbase + (wave * RT + t) * 16 + l16;
        if (R >= na) continue;
        const int meta = a_meta[R], nt = meta & 255, loc = (meta >> 8) & 255, bd = meta >> 16;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int s = s0 + F::row(lg, v);
            if (s >= nsp) continue;
            const int i = s / ny, j = s % ny;
            const double val = -dxdy * wx[i] * wy[j] * sig[s] * double(acc[t][v]);
            const int addr = a_dof[R] + s * nt + loc;
            const bool masked = bd && on_bdry(bd, i, j, nx, ny);
            if (accumulate) {
                if (!masked) y[addr] = T(double(y[addr]) + val);
            } else {
                y[addr] = T(masked ? 0.0 : val);
            }
        }
    }
}

template <typename T>
__global__ void k_zero(T *y, int64_t n) {
    int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    for (; i < n; i += int64_t(gridDim.x) * blockDim.x) y[i] = T(0);
}

// keep == 0: zero inflow-boundary DOFs; keep == 1: zero interior DOFs.
template <typename T>
__global__ __launch_bounds__(64) void k_mask(DevMesh m, int keep, T *__restrict__ x) {
    const int k = blockIdx.x;
    const int c = m.cell_col[k];
    const int nx = m.col_nx[c], ny = m.col_ny[c], nth = m.cell_nth[k];
    const int bd = m.cell_bdry[k];
    const int off = m.cell_dof_off[k];
    for (int t = threadIdx.x; t < nx * ny * nth; t += 64) {
        const int i = t / (ny * nth), j = (t / nth) % ny;
        const bool b = bd && on_bdry(bd, i, j, nx, ny);
        if (b == (keep == 0)) x[off + t] = T(0);
    }
}

// Generic per-cell transport + per-column scattering on all cells / columns
// (cells, cols null) or on the given lists (the columns no fused bin takes).
template <typename T>
static int launch_apply(rtx_plan *p, uint32_t ops, const void *dx, void *dy, hipStream_t st,
                        const int32_t *cells = nullptr, int ncells = -1,
                        const int32_t *items = nullptr, int nitems = -1) {
    const T *x = static_cast<const T *>(dx);
    T *y = static_cast<T *>(dy);
    const bool acc = (ops & RTX_OP_ACCUM) != 0;
    const bool tp = (ops & (RTX_OP_MASS | RTX_OP_ICONV | RTX_OP_BCONV | RTX_OP_QUADW)) != 0;
    const bool sc = (ops & RTX_OP_SCAT) != 0;
    if (ncells < 0) ncells = p->m.ncell;
    if (!items) {
        items = p->sc_all;
        nitems = p->n_sc_all;
    }
    if (!tp && !sc) {
        if (!acc) k_zero<T><<<1024, 256, 0, st>>>(y, p->m.ndof);
        RTX_HIP(hipGetLastError());
        return RTX_OK;
    }
    if (tp && ncells > 0)
        k_transport<T><<<8 * ((ncells + 7) / 8), TT, size_t(2) * p->max_cell_ndof * sizeof(double),
                         st>>>(p->m, ops, acc ? 1 : 0, cells, ncells, x, y);
    if (sc && nitems > 0)
        k_scatter_mma<T><<<nitems, SCAT_THREADS, 0, st>>>(p->m, ops, (acc || tp) ? 1 : 0, items,
                                                          x, y);
    RTX_HIP(hipGetLastError());
    return RTX_OK;
}

// The fused per-column kernel holds a whole column (values and accumulators)
// in LDS and tiles its spatial nodes by 16 up to 7 tiles.

static bool any_fused(const rtx_plan *p, bool all_bins) {
    for (const FusedBin &fb : p->fbin)
        if (fb.n > 0 && (all_bins || !fb.generic)) return true;
    return false;
}

}  // namespace rtx

using namespace rtx;

/* Select the apply path: -1 automatic, 0 the generic per-cell + per-column
   kernels, 1 the fused per-part kernel, 2 the uniform-bin kernel, 3 the hp
   column kernel. */
extern "C" int rtx_set_path(rtx_plan *p, int mode) {
    RTX_CHECK(p && mode >= -1 && mode <= 3, "rtx_set_path: bad argument");
    p->fused_mode = mode;
    return RTX_OK;
}

extern "C" int rtx_apply(rtx_plan *p, uint32_t op_mask, int precision, const void *d_x,
                         void *d_y, void *stream) {
    RTX_CHECK(p && d_x && d_y, "rtx_apply: null argument");
    if (!p->has_coeffs) {
        set_error("rtx_apply: coefficients not set (call rtx_plan_set_coeffs)");
        return RTX_ERR_STATE;
    }
    RTX_CHECK(d_x != d_y, "rtx_apply: x and y must not alias");
    RTX_CHECK(p->max_nang <= MAX_NANG, "rtx_apply: a column has %d angular nodes (max %d)",
              p->max_nang, MAX_NANG);
    hipStream_t st = static_cast<hipStream_t>(stream);
    RTX_CHECK(precision >= RTX_PREC_F64 && precision <= RTX_PREC_MIXED,
              "rtx_apply: unsupported precision %d", precision);
    const int acc = (op_mask & RTX_OP_ACCUM) ? 1 : 0;
    if ((p->fused_mode == -1 || p->fused_mode == 2) && p->uni_ok) {
        const int rc = launch_uniform(p, op_mask, precision, d_x, d_y, st, acc);
        if (rc <= 0) return rc;
    }
    RTX_CHECK(p->fused_mode != 2, "rtx_apply: the uniform-bin path was requested but the mesh "
                                  "is not a compiled uniform bin");
    if (p->fused_mode == 3 || (p->fused_mode == -1 && p->hp_ok && p->hp_default)) {
        // hp column kernel (mass, scattering, both convections, restriction)
        RTX_CHECK(p->hp_ok, "rtx_apply: the hp column path was requested but no column fits it");
        RTX_CHECK(!(op_mask & RTX_OP_QUADW) || p->fused_mode != 3,
                  "rtx_apply: the hp column path does not apply quadrature weights");
        if (!(op_mask & RTX_OP_QUADW)) {
            int rc = launch_hpcol(p, op_mask, precision, d_x, d_y, st, acc);
            if (rc || p->hp_over_cols.empty()) return rc;
            if (precision == RTX_PREC_F64)
                return launch_apply<double>(p, op_mask, d_x, d_y, st, p->hp_gen_cells,
                                            p->hp_n_gen_cells, p->hp_gen_items,
                                            p->hp_n_gen_items);
            return launch_apply<float>(p, op_mask, d_x, d_y, st, p->hp_gen_cells,
                                       p->hp_n_gen_cells, p->hp_gen_items, p->hp_n_gen_items);
        }
    }
    const bool forced = p->fused_mode == 1;
    RTX_CHECK(!forced || p->h_over_cols.empty(),
              "rtx_apply: the fused per-column path was requested but %d columns do not fit "
              "it (largest column: DOFs %d, cells %d, nodes %d)",
              int(p->h_over_cols.size()), p->max_col_dofs, p->max_col_cells, p->max_nsp);
    if ((p->fused_mode == -1 || forced) && any_fused(p, forced)) {
        // order bins: fused bins, then the generic kernels on the rest
        const bool tp = (op_mask & (RTX_OP_MASS | RTX_OP_ICONV | RTX_OP_BCONV | RTX_OP_QUADW)) != 0;
        const bool sc = (op_mask & RTX_OP_SCAT) != 0;
        if (!tp && !sc && !acc) {
            if (precision == RTX_PREC_F64)
                k_zero<double><<<1024, 256, 0, st>>>(static_cast<double *>(d_y), p->m.ndof);
            else
                k_zero<float><<<1024, 256, 0, st>>>(static_cast<float *>(d_y), p->m.ndof);
            RTX_HIP(hipGetLastError());
            return RTX_OK;
        }
        int rc = launch_fused(p, op_mask, precision, d_x, d_y, st, acc, -1, forced);
        if (rc || forced || p->n_gen_cols == 0) return rc;
        if (precision == RTX_PREC_F64)
            return launch_apply<double>(p, op_mask, d_x, d_y, st, p->gen_cells, p->n_gen_cells,
                                        p->sc_gen, p->n_sc_gen);
        return launch_apply<float>(p, op_mask, d_x, d_y, st, p->gen_cells, p->n_gen_cells,
                                   p->sc_gen, p->n_sc_gen);
    }
    switch (precision) {
        case RTX_PREC_F64: return launch_apply<double>(p, op_mask, d_x, d_y, st);
        case RTX_PREC_F32: return launch_apply<float>(p, op_mask, d_x, d_y, st);
        case RTX_PREC_MIXED: return launch_apply<float>(p, op_mask, d_x, d_y, st);
        default: set_error("rtx_apply: unsupported precision %d", precision); return RTX_ERR_ARG;
    }
}

extern "C" int rtx_mask(rtx_plan *p, int keep, int precision, void *d_x, void *stream) {
    RTX_CHECK(p && d_x && (keep == 0 || keep == 1), "rtx_mask: bad argument");
    hipStream_t st = static_cast<hipStream_t>(stream);
    if (precision == RTX_PREC_F64)
        k_mask<double><<<p->m.ncell, 64, 0, st>>>(p->m, keep, static_cast<double *>(d_x));
    else
        k_mask<float><<<p->m.ncell, 64, 0, st>>>(p->m, keep, static_cast<float *>(d_x));
    RTX_HIP(hipGetLastError());
    return RTX_OK;
}

namespace {

// Mean duration (us) of `reps` launches of fn on the stream.
template <typename F>
float time_launches(F fn, int reps, hipStream_t st, int *rc) {
    hipEvent_t e0, e1;
    *rc = 0;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
        *rc = RTX_ERR_HIP;
        return 0.f;
    }
    *rc = fn();                       // warm-up (also sets the LDS attribute)
    (void)hipEventRecord(e0, st);
    for (int i = 0; i < reps && *rc == 0; ++i) *rc = fn();
    (void)hipEventRecord(e1, st);
    (void)hipEventSynchronize(e1);
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return 1000.f * ms / float(reps > 0 ? reps : 1);
}

}  // namespace

/* Autotune the apply's launch shapes on this plan: for the uniform-bin
   kernel and for every fused order bin, time each valid number of column
   parts H (1, 2, 4) with `reps` launches of the interior operator on d_x ->
   d_y (device vectors in `precision`) and keep the fastest; then time the
   bin's columns on the generic kernels and move the bin there if they are
   faster.  Synchronises the stream.  The choice holds for every precision. */
extern "C" int rtx_autotune(rtx_plan *p, int precision, const void *d_x, void *d_y, int reps,
                            void *stream) {
    RTX_CHECK(p && d_x && d_y && d_x != d_y && reps >= 1, "rtx_autotune: bad argument");
    RTX_CHECK(p->has_coeffs, "rtx_autotune: coefficients not set");
    hipStream_t st = static_cast<hipStream_t>(stream);
    const uint32_t ops = RTX_OP_SYSTEM | RTX_OP_RESTRICT;
    int rc = 0;
    if (p->uni_ok) {
        const int mask = uniform_valid_parts(p);
        float best = 1e30f;
        int bestH = p->uni_H;
        // 0: the output-layout kernel; 1, 2, 4: the line-pass kernel's parts
        for (int H : {0, 1, 2, 4}) {
            if (H && !(mask & H)) continue;
            p->uni_H = H;
            const float us = time_launches(
                [&]() { return launch_uniform(p, ops, precision, d_x, d_y, st, 0); }, reps, st, &rc);
            if (rc) return rc;
            if (us < best) { best = us; bestH = H; }
        }
        p->uni_H = bestH;
    }
    // the order bins and generic kernels are tuned with the hp column kernel
    // off; it is compared against their best afterwards
    p->hp_default = false;
    bool changed = false;
    for (int bi = 0; bi < RTX_NBINS && !p->uni_ok; ++bi) {   // a uniform mesh uses no bin
        FusedBin &fb = p->fbin[bi];
        if (fb.n == 0) continue;
        const int mask = fused_bin_valid_parts(p, bi, precision);
        float best = 1e30f;
        int bestH = fb.H;
        for (int H : {1, 2, 4}) {
            if (!(mask & H)) continue;
            fb.H = H;
            const float us = time_launches(
                [&]() { return launch_fused(p, ops, precision, d_x, d_y, st, 0, bi); }, reps, st,
                &rc);
            if (rc) return rc;
            if (us < best) { best = us; bestH = H; }
        }
        fb.H = bestH;
        // the same columns on the generic kernels
        auto gen = [&]() {
            if (precision == RTX_PREC_F64)
                return launch_apply<double>(p, ops, d_x, d_y, st, fb.cells, fb.ncells, fb.items,
                                            fb.nitems);
            return launch_apply<float>(p, ops, d_x, d_y, st, fb.cells, fb.ncells, fb.items,
                                       fb.nitems);
        };
        const float ug = time_launches(gen, reps, st, &rc);
        if (rc) return rc;
        const bool g = ug < best;
        changed |= g != fb.generic;
        fb.generic = g;
    }
    if (changed && (rc = rebuild_generic(p))) return rc;
    // the mix against every bin on the generic kernels (one launch pair)
    if (!p->uni_ok && any_fused(p, false)) {
        bool saved[RTX_NBINS];
        for (int bi = 0; bi < RTX_NBINS; ++bi) saved[bi] = p->fbin[bi].generic;
        auto whole = [&]() { return rtx_apply(p, ops, precision, d_x, d_y, stream); };
        const int mode = p->fused_mode;
        p->fused_mode = -1;
        const float t_mix = time_launches(whole, reps, st, &rc);
        if (rc) return rc;
        for (int bi = 0; bi < RTX_NBINS; ++bi) p->fbin[bi].generic = true;
        if ((rc = rebuild_generic(p))) return rc;
        const float t_gen = time_launches(whole, reps, st, &rc);
        if (rc) return rc;
        if (t_mix <= t_gen) {
            for (int bi = 0; bi < RTX_NBINS; ++bi) p->fbin[bi].generic = saved[bi];
            if ((rc = rebuild_generic(p))) return rc;
        }
        p->fused_mode = mode;
    }
    // the hp column kernel against the bins and generic kernels chosen above
    if (!p->uni_ok && p->hp_ok) {
        auto whole = [&]() { return rtx_apply(p, ops, precision, d_x, d_y, stream); };
        const int mode = p->fused_mode;
        p->fused_mode = -1;
        p->hp_default = true;
        const float t_hp = time_launches(whole, reps, st, &rc);
        if (rc) return rc;
        p->hp_default = false;
        const float t_rest = time_launches(whole, reps, st, &rc);
        if (rc) return rc;
        p->hp_default = t_hp <= t_rest;
        p->fused_mode = mode;
    }
    return RTX_OK;
}

/* Run fused bin `bin` on the generic kernels (generic = 1) or fused (0). */
extern "C" int rtx_set_bin_generic(rtx_plan *p, int bin, int generic) {
    RTX_CHECK(p && bin >= 0 && bin < RTX_NBINS && (generic == 0 || generic == 1),
              "rtx_set_bin_generic: bad argument");
    if (p->fbin[bin].generic == (generic != 0)) return RTX_OK;
    p->fbin[bin].generic = generic != 0;
    return rebuild_generic(p);
}

/* Force the line-pass uniform kernel with 1, 2 or 4 parts per column; 0
   restores the default (FP64: the output-layout kernel, one column per
   workgroup, when every column is resident at once, else the line-pass
   kernel; FP32 and mixed: the line-pass kernel with one part). */
extern "C" int rtx_set_uniform_parts(rtx_plan *p, int H) {
    RTX_CHECK(p, "rtx_set_uniform_parts: null plan");
    RTX_CHECK(H == 0 || (p->uni_ok && (H == 1 || H == 2 || H == 4) &&
                         (uniform_valid_parts(p) & H)),
              "rtx_set_uniform_parts: %d parts not valid for this plan", H);
    p->uni_H = H;
    return RTX_OK;
}

/* Launch-shape summary: out[0] uniform bin usable, [1] its parts per column,
   [2 + 3 b .. 4 + 3 b] parts, columns and generic flag of fused bin b
   (b < 6), [20], [21] columns and cells on the generic kernels, [22] the hp
   column kernel is the default path, [23] columns it takes, [24] the uniform
   bin's FP64 apply runs the output-layout kernel (1: the default when every
   column is resident at once, four per CU) or the line-pass kernel (0).
   Returns the count (25). */
extern "C" int rtx_plan_info(const rtx_plan *p, int64_t *out, int n) {
    constexpr int NV = 7 + 3 * RTX_NBINS;
    if (!p) return 0;
    int64_t v[NV] = {};
    v[0] = p->uni_ok ? 1 : 0;
    v[1] = p->uni_ok ? uniform_default_parts(p) : 0;
    for (int b = 0; b < RTX_NBINS; ++b) {
        v[2 + 3 * b] = p->fbin[b].H;
        v[3 + 3 * b] = p->fbin[b].n;
        v[4 + 3 * b] = p->fbin[b].generic ? 1 : 0;
    }
    v[20] = p->n_gen_cols;
    v[21] = p->n_gen_cells;
    v[22] = p->hp_default ? 1 : 0;
    int hp_cols = 0;
    for (int k = 0; k < HP_NCLS; ++k) hp_cols += p->hp_cls[k].n;
    v[23] = hp_cols;
    v[24] = (p->uni_ok && p->uni_H == 0 && p->m.ncol <= 4 * p->n_cu) ? 1 : 0;
    for (int i = 0; i < NV && i < n && out; ++i) out[i] = v[i];
    return NV;
}
