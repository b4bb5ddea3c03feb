// Fused per-column transport operator (the fast path of rtx_apply).
//
// One workgroup per spatial column; the column's DOFs are contiguous in the
// reference's global order (columns by key, cells by key, then
// beta = nth ny i + nth j + a), so the column is read once, coalesced, into
// LDS, every term is accumulated in LDS, and the result is written once:
//
//   set-up   per-column constants (weights, weighted derivative matrices,
//            kappa/sigma at the nodes, per-cell angles) -> LDS
//   phase 1  x (column) and the upwind neighbour traces of every link -> LDS
//            (inflow-boundary mask applied)
//   phase 2  mass + (-IC) + outflow faces, per DOF  (mass_matrix.py:140-173,
//            interior_convection_matrix.py:161-243,
//            boundary_convection_matrix.py:284-346)
//   phase 3  upwind inflow couplings, one (link, face node) per thread, LDS
//            atomics for faces shared by two links
//            (boundary_convection_matrix.py:348-524)
//   phase 4  scattering: per spatial node s, y[s, R] -= dx dy w w sigma_s
//            sum_A G^T[A, R] x[s, A] as an MFMA GEMM with M = spatial nodes,
//            N = output angles, K = input angles (scattering_matrix.py:142-237):
//              F64   -> v_mfma_f64_16x16x4f64   (x and G in FP64)
//              F32   -> v_mfma_f32_16x16x4f32   (x and G in FP32)
//              MIXED -> v_mfma_f32_16x16x16bf16 (x and G in BF16, FP32 accum.)
//   phase 5  LDS -> y                               (mask, optional accumulate)
//
// Wavefront w owns the output-angle tiles w, w + nwaves, ... so no two lanes
// write the same LDS entry in phase 4.
#include <type_traits>

#include "rtx_internal.h"

namespace rtx {

typedef double dbl4 __attribute__((ext_vector_type(4)));
typedef float flt4 __attribute__((ext_vector_type(4)));
typedef short sht4 __attribute__((ext_vector_type(4)));

constexpr int FNM = RTX_NMAX;
constexpr int FMAX_MT = 7;        // spatial-node tiles of 16 (nx * ny <= 112)
constexpr int FMAX_CELLS = 64;    // cells per column on this path
constexpr int FMAX_LINKS = 256;   // upwind links per column on this path
constexpr int FMAX_NSP = FMAX_MT * 16;

__device__ inline bool f_on_bdry(int bd, int i, int j, int nx, int ny) {
    return ((bd & 1) && i == nx - 1) || ((bd & 2) && j == ny - 1) || ((bd & 4) && i == 0) ||
           ((bd & 8) && j == 0);
}

__device__ inline uint16_t f_bf16(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);
    return uint16_t(u >> 16);
}

// Largest i with starts[i] <= t (starts ascending, starts[0] = 0).
__device__ inline int f_find(const int *starts, int n, int t) {
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (starts[mid] <= t) lo = mid; else hi = mid - 1;
    }
    return lo;
}

// Exclusive prefix sum of v[0..n) into out[0..n], n <= nthreads; returns total.
__device__ inline int f_scan(int *buf, int n, int v_own, int *out) {
    const int tid = threadIdx.x;
    if (tid < n) buf[tid] = v_own;
    __syncthreads();
    if (tid == 0) {
        int acc = 0;
        for (int i = 0; i < n; ++i) {
            out[i] = acc;
            acc += buf[i];
        }
        out[n] = acc;
    }
    __syncthreads();
    return out[n];
}

// A: arithmetic type of the transport terms and of the LDS accumulators.
// T: storage type of x and y.  PREC: RTX_PREC_*.
template <typename T, typename A, int PREC>
__global__ __launch_bounds__(512) void k_fused(DevMesh m, uint32_t ops, int accumulate,
                                              const T *__restrict__ x, T *__restrict__ y) {
    extern __shared__ __align__(16) unsigned char f_lds[];
    __shared__ int cst[FMAX_CELLS + 1];      // column-local cell starts
    __shared__ int cbd[FMAX_CELLS];          // masked faces per cell
    __shared__ int cnt[FMAX_CELLS];          // nth per cell
    __shared__ A cdth[FMAX_CELLS];
    __shared__ A cth[FMAX_CELLS * FNM], sth[FMAX_CELLS * FNM], cwt[FMAX_CELLS * FNM];
    __shared__ A wxs[FNM], wys[FNM], WDx[FNM * FNM], WDy[FNM * FNM];
    __shared__ A kap[FMAX_NSP], sgw[FMAX_NSP];
    __shared__ int l_tr[FMAX_LINKS + 1], l_it[FMAX_LINKS + 1], scan_buf[512];
    const int c = blockIdx.x, tid = threadIdx.x, nthr = blockDim.x;
    const int lane = tid & 63, wave = tid >> 6, nwaves = nthr >> 6;
    const int k0 = m.col_cell_start[c], k1 = m.col_cell_start[c + 1], ncells = k1 - k0;
    const int nx = m.col_nx[c], ny = m.col_ny[c], nsp = nx * ny;
    const int na = m.col_nang[c];
    const int cdof = m.cell_dof_off[k0];
    const int L0 = m.link_start[k0], nl = m.link_start[k1] - L0;
    const bool restrict_ = (ops & RTX_OP_RESTRICT) != 0;
    const double dx = m.col_geom[4 * c + 2] - m.col_geom[4 * c + 0];
    const double dy = m.col_geom[4 * c + 3] - m.col_geom[4 * c + 1];

    // ---- set-up: per-column constants
    for (int kk = tid; kk <= ncells; kk += nthr) {
        if (kk < ncells) {
            const int k = k0 + kk;
            cst[kk] = m.cell_dof_off[k] - cdof;
            cbd[kk] = restrict_ ? m.cell_bdry[k] : 0;
            cnt[kk] = m.cell_nth[k];
            cdth[kk] = A(m.cell_geom[2 * k + 1] - m.cell_geom[2 * k]);
        } else {
            cst[kk] = m.cell_dof_off[k1 - 1] - cdof + nsp * m.cell_nth[k1 - 1];
        }
    }
    for (int e = tid; e < ncells * FNM; e += nthr) {
        const int kk = e / FNM, r = e % FNM;
        const int nt = m.cell_nth[k0 + kk];
        if (r < nt) {
            const double th0 = m.cell_geom[2 * (k0 + kk)], th1 = m.cell_geom[2 * (k0 + kk) + 1];
            double s, cc;
            sincos(th0 + (th1 - th0) * 0.5 * (m.lg_x[nt * FNM + r] + 1.0), &s, &cc);
            cth[e] = A(cc);
            sth[e] = A(s);
            cwt[e] = A(m.lg_w[nt * FNM + r]);
        }
    }
    for (int e = tid; e < FNM * FNM; e += nthr) {
        const int p = e / FNM, i = e % FNM;
        if (p < nx && i < nx) WDx[e] = A(m.lgl_w[nx * FNM + i] * m.lgl_D[(size_t(nx) * FNM + p) * FNM + i]);
        if (p < ny && i < ny) WDy[e] = A(m.lgl_w[ny * FNM + i] * m.lgl_D[(size_t(ny) * FNM + p) * FNM + i]);
    }
    if (tid < FNM) {
        wxs[tid] = A(tid < nx ? m.lgl_w[nx * FNM + tid] : 0.0);
        wys[tid] = A(tid < ny ? m.lgl_w[ny * FNM + tid] : 0.0);
    }
    for (int s = tid; s < nsp; s += nthr) {
        const int p = s / ny, q = s % ny;
        kap[s] = A(m.kappa[m.col_sp_off[c] + s]);
        sgw[s] = A(dx * dy * m.lgl_w[nx * FNM + p] * m.lgl_w[ny * FNM + q] *
                   m.sigma[m.col_sp_off[c] + s]);
    }
    // per-link trace sizes and phase-3 item counts (prefix sums)
    int tr_own = 0, it_own = 0;
    if (tid < nl) {
        const int l = L0 + tid, s = m.link_src[l], F = m.link_face[l];
        const int c1 = m.cell_col[s], tk = m.link_target[l];
        const int n1 = (F & 1) ? m.col_nx[c1] : m.col_ny[c1];
        const int n0 = (F & 1) ? nx : ny;
        tr_own = n1 * m.cell_nth[s];
        it_own = n0 * m.cell_nth[tk];
    }
    f_scan(scan_buf, nl, tr_own, l_tr);
    f_scan(scan_buf, nl, it_own, l_it);
    const int ncd = cst[ncells];
    A *xs = reinterpret_cast<A *>(f_lds);
    A *ys = xs + ncd;
    A *tr = ys + ncd;
    int *a_base = reinterpret_cast<int *>(tr + l_tr[nl]);
    int *a_nt = a_base + na;
    int *a_loc = a_nt + na;
    for (int kk = 0; kk < ncells; ++kk) {
        const int nt = cnt[kk], ao = m.cell_ang_off[k0 + kk];
        for (int a = tid; a < nt; a += nthr) {
            a_base[ao + a] = cst[kk];
            a_nt[ao + a] = nt;
            a_loc[ao + a] = a;
        }
    }

    // ---- phase 1: column values and neighbour traces
    for (int t = tid; t < ncd; t += nthr) {
        A v = A(x[cdof + t]);
        const int kk = f_find(cst, ncells, t);
        if (cbd[kk]) {
            const int lt = t - cst[kk], nt = cnt[kk];
            if (f_on_bdry(cbd[kk], lt / (ny * nt), (lt / nt) % ny, nx, ny)) v = A(0);
        }
        xs[t] = v;
    }
    if (ops & RTX_OP_BCONV) {
        for (int e = tid; e < l_tr[nl]; e += nthr) {
            const int li = f_find(l_tr, nl, e), l = L0 + li;
            const int s = m.link_src[l], F = m.link_face[l];
            const int c1 = m.cell_col[s];
            const int nx1 = m.col_nx[c1], ny1 = m.col_ny[c1], nt1 = m.cell_nth[s];
            const int le = e - l_tr[li], m1 = le / nt1, a = le % nt1;
            int i1, j1;
            if (F == 0) { i1 = 0; j1 = m1; }
            else if (F == 2) { i1 = nx1 - 1; j1 = m1; }
            else if (F == 1) { i1 = m1; j1 = 0; }
            else { i1 = m1; j1 = ny1 - 1; }
            const int bd1 = restrict_ ? m.cell_bdry[s] : 0;
            A v = A(x[m.cell_dof_off[s] + (i1 * ny1 + j1) * nt1 + a]);
            if (bd1 && f_on_bdry(bd1, i1, j1, nx1, ny1)) v = A(0);
            tr[e] = v;
        }
    }
    __syncthreads();

    // ---- phase 2: cell-local terms
    for (int t = tid; t < ncd; t += nthr) {
        const int kk = f_find(cst, ncells, t);
        const int nt = cnt[kk];
        const int lt = t - cst[kk], nyt = ny * nt;
        const int p = lt / nyt, q = (lt / nt) % ny, r = lt % nt;
        const A dth = cdth[kk], wr = cwt[kk * FNM + r];
        const A *xc = xs + cst[kk];
        A acc = A(0);
        if (ops & RTX_OP_MASS)
            acc += A(dx * dy * 0.125) * dth * wxs[p] * wys[q] * wr * kap[p * ny + q] * xc[lt];
        if (ops & RTX_OP_QUADW) acc += A(dx * dy * 0.125) * dth * wxs[p] * wys[q] * wr * xc[lt];
        if (ops & RTX_OP_ICONV) {
            A sx = A(0), sy = A(0);
            for (int i = 0; i < nx; ++i) sx += WDx[p * FNM + i] * xc[i * nyt + q * nt + r];
            for (int j = 0; j < ny; ++j) sy += WDy[q * FNM + j] * xc[p * nyt + j * nt + r];
            acc -= A(dy * 0.25) * dth * wys[q] * wr * cth[kk * FNM + r] * sx +
                   A(dx * 0.25) * dth * wxs[p] * wr * sth[kk * FNM + r] * sy;
        }
        const int quad = m.cell_quad[k0 + kk];
        if ((ops & RTX_OP_BCONV) && quad >= 0) {
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int F = (quad + h) & 3;
                const bool on = (F == 0 && p == nx - 1) || (F == 2 && p == 0) ||
                                (F == 1 && q == ny - 1) || (F == 3 && q == 0);
                if (on) {
                    const A c_ = cth[kk * FNM + r], s_ = sth[kk * FNM + r];
                    const A tf = F == 0 ? c_ : (F == 1 ? s_ : (F == 2 ? -c_ : -s_));
                    acc += A(((F & 1) ? dx : dy) * 0.25) * dth * ((F & 1) ? wxs[p] : wys[q]) * wr *
                           tf * xc[lt];
                }
            }
        }
        ys[t] = acc;
    }
    __syncthreads();

    // ---- phase 3: upwind inflow couplings, one (link, face node) per thread
    if (ops & RTX_OP_BCONV) {
        for (int e = tid; e < l_it[nl]; e += nthr) {
            const int li = f_find(l_it, nl, e), l = L0 + li;
            const int s = m.link_src[l], F = m.link_face[l];
            const int kk = m.link_target[l] - k0;
            const int nt = cnt[kk], nyt = ny * nt;
            const int c1 = m.cell_col[s];
            const int nt1 = m.cell_nth[s];
            const int n1 = (F & 1) ? m.col_nx[c1] : m.col_ny[c1];
            const int n0 = (F & 1) ? nx : ny;
            const int le = e - l_it[li], m0 = le / nt, r = le % nt;
            const double *Es = m.Es_pool + m.link_Es_off[l];
            const double *Et = m.Eth_pool + m.link_Eth_off[l];
            A et[FNM], es[FNM];
#pragma unroll
            for (int a = 0; a < FNM; ++a) {
                et[a] = a < nt1 ? A(Et[a * nt + r]) : A(0);
                es[a] = a < n1 ? A(Es[a * n0 + m0]) : A(0);
            }
            const A *trl = tr + l_tr[li];
            A acc = A(0);
            for (int m1 = 0; m1 < n1; ++m1) {
                A inner = A(0);
#pragma unroll
                for (int a = 0; a < FNM; ++a)
                    if (a < nt1) inner += et[a] * trl[m1 * nt1 + a];
                acc += es[m1] * inner;
            }
            int node;
            if (F == 0) node = (nx - 1) * nyt + m0 * nt + r;
            else if (F == 2) node = m0 * nt + r;
            else if (F == 1) node = m0 * nyt + (ny - 1) * nt + r;
            else node = m0 * nyt + r;
            atomicAdd(&ys[cst[kk] + node], A(((F & 1) ? dx : dy) * 0.25) * cdth[kk] * acc);
        }
    }
    __syncthreads();

    // ---- phase 4: scattering on MFMA
    if (ops & RTX_OP_SCAT) {
        const int MT = (nsp + 15) / 16;
        const int NTt = (na + 15) / 16;
        const int l16 = lane & 15, lg = lane >> 4;
        for (int ntile = wave; ntile < NTt; ntile += nwaves) {
            const int R = ntile * 16 + l16;
            if constexpr (PREC == RTX_PREC_F64 || PREC == RTX_PREC_F32) {
                typedef typename std::conditional<PREC == RTX_PREC_F64, dbl4, flt4>::type acc4;
                typedef typename std::conditional<PREC == RTX_PREC_F64, double, float>::type E;
                const E *G = (PREC == RTX_PREC_F64) ? (const E *)(m.G + m.col_G_off[c])
                                                    : (const E *)(m.G_f + m.col_G_off[c]);
                acc4 acc[FMAX_MT];
#pragma unroll
                for (int t = 0; t < FMAX_MT; ++t) acc[t] = acc4{0, 0, 0, 0};
                const int nks = (na + 3) / 4;
                for (int ks0 = 0; ks0 < nks; ks0 += 4) {
                    E bb[4];
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        const int Ak = (ks0 + u) * 4 + lg;
                        bb[u] = (Ak < na && R < na) ? G[size_t(Ak) * na + R] : E(0);
                    }
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        const int Ak = (ks0 + u) * 4 + lg;
                        const bool kok = Ak < na;
                        const int base = kok ? a_base[Ak] : 0, st = kok ? a_nt[Ak] : 0,
                                  loc = kok ? a_loc[Ak] : 0;
#pragma unroll
                        for (int t = 0; t < FMAX_MT; ++t) {
                            if (t < MT) {
                                const int s = t * 16 + l16;
                                const E a = (kok && s < nsp) ? E(xs[base + s * st + loc]) : E(0);
                                if constexpr (PREC == RTX_PREC_F64)
                                    acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, bb[u], acc[t], 0, 0, 0);
                                else
                                    acc[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, bb[u], acc[t], 0, 0, 0);
                            }
                        }
                    }
                }
                if (R < na) {
                    const int ob = a_base[R], ost = a_nt[R], oloc = a_loc[R];
#pragma unroll
                    for (int t = 0; t < FMAX_MT; ++t) {
                        if (t < MT) {
#pragma unroll
                            for (int v = 0; v < 4; ++v) {
                                // f64 D row = l/16 + 4 v; f32 D row = 4 (l/16) + v
                                const int s = t * 16 + ((PREC == RTX_PREC_F64) ? (lg + 4 * v) : (4 * lg + v));
                                if (s < nsp) ys[ob + s * ost + oloc] -= sgw[s] * A(acc[t][v]);
                            }
                        }
                    }
                }
            } else {   // MIXED: BF16 operands, FP32 accumulation
                const uint16_t *G = m.G_bf16 + m.col_G_off[c];
                flt4 acc[FMAX_MT];
#pragma unroll
                for (int t = 0; t < FMAX_MT; ++t) acc[t] = flt4{0, 0, 0, 0};
                for (int ks = 0; ks < (na + 15) / 16; ++ks) {
                    sht4 b;
                    int base[4], st[4], loc[4];
                    bool kok[4];
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        const int Ak = ks * 16 + 4 * lg + u;
                        kok[u] = Ak < na;
                        b[u] = (kok[u] && R < na) ? short(G[size_t(Ak) * na + R]) : short(0);
                        base[u] = kok[u] ? a_base[Ak] : 0;
                        st[u] = kok[u] ? a_nt[Ak] : 0;
                        loc[u] = kok[u] ? a_loc[Ak] : 0;
                    }
#pragma unroll
                    for (int t = 0; t < FMAX_MT; ++t) {
                        if (t < MT) {
                            const int s = t * 16 + l16;
                            sht4 a;
#pragma unroll
                            for (int u = 0; u < 4; ++u)
                                a[u] = (kok[u] && s < nsp)
                                           ? short(f_bf16(float(xs[base[u] + s * st[u] + loc[u]])))
                                           : short(0);
                            acc[t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc[t], 0, 0, 0);
                        }
                    }
                }
                if (R < na) {
                    const int ob = a_base[R], ost = a_nt[R], oloc = a_loc[R];
#pragma unroll
                    for (int t = 0; t < FMAX_MT; ++t) {
                        if (t < MT) {
#pragma unroll
                            for (int v = 0; v < 4; ++v) {
                                const int s = t * 16 + 4 * lg + v;
                                if (s < nsp) ys[ob + s * ost + oloc] -= sgw[s] * A(acc[t][v]);
                            }
                        }
                    }
                }
            }
        }
    }
    __syncthreads();

    // ---- phase 5: write back
    for (int t = tid; t < ncd; t += nthr) {
        const int kk = f_find(cst, ncells, t);
        bool masked = false;
        if (cbd[kk]) {
            const int lt = t - cst[kk], nt = cnt[kk];
            masked = f_on_bdry(cbd[kk], lt / (ny * nt), (lt / nt) % ny, nx, ny);
        }
        const A v = masked ? A(0) : ys[t];
        if (accumulate) {
            if (!masked) y[cdof + t] = T(A(y[cdof + t]) + v);
        } else {
            y[cdof + t] = T(v);
        }
    }
}

// Dynamic LDS of the fused kernel: column values and accumulators, neighbour
// traces, per-angle tables.
size_t fused_lds_bytes(int ncd, int na, int ntrace, int prec) {
    const size_t a = (prec == RTX_PREC_F64) ? sizeof(double) : sizeof(float);
    return (2 * size_t(ncd) + size_t(ntrace)) * a + 3 * size_t(na) * sizeof(int);
}

int launch_fused(rtx_plan *p, uint32_t ops, int prec, const void *dx, void *dy, hipStream_t st,
                 int accumulate) {
    const size_t lds = fused_lds_bytes(p->max_col_dofs, p->max_nang, p->max_col_trace, prec);
    // enough wavefronts for the output-angle tiles, and at least one thread per
    // link of a column (the link prefix sums are one thread per link)
    int waves = (p->max_nang + 15) / 16;
    const int wl = (p->max_col_links + 63) / 64;
    if (waves < wl) waves = wl;
    if (waves < 2) waves = 2;
    if (waves > 8) waves = 8;
    RTX_CHECK(p->max_col_links <= 64 * waves, "fused apply: %d links in a column",
              p->max_col_links);
    const int nthr = 64 * waves;
    if (prec == RTX_PREC_F64) {
        RTX_HIP(hipFuncSetAttribute((const void *)k_fused<double, double, RTX_PREC_F64>,
                                    hipFuncAttributeMaxDynamicSharedMemorySize, int(lds)));
        k_fused<double, double, RTX_PREC_F64><<<p->m.ncol, nthr, lds, st>>>(
            p->m, ops, accumulate, (const double *)dx, (double *)dy);
    } else if (prec == RTX_PREC_F32) {
        RTX_HIP(hipFuncSetAttribute((const void *)k_fused<float, float, RTX_PREC_F32>,
                                    hipFuncAttributeMaxDynamicSharedMemorySize, int(lds)));
        k_fused<float, float, RTX_PREC_F32><<<p->m.ncol, nthr, lds, st>>>(
            p->m, ops, accumulate, (const float *)dx, (float *)dy);
    } else {
        RTX_HIP(hipFuncSetAttribute((const void *)k_fused<float, float, RTX_PREC_MIXED>,
                                    hipFuncAttributeMaxDynamicSharedMemorySize, int(lds)));
        k_fused<float, float, RTX_PREC_MIXED><<<p->m.ncol, nthr, lds, st>>>(
            p->m, ops, accumulate, (const float *)dx, (float *)dy);
    }
    RTX_HIP(hipGetLastError());
    return RTX_OK;
}

}  // namespace rtx
