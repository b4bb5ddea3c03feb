// Fused per-column transport operator (the fast path of rtx_apply).
//
// The DOFs of a spatial column are contiguous in the reference's global
// order (columns by key, cells by key, then beta = nth ny i + nth j + a).
// Each column is processed by H workgroups ("parts"); part h owns the
// angular cells [h ncells / H, (h + 1) ncells / H) of the column, i.e. a
// contiguous range of output DOFs and output angles.  Every part stages the
// whole column's input in LDS (the scattering integral needs every input
// angle), but accumulates only its own outputs, so LDS per workgroup stays
// small enough for several workgroups per CU.  The parts of a column are
// given block indices congruent modulo 8 so they run on the same XCD and
// share its L2 for the column input.
//
//   set-up   per-column constants and tables -> LDS (batched global loads)
//   phase 1  x (whole column) and the upwind neighbour traces of the part's
//            links -> LDS (inflow-boundary mask applied)
//   phase 2  mass + (-IC) + outflow faces on the part's DOFs
//            (mass_matrix.py:140-173, interior_convection_matrix.py:161-243,
//            boundary_convection_matrix.py:284-346)
//   phase 3  upwind inflow couplings, one (link, face node) per thread, LDS
//            atomics for faces shared by two links
//            (boundary_convection_matrix.py:348-524)
//   phase 4  scattering: per spatial node s, y[s, R] -= dx dy w w sigma_s
//            sum_A G^T[A, R] x[s, A] for the part's output angles R, as an
//            MFMA GEMM with M = spatial nodes, N = output angles, K = input
//            angles (scattering_matrix.py:142-237):
//              F64   -> v_mfma_f64_16x16x4f64   (x and G in FP64)
//              F32   -> v_mfma_f32_16x16x4f32   (x and G in FP32)
//              MIXED -> v_mfma_f32_16x16x16bf16 (x and G in BF16, FP32 accum.)
//   phase 5  LDS -> y for the part's DOFs (mask, optional accumulate)
#include <algorithm>
#include <type_traits>

#include "rtx_internal.h"

namespace rtx {

typedef double dbl4 __attribute__((ext_vector_type(4)));
typedef float flt4 __attribute__((ext_vector_type(4)));
typedef short sht4 __attribute__((ext_vector_type(4)));

constexpr int FNM = RTX_NMAX;
constexpr int FMAX_MT = 7;        // spatial-node tiles of 16 (nx * ny <= 112)
constexpr int FMAX_CELLS = 64;    // cells per column on this path
constexpr int FMAX_LINKS = 256;   // upwind links per column part on this path
constexpr int FMAX_NSP = FMAX_MT * 16;
constexpr int FXCD = 8;           // XCDs; block b runs on XCD b % 8
constexpr int FTHREADS = 256;

__device__ inline bool f_on_bdry(int bd, int i, int j, int nx, int ny) {
    return ((bd & 1) && i == nx - 1) || ((bd & 2) && j == ny - 1) || ((bd & 4) && i == 0) ||
           ((bd & 8) && j == 0);
}

__device__ inline uint16_t f_bf16(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);
    return uint16_t(u >> 16);
}

// Largest i in [lo, hi) with starts[i] <= t (starts ascending, starts[lo] <= t).
__device__ inline int f_find(const int *starts, int lo, int hi, int t) {
    hi -= 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (starts[mid] <= t) lo = mid; else hi = mid - 1;
    }
    return lo;
}

// A: arithmetic type of the transport terms and of the LDS accumulators.
// T: storage type of x and y.  PREC: RTX_PREC_*.  H: parts per column.
template <typename T, typename A, int PREC>
__global__ __launch_bounds__(FTHREADS) void k_fused(DevMesh m, uint32_t ops, int accumulate,
                                                   int H, const int32_t *__restrict__ cols,
                                                   int ncols, const T *__restrict__ x,
                                                   T *__restrict__ y) {
    extern __shared__ __align__(16) unsigned char f_lds[];
    __shared__ int cst[FMAX_CELLS + 1];      // column-local cell starts
    __shared__ int cbd[FMAX_CELLS];          // masked faces per cell
    __shared__ int cnt[FMAX_CELLS];          // nth per cell
    __shared__ int cao[FMAX_CELLS + 1];      // first column angle of each cell
    __shared__ int cqd[FMAX_CELLS];          // quadrant per cell (-1: none)
    __shared__ A cdth[FMAX_CELLS];
    __shared__ A wxs[FNM], wys[FNM], WDx[FNM * FNM], WDy[FNM * FNM];
    __shared__ A kap[FMAX_NSP], sgw[FMAX_NSP];
    __shared__ int lm0[FMAX_LINKS], lm1[FMAX_LINKS], l_tr[FMAX_LINKS + 1], l_it[FMAX_LINKS + 1];
    __shared__ int l_es[FMAX_LINKS], l_et[FMAX_LINKS];

    // XCD-aware block -> (column of the bin, part): the H parts of a column
    // share b % 8
    const int b = blockIdx.x;
    const int grp = b / (FXCD * H), rem = b % (FXCD * H);
    const int h = rem / FXCD, ci = grp * FXCD + rem % FXCD;
    if (ci >= ncols) return;
    const int c = cols[ci];
    const int tid = threadIdx.x, nthr = blockDim.x;
    const int lane = tid & 63, wave = tid >> 6, nwaves = nthr >> 6;
    const int k0 = m.col_cell_start[c], ncells = m.col_cell_start[c + 1] - k0;
    const int kk_lo = (h * ncells) / H, kk_hi = ((h + 1) * ncells) / H;
    if (kk_lo >= kk_hi) return;
    const int nx = m.col_nx[c], ny = m.col_ny[c], nsp = nx * ny;
    const int na = m.col_nang[c];
    const int ncd = nsp * na;
    const int cdof = m.cell_dof_off[k0];
    const int Lc = m.link_start[k0];                       // first link of the column
    const int L0 = m.link_start[k0 + kk_lo], nl = m.link_start[k0 + kk_hi] - L0;
    const int ga0 = m.col_ang_start[c];
    const bool restrict_ = (ops & RTX_OP_RESTRICT) != 0;
    const bool bconv = (ops & RTX_OP_BCONV) != 0;
    const double dx = m.col_geom[4 * c + 2] - m.col_geom[4 * c + 0];
    const double dy = m.col_geom[4 * c + 3] - m.col_geom[4 * c + 1];

    // ---- batched loads: column input, set-up items
    constexpr int XU = 8;
    A xv[XU];
#pragma unroll
    for (int u = 0; u < XU; ++u) {
        const int t = tid + u * nthr;
        xv[u] = (t < ncd) ? A(x[cdof + t]) : A(0);
    }
    int r_cst = 0, r_bd = 0, r_nt = 0, r_ao = 0, r_q = -1, r_pk = 0;
    double r_g0 = 0.0, r_g1 = 0.0, r_c = 0.0, r_s = 0.0, r_w = 0.0, r_kap = 0.0, r_sig = 0.0;
    double r_wd = 0.0, r_wdy = 0.0;
    int r_lm0 = 0, r_lm1 = 0, r_ltr = 0, r_lit = 0, r_les = 0, r_let = 0;
    if (tid < ncells) {
        const int k = k0 + tid;
        r_cst = m.cell_dof_off[k];
        r_bd = m.cell_bdry[k];
        r_nt = m.cell_nth[k];
        r_ao = m.cell_ang_off[k];
        r_q = m.cell_quad[k];
        r_g0 = m.cell_geom[2 * k];
        r_g1 = m.cell_geom[2 * k + 1];
    }
    if (tid < na) {
        r_pk = m.ang_pack[ga0 + tid];
        r_c = m.ang_trig[3 * (ga0 + tid)];
        r_s = m.ang_trig[3 * (ga0 + tid) + 1];
        r_w = m.ang_trig[3 * (ga0 + tid) + 2];
    }
    if (tid < nsp) {
        r_kap = m.kappa[m.col_sp_off[c] + tid];
        r_sig = m.sigma[m.col_sp_off[c] + tid] * m.lgl_w[nx * FNM + tid / ny] *
                m.lgl_w[ny * FNM + tid % ny];
    }
    const int wp = tid / FNM, wi = tid % FNM;
    if (tid < FNM * FNM) {
        if (wp < nx && wi < nx) r_wd = m.lgl_w[nx * FNM + wi] * m.lgl_D[(size_t(nx) * FNM + wp) * FNM + wi];
        if (wp < ny && wi < ny) r_wdy = m.lgl_w[ny * FNM + wi] * m.lgl_D[(size_t(ny) * FNM + wp) * FNM + wi];
    }
    if (bconv && tid < nl) {
        const int l = L0 + tid;
        r_lm0 = m.link_meta[4 * l + 0];
        r_lm1 = m.link_meta[4 * l + 1];
        r_ltr = m.link_meta[4 * l + 2];
        r_lit = m.link_meta[4 * l + 3];
        r_les = m.link_Es_off[l];
        r_let = m.link_Eth_off[l];
    }
    // column-local offsets of the part's first trace and face item
    int tr0 = 0, it0 = 0, ntr_own = 0, nit_own = 0;
    if (bconv && nl > 0) {
        tr0 = m.link_meta[4 * L0 + 2];
        it0 = m.link_meta[4 * L0 + 3];
        const int Lend = m.link_start[k0 + kk_hi];
        const int Lcol_end = m.link_start[k0 + ncells];
        const int tr_end = (Lend < Lcol_end) ? m.link_meta[4 * Lend + 2] : m.col_ntrace[c];
        const int it_end = (Lend < Lcol_end) ? m.link_meta[4 * Lend + 3] : m.col_nitems[c];
        ntr_own = tr_end - tr0;
        nit_own = it_end - it0;
    }
    (void)Lc;

    // ---- LDS layout
    A *xs = reinterpret_cast<A *>(f_lds);          // ncd (whole column)
    A *cth = xs + ncd;                             // na
    A *sth = cth + na;                             // na
    A *cw = sth + na;                              // na
    int *apk = reinterpret_cast<int *>(cw + na);   // na (int)
    A *ys = reinterpret_cast<A *>(f_lds) + ncd + 3 * na + (na + (sizeof(A) / sizeof(int)) - 1) / (sizeof(A) / sizeof(int));
    // ys and tr follow, sized for this part
    if (tid < ncells) {
        cst[tid] = r_cst - cdof;
        cbd[tid] = restrict_ ? r_bd : 0;
        cnt[tid] = r_nt;
        cao[tid] = r_ao;
        cqd[tid] = r_q;
        cdth[tid] = A(r_g1 - r_g0);
    }
    if (tid == 0) {
        cst[ncells] = ncd;
        cao[ncells] = na;
    }
    if (tid < na) {
        apk[tid] = r_pk;
        cth[tid] = A(r_c);
        sth[tid] = A(r_s);
        cw[tid] = A(r_w);
    }
    if (tid < nsp) {
        kap[tid] = A(r_kap);
        sgw[tid] = A(dx * dy * r_sig);
    }
    if (tid < FNM * FNM) {
        WDx[tid] = A(r_wd);
        WDy[tid] = A(r_wdy);
    }
    if (tid < FNM) {
        wxs[tid] = A(tid < nx ? m.lgl_w[nx * FNM + tid] : 0.0);
        wys[tid] = A(tid < ny ? m.lgl_w[ny * FNM + tid] : 0.0);
    }
    if (bconv) {
        if (tid < nl) {
            lm0[tid] = r_lm0;
            lm1[tid] = r_lm1;
            l_tr[tid] = r_ltr - tr0;
            l_it[tid] = r_lit - it0;
            l_es[tid] = r_les;
            l_et[tid] = r_let;
        }
        if (tid == 0) {
            l_tr[nl] = ntr_own;
            l_it[nl] = nit_own;
        }
        for (int li = tid + nthr; li < nl; li += nthr) {
            const int l = L0 + li;
            lm0[li] = m.link_meta[4 * l + 0];
            lm1[li] = m.link_meta[4 * l + 1];
            l_tr[li] = m.link_meta[4 * l + 2] - tr0;
            l_it[li] = m.link_meta[4 * l + 3] - it0;
            l_es[li] = m.link_Es_off[l];
            l_et[li] = m.link_Eth_off[l];
        }
    }
    for (int a = tid + nthr; a < na; a += nthr) {
        apk[a] = m.ang_pack[ga0 + a];
        cth[a] = A(m.ang_trig[3 * (ga0 + a)]);
        sth[a] = A(m.ang_trig[3 * (ga0 + a) + 1]);
        cw[a] = A(m.ang_trig[3 * (ga0 + a) + 2]);
    }
    __syncthreads();
    const int own0 = cst[kk_lo], nown = cst[kk_hi] - own0;   // the part's DOFs
    const int oa0 = cao[kk_lo], noa = cao[kk_hi] - oa0;      // the part's angles
    A *tr = ys + nown;

    // ---- phase 1: column input (masked) and the part's neighbour traces
#pragma unroll
    for (int u = 0; u < XU; ++u) {
        const int t = tid + u * nthr;
        if (t < ncd) {
            A v = xv[u];
            if (restrict_) {
                const int kk = f_find(cst, 0, ncells, t);
                if (cbd[kk]) {
                    const int lt = t - cst[kk], nt = cnt[kk];
                    if (f_on_bdry(cbd[kk], lt / (ny * nt), (lt / nt) % ny, nx, ny)) v = A(0);
                }
            }
            xs[t] = v;
        }
    }
    for (int t = tid + XU * nthr; t < ncd; t += nthr) {
        A v = A(x[cdof + t]);
        if (restrict_) {
            const int kk = f_find(cst, 0, ncells, t);
            if (cbd[kk]) {
                const int lt = t - cst[kk], nt = cnt[kk];
                if (f_on_bdry(cbd[kk], lt / (ny * nt), (lt / nt) % ny, nx, ny)) v = A(0);
            }
        }
        xs[t] = v;
    }
    if (bconv) {
        constexpr int TU = 4;   // trace values per thread in flight
        for (int e0 = tid; e0 < ntr_own; e0 += TU * nthr) {
            int addr[TU];
            bool zero[TU];
            A v[TU];
#pragma unroll
            for (int u = 0; u < TU; ++u) {
                const int e = e0 + u * nthr;
                addr[u] = -1;
                zero[u] = false;
                if (e < ntr_own) {
                    const int li = f_find(l_tr, 0, nl, e);
                    const int w1 = lm1[li];
                    const int F = w1 & 3, nx1 = (w1 >> 2) & 15, ny1 = (w1 >> 6) & 15;
                    const int nt1 = (w1 >> 10) & 15, bd1 = restrict_ ? ((w1 >> 14) & 15) : 0;
                    const int le = e - l_tr[li], m1 = le / nt1, a = le % nt1;
                    int i1, j1;
                    if (F == 0) { i1 = 0; j1 = m1; }
                    else if (F == 2) { i1 = nx1 - 1; j1 = m1; }
                    else if (F == 1) { i1 = m1; j1 = 0; }
                    else { i1 = m1; j1 = ny1 - 1; }
                    addr[u] = lm0[li] + (i1 * ny1 + j1) * nt1 + a;
                    zero[u] = bd1 && f_on_bdry(bd1, i1, j1, nx1, ny1);
                }
            }
#pragma unroll
            for (int u = 0; u < TU; ++u) v[u] = (addr[u] >= 0) ? A(x[addr[u]]) : A(0);
#pragma unroll
            for (int u = 0; u < TU; ++u) {
                const int e = e0 + u * nthr;
                if (e < ntr_own) tr[e] = zero[u] ? A(0) : v[u];
            }
        }
    }
    __syncthreads();

    // ---- phase 2: cell-local terms on the part's DOFs
    for (int e = tid; e < nown; e += nthr) {
        const int t = own0 + e;
        const int kk = f_find(cst, kk_lo, kk_hi, t);
        const int nt = cnt[kk];
        const int lt = t - cst[kk], nyt = ny * nt;
        const int p = lt / nyt, q = (lt / nt) % ny, r = lt % nt;
        const int ca = cao[kk] + r;
        const A dth = cdth[kk], wr = cw[ca];
        const A *xc = xs + cst[kk];
        A acc = A(0);
        if (ops & RTX_OP_MASS)
            acc += A(dx * dy * 0.125) * dth * wxs[p] * wys[q] * wr * kap[p * ny + q] * xc[lt];
        if (ops & RTX_OP_QUADW) acc += A(dx * dy * 0.125) * dth * wxs[p] * wys[q] * wr * xc[lt];
        if (ops & RTX_OP_ICONV) {
            A sx = A(0), sy = A(0);
            for (int i = 0; i < nx; ++i) sx += WDx[p * FNM + i] * xc[i * nyt + q * nt + r];
            for (int j = 0; j < ny; ++j) sy += WDy[q * FNM + j] * xc[p * nyt + j * nt + r];
            acc -= A(dy * 0.25) * dth * wys[q] * wr * cth[ca] * sx +
                   A(dx * 0.25) * dth * wxs[p] * wr * sth[ca] * sy;
        }
        const int quad = cqd[kk];
        if (bconv && quad >= 0) {
#pragma unroll
            for (int hh = 0; hh < 2; ++hh) {
                const int F = (quad + hh) & 3;
                const bool on = (F == 0 && p == nx - 1) || (F == 2 && p == 0) ||
                                (F == 1 && q == ny - 1) || (F == 3 && q == 0);
                if (on) {
                    const A c_ = cth[ca], s_ = sth[ca];
                    const A tf = F == 0 ? c_ : (F == 1 ? s_ : (F == 2 ? -c_ : -s_));
                    acc += A(((F & 1) ? dx : dy) * 0.25) * dth * ((F & 1) ? wxs[p] : wys[q]) * wr *
                           tf * xc[lt];
                }
            }
        }
        ys[e] = acc;
    }
    __syncthreads();

    // ---- phase 3: upwind inflow couplings, one (link, face node) per thread
    if (bconv) {
        for (int e = tid; e < nit_own; e += nthr) {
            const int li = f_find(l_it, 0, nl, e);
            const int w1 = lm1[li];
            const int F = w1 & 3, nt1 = (w1 >> 10) & 15, kk = w1 >> 18;
            const int nt = cnt[kk], nyt = ny * nt;
            const int n1 = (F & 1) ? ((w1 >> 2) & 15) : ((w1 >> 6) & 15);
            const int n0 = (F & 1) ? nx : ny;
            const int le = e - l_it[li], m0 = le / nt, r = le % nt;
            const double *Es = m.Es_pool + l_es[li];
            const double *Et = m.Eth_pool + l_et[li];
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
            atomicAdd(&ys[cst[kk] - own0 + node], A(((F & 1) ? dx : dy) * 0.25) * cdth[kk] * acc);
        }
    }
    __syncthreads();

    // ---- phase 4: scattering on MFMA for the part's output angles
    if (ops & RTX_OP_SCAT) {
        const int MT = (nsp + 15) / 16;
        const int NTt = (noa + 15) / 16;
        const int l16 = lane & 15, lg = lane >> 4;
        for (int ntile = wave; ntile < NTt; ntile += nwaves) {
            const int Rl = ntile * 16 + l16;          // part-local output angle
            const int R = oa0 + Rl;                   // column angle
            const bool rok = Rl < noa;
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
                        bb[u] = (Ak < na && rok) ? G[size_t(Ak) * na + R] : E(0);
                    }
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        if (ks0 + u < nks) {
                            const int Ak = (ks0 + u) * 4 + lg;
                            const bool kok = Ak < na;
                            const int pk = kok ? apk[Ak] : 0;
                            const int base = pk & 0xFFFFF, st = (pk >> 20) & 31, loc = pk >> 25;
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
                }
                if (rok) {
                    const int opk = apk[R];
                    const int ob = (opk & 0xFFFFF) - own0, ost = (opk >> 20) & 31, oloc = opk >> 25;
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
                    sht4 bv;
                    int base[4], st[4], loc[4];
                    bool kok[4];
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        const int Ak = ks * 16 + 4 * lg + u;
                        kok[u] = Ak < na;
                        bv[u] = (kok[u] && rok) ? short(G[size_t(Ak) * na + R]) : short(0);
                        const int pk = kok[u] ? apk[Ak] : 0;
                        base[u] = pk & 0xFFFFF;
                        st[u] = (pk >> 20) & 31;
                        loc[u] = pk >> 25;
                    }
#pragma unroll
                    for (int t = 0; t < FMAX_MT; ++t) {
                        if (t < MT) {
                            const int s = t * 16 + l16;
                            sht4 av;
#pragma unroll
                            for (int u = 0; u < 4; ++u)
                                av[u] = (kok[u] && s < nsp)
                                            ? short(f_bf16(float(xs[base[u] + s * st[u] + loc[u]])))
                                            : short(0);
                            acc[t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av, bv, acc[t], 0, 0, 0);
                        }
                    }
                }
                if (rok) {
                    const int opk = apk[R];
                    const int ob = (opk & 0xFFFFF) - own0, ost = (opk >> 20) & 31, oloc = opk >> 25;
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

    // ---- phase 5: write back the part's DOFs
    for (int e = tid; e < nown; e += nthr) {
        const int t = own0 + e;
        bool masked = false;
        if (restrict_) {
            const int kk = f_find(cst, kk_lo, kk_hi, t);
            if (cbd[kk]) {
                const int lt = t - cst[kk], nt = cnt[kk];
                masked = f_on_bdry(cbd[kk], lt / (ny * nt), (lt / nt) % ny, nx, ny);
            }
        }
        const A v = masked ? A(0) : ys[e];
        if (accumulate) {
            if (!masked) y[cdof + t] = T(A(y[cdof + t]) + v);
        } else {
            y[cdof + t] = T(v);
        }
    }
}

// Dynamic LDS of one part: whole-column input, per-angle tables, the part's
// accumulators and traces.
size_t fused_bytes(int64_t col_dofs, int64_t nang, int64_t part_dofs, int64_t part_trace,
                   size_t a) {
    const size_t ints_as_a = (size_t(nang) + a / sizeof(int) - 1) / (a / sizeof(int));
    return (size_t(col_dofs) + 3 * size_t(nang) + ints_as_a + size_t(part_dofs) +
            size_t(part_trace)) * a;
}

constexpr size_t FUSED_CAP = 96 * 1024;   // dynamic LDS cap of the fused path

int fused_column_ok(int ncells, int nsp, int nang, int part_links, size_t bytes64) {
    return ncells <= FMAX_CELLS && nsp <= FMAX_NSP && part_links <= FMAX_LINKS &&
           nang <= 1024 && bytes64 <= FUSED_CAP;
}

static int h_index(int H) { return H == 1 ? 0 : (H == 2 ? 1 : 2); }

size_t fused_lds_bytes(const rtx_plan *p, int H, int prec) {
    // largest part over the bins (for reporting; launches size each bin)
    size_t mx = 0;
    for (const FusedBin &fb : p->fbin)
        if (fb.n > 0) mx = std::max(mx, fb.bytes[prec == RTX_PREC_F64 ? 0 : 1][h_index(H)]);
    return mx;
}

int fused_parts(const rtx_plan *p, int prec) {
    (void)prec;
    int H = 1;
    for (const FusedBin &fb : p->fbin)
        if (fb.n > 0) H = std::max(H, fb.H);
    return H;
}

int fused_bin_valid_parts(const rtx_plan *p, int bin, int prec) {
    const FusedBin &fb = p->fbin[bin];
    int mask = 0;
    const int Hs[3] = {1, 2, 4};
    for (int i = 0; i < 3; ++i)
        if (fb.n > 0 && Hs[i] <= fb.min_cells && fb.part_links[i] <= FMAX_LINKS &&
            fb.bytes[prec == RTX_PREC_F64 ? 0 : 1][i] <= FUSED_CAP)
            mask |= Hs[i];
    return mask;
}

template <typename T, typename A, int PREC>
static int launch_bin(const rtx_plan *p, const FusedBin &fb, uint32_t ops, const void *dx, void *dy,
                      hipStream_t st, int accumulate) {
    const int H = fb.H, hi = h_index(H);
    RTX_CHECK(fb.part_links[hi] <= FMAX_LINKS, "fused apply: %d links in a column part",
              fb.part_links[hi]);
    // each column carves its own part from the front of the dynamic LDS
    const size_t lds = fb.bytes[sizeof(A) == sizeof(double) ? 0 : 1][hi];
    RTX_CHECK(lds <= FUSED_CAP, "fused apply: %zu bytes of LDS per part", lds);
    const int nblocks = ((fb.n + FXCD - 1) / FXCD) * FXCD * H;
    RTX_HIP(hipFuncSetAttribute((const void *)k_fused<T, A, PREC>,
                                hipFuncAttributeMaxDynamicSharedMemorySize, int(lds)));
    k_fused<T, A, PREC><<<nblocks, FTHREADS, lds, st>>>(p->m, ops, accumulate, H, fb.cols, fb.n,
                                                        (const T *)dx, (T *)dy);
    RTX_HIP(hipGetLastError());
    return RTX_OK;
}

int launch_fused(rtx_plan *p, uint32_t ops, int prec, const void *dx, void *dy, hipStream_t st,
                 int accumulate, int only_bin, bool all_bins) {
    for (int bi = 0; bi < RTX_NBINS; ++bi) {
        const FusedBin &fb = p->fbin[bi];
        if (fb.n == 0 || (only_bin >= 0 && bi != only_bin)) continue;
        if (fb.generic && !all_bins && only_bin < 0) continue;
        int rc;
        if (prec == RTX_PREC_F64)
            rc = launch_bin<double, double, RTX_PREC_F64>(p, fb, ops, dx, dy, st, accumulate);
        else if (prec == RTX_PREC_F32)
            rc = launch_bin<float, float, RTX_PREC_F32>(p, fb, ops, dx, dy, st, accumulate);
        else
            rc = launch_bin<float, float, RTX_PREC_MIXED>(p, fb, ops, dx, dy, st, accumulate);
        if (rc) return rc;
    }
    return RTX_OK;
}

}  // namespace rtx
