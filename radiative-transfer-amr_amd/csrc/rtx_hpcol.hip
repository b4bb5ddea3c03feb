// Matrix-free DG transport operator for uneven-order hp meshes: one fused
// pass per spatial column.
//
// A column's DOFs are contiguous in the reference's global order (columns by
// key, cells by key, then beta = nth ny i + nth j + a inside a cell).  One
// workgroup of 512 threads owns one column:
//
//   phase 0  the column's cell, angle, node and link tables -> LDS
//   phase 1  x of the column -> LDS, node-major (row s = spatial node, one
//            entry per column angle; inflow-boundary DOFs zeroed), and, for
//            every upwind link, the neighbour's trace contracted with E^theta
//            over the neighbour's angles: inn[m1, r] = sum_a E^th[a, r] tr[m1, a]
//            (boundary_convection_matrix.py:348-524, the angular half)
//   phase 2  each wavefront takes tiles of 16 output angles R; for all node
//            tiles of the column it accumulates the scattering contraction
//            sum_A x[s, A] G^T[A, R] on the matrix cores
//            (scattering_matrix.py:142-237), then, in the accumulator layout,
//            adds for each of its (s, R) the mass (mass_matrix.py:140-173),
//            interior convection (interior_convection_matrix.py:161-243), the
//            outflow faces (boundary_convection_matrix.py:284-346) and the
//            upwind inflow faces (the spatial half, E^s applied to inn), and
//            stores y.
//
// Every output has one owner and a fixed order of additions, so results are
// bit-reproducible; x is read from HBM once (plus the neighbours' face
// traces), G once and y written once.  Columns are grouped by the LDS they
// need into two launches, so columns of up to about 70 KB run two
// workgroups per CU.  Arithmetic: FP64 (v_mfma_f64_16x16x4f64), FP32
// (v_mfma_f32_16x16x4f32) or mixed (FP32 transport terms, BF16 scattering
// operands on v_mfma_f32_16x16x16bf16_1k with FP32 accumulation).
#include <algorithm>
#include <type_traits>

#include "rtx_internal.h"

namespace rtx {

constexpr int HP_THREADS = 512;
constexpr int HP_WAVES = HP_THREADS / 64;
constexpr int HP_NM = RTX_NMAX;
constexpr int HP_MT = (RTX_NMAX * RTX_NMAX + 15) / 16;   // node tiles of 16
constexpr int HP_XCD = 8;
constexpr int HP_TG = 2;          // node tiles per scattering pass (accumulators in registers)

typedef double hp_d4 __attribute__((ext_vector_type(4)));
typedef float hp_f4 __attribute__((ext_vector_type(4)));
typedef short hp_s4 __attribute__((ext_vector_type(4)));

__device__ inline bool hp_on_bdry(int bd, int i, int j, int nx, int ny) {
    return ((bd & 1) && i == nx - 1) || ((bd & 2) && j == ny - 1) || ((bd & 4) && i == 0) ||
           ((bd & 8) && j == 0);
}

// floor(v / n) for 0 <= v < 2^16 and 1 <= n <= 16, by a float product: the
// quotient's fractional part stays >= 0.5 / n away from an integer.
__device__ inline int hp_div(int v, float rcp) { return int((float(v) + 0.5f) * rcp); }

__device__ inline uint16_t hp_bf16(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);
    return uint16_t(u >> 16);
}

// Largest i in [lo, hi) with starts[i] <= t (starts ascending, starts[lo] <= t).
__device__ inline int hp_find(const int *starts, int lo, int hi, int t) {
    hi -= 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (starts[mid] <= t) lo = mid; else hi = mid - 1;
    }
    return lo;
}

// Row stride (elements) of the node-major LDS copy: at least na rounded up
// to the K step, and congruent to c modulo M, so that the 16 rows x K lanes of
// one matrix-core operand read fall on distinct banks (8-byte elements, 16
// rows x 2 k per half-wave: stride/2 odd, c = 2, M = 4; 4-byte, 16 x 4:
// stride/4 odd, c = 4, M = 8; 2-byte BF16 quads, 16 x 2 per half-wave:
// stride/8 odd, c = 8, M = 16).
__host__ __device__ inline int hp_stride(int na, int kstep, int c, int M) {
    const int n = (na + kstep - 1) / kstep * kstep;
    return n + ((c - n) % M + M) % M;
}

struct HpLayout {           // dynamic LDS of one column, in bytes from the base
    int xs, xb, inn, cth, sth, wr, acell, total;
    int XS, XSB;
};

// T: storage of x and y; A: arithmetic and LDS element type; PREC: RTX_PREC_*.
template <typename A, int PREC>
__host__ __device__ inline HpLayout hp_layout(int nsp, int na, int ninn) {
    HpLayout L;
    const int ea = int(sizeof(A));
    L.XS = (ea == 8) ? hp_stride(na, 4, 2, 4) : hp_stride(na, PREC == RTX_PREC_MIXED ? 16 : 4, 4, 8);
    L.XSB = (PREC == RTX_PREC_MIXED) ? hp_stride(na, 16, 8, 16) : 0;
    int o = 0;
    L.xs = o;
    o += nsp * L.XS * ea;
    o = (o + 15) & ~15;
    L.xb = o;
    o += nsp * L.XSB * 2;
    o = (o + 15) & ~15;
    L.inn = o;
    o += ninn * ea;
    L.cth = o;
    o += na * ea;
    L.sth = o;
    o += na * ea;
    L.wr = o;
    o += na * ea;
    L.acell = o;
    o += na * 4;
    L.total = (o + 15) & ~15;
    return L;
}

template <typename T, typename A, int PREC>
__global__ __launch_bounds__(HP_THREADS) __attribute__((amdgpu_waves_per_eu(4)))
void k_hpcol(DevMesh m, HpTables h, uint32_t ops, int accumulate,
             const int32_t *__restrict__ cols, int ncols, const T *__restrict__ x,
             T *__restrict__ y) {
    extern __shared__ __align__(16) unsigned char hp_lds[];
    __shared__ int c_nth[HP_MAX_CELLS], c_bd[HP_MAX_CELLS], c_lf[HP_MAX_CELLS];
    __shared__ int c_quad[HP_MAX_CELLS], c_ao[HP_MAX_CELLS + 1], c_l0[HP_MAX_CELLS + 1];
    __shared__ A c_dth[HP_MAX_CELLS];
    __shared__ int l_w[HP_MAX_LINKS], l_inn[HP_MAX_LINKS + 1], l_es[HP_MAX_LINKS];
    __shared__ int n_pq[HP_NM * HP_NM];                 // node s -> p | q << 8
    __shared__ A w2k[HP_NM * HP_NM], sgw[HP_NM * HP_NM];
    __shared__ A wxs[HP_NM], wys[HP_NM], WDx[HP_NM * HP_NM], WDy[HP_NM * HP_NM];

    // XCD-aware block -> column: block b runs on XCD b % 8, and XCD j takes
    // the j-th contiguous run of the (key-ordered) list, so the neighbour
    // columns whose traces a column reads are mostly cached on the same XCD
    const int per = (ncols + HP_XCD - 1) / HP_XCD;
    const int bi = blockIdx.x % HP_XCD, bj = blockIdx.x / HP_XCD;
    if (bj >= per) return;
    const int ci = bi * per + bj;
    if (ci >= ncols) return;
    const int c = cols[ci];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int l16 = lane & 15, lg = lane >> 4;
    const int k0 = m.col_cell_start[c], ncells = m.col_cell_start[c + 1] - k0;
    const int nx = m.col_nx[c], ny = m.col_ny[c], nsp = nx * ny;
    const int na = m.col_nang[c];
    const int ncd = nsp * na;
    const int cdof = m.cell_dof_off[k0];
    const int L0 = m.link_start[k0], nl = m.link_start[k0 + ncells] - L0;
    const int ga0 = m.col_ang_start[c];
    const int ninn = h.col_inn[c];
    const bool restrict_ = (ops & RTX_OP_RESTRICT) != 0;
    const bool bconv = (ops & RTX_OP_BCONV) != 0;
    const double x0 = m.col_geom[4 * c + 0], y0 = m.col_geom[4 * c + 1];
    const double dx = m.col_geom[4 * c + 2] - x0, dy = m.col_geom[4 * c + 3] - y0;
    const float rnsp = 1.0f / float(nsp);
    const HpLayout Lo = hp_layout<A, PREC>(nsp, na, ninn);
    A *xs = reinterpret_cast<A *>(hp_lds + Lo.xs);
    uint16_t *xb = reinterpret_cast<uint16_t *>(hp_lds + Lo.xb);
    A *inn = reinterpret_cast<A *>(hp_lds + Lo.inn);
    A *cth = reinterpret_cast<A *>(hp_lds + Lo.cth);
    A *sth = reinterpret_cast<A *>(hp_lds + Lo.sth);
    A *wr = reinterpret_cast<A *>(hp_lds + Lo.wr);
    int *acell = reinterpret_cast<int *>(hp_lds + Lo.acell);
    const int XS = Lo.XS, XSB = Lo.XSB;

    // ---- phase 0: tables
    if (tid < ncells) {
        const int k = k0 + tid;
        const int nt = m.cell_nth[k], ao = m.cell_ang_off[k];
        const int lb = m.link_start[k], le = m.link_start[k + 1];
        c_nth[tid] = nt;
        c_bd[tid] = restrict_ ? m.cell_bdry[k] : 0;
        c_quad[tid] = m.cell_quad[k];
        c_ao[tid] = ao;
        c_l0[tid] = lb - L0;
        c_dth[tid] = A(m.cell_geom[2 * k + 1] - m.cell_geom[2 * k]);
        int lf = 0;
        for (int l = lb; l < le; ++l) lf |= 1 << m.link_face[l];
        c_lf[tid] = lf;
        for (int a = 0; a < nt; ++a) acell[ao + a] = tid | (a << 8);
    }
    if (tid == 0) {
        c_ao[ncells] = na;
        c_l0[ncells] = nl;
        l_inn[nl] = ninn;
    }
    for (int a = tid; a < na; a += HP_THREADS) {
        cth[a] = A(m.ang_trig[3 * (ga0 + a)]);
        sth[a] = A(m.ang_trig[3 * (ga0 + a) + 1]);
        wr[a] = A(m.ang_trig[3 * (ga0 + a) + 2]);
    }
    for (int li = tid; li < nl; li += HP_THREADS) {
        const int l = L0 + li;
        const int s = m.link_src[l], F = m.link_face[l], c1 = m.cell_col[s];
        const int n1 = (F & 1) ? m.col_nx[c1] : m.col_ny[c1];
        l_w[li] = F | (n1 << 2);
        l_inn[li] = h.link_inn[l];
        l_es[li] = m.link_Es_off[l];
    }
    if (tid < nsp) {
        const int i = tid / ny, j = tid - (tid / ny) * ny;
        const double w2 = m.lgl_w[nx * HP_NM + i] * m.lgl_w[ny * HP_NM + j];
        n_pq[tid] = i | (j << 8);
        w2k[tid] = A(w2 * m.kappa[m.col_sp_off[c] + tid]);
        sgw[tid] = A(dx * dy * m.sigma[m.col_sp_off[c] + tid] * w2);
    }
    if (tid < HP_NM * HP_NM) {
        const int pp = tid / HP_NM, ii = tid % HP_NM;
        WDx[tid] = (pp < nx && ii < nx)
                       ? A(m.lgl_w[nx * HP_NM + ii] * m.lgl_D[(size_t(nx) * HP_NM + pp) * HP_NM + ii])
                       : A(0);
        WDy[tid] = (pp < ny && ii < ny)
                       ? A(m.lgl_w[ny * HP_NM + ii] * m.lgl_D[(size_t(ny) * HP_NM + pp) * HP_NM + ii])
                       : A(0);
    }
    if (tid < HP_NM) {
        wxs[tid] = A(tid < nx ? m.lgl_w[nx * HP_NM + tid] : 0.0);
        wys[tid] = A(tid < ny ? m.lgl_w[ny * HP_NM + tid] : 0.0);
    }
    __syncthreads();

    // ---- phase 1: x (node-major, masked) and the angle-contracted traces;
    // loads are issued in batches so their latencies overlap.  DOF t of the
    // column lies in the cell of angle t / nsp (cell kk's DOFs start at
    // nsp * c_ao[kk]).
    constexpr int XU = 8;
    for (int t0 = tid; t0 < ncd; t0 += XU * HP_THREADS) {
        A vv[XU];
#pragma unroll
        for (int u = 0; u < XU; ++u) {
            const int t = t0 + u * HP_THREADS;
            vv[u] = (t < ncd) ? A(x[cdof + t]) : A(0);
        }
#pragma unroll
        for (int u = 0; u < XU; ++u) {
            const int t = t0 + u * HP_THREADS;
            if (t >= ncd) break;
            A v = vv[u];
            const int kk = acell[hp_div(t, rnsp)] & 255;
            const int nt = c_nth[kk], ao = c_ao[kk];
            const int lt = t - nsp * ao;
            const int s = hp_div(lt, 1.0f / float(nt)), a = lt - s * nt;
            if (c_bd[kk]) {
                const int pq = n_pq[s];
                if (hp_on_bdry(c_bd[kk], pq & 255, pq >> 8, nx, ny)) v = A(0);
            }
            xs[s * XS + ao + a] = v;
            if constexpr (PREC == RTX_PREC_MIXED) xb[s * XSB + ao + a] = hp_bf16(float(v));
        }
    }
    // zero padding columns [na, stride) of every row
    for (int t = tid; t < nsp * (XS - na); t += HP_THREADS) {
        const int s = t / (XS - na), a = na + t % (XS - na);
        xs[s * XS + a] = A(0);
    }
    if constexpr (PREC == RTX_PREC_MIXED) {
        for (int t = tid; t < nsp * (XSB - na); t += HP_THREADS) {
            const int s = t / (XSB - na), a = na + t % (XSB - na);
            xb[s * XSB + a] = 0;
        }
    }
    if (bconv) {
        for (int e = tid; e < ninn; e += HP_THREADS) {
            const int li = hp_find(l_inn, 0, nl, e);
            const int l = L0 + li;
            const int w1 = m.link_meta[4 * l + 1];
            const int F = w1 & 3, nx1 = (w1 >> 2) & 15, ny1 = (w1 >> 6) & 15;
            const int nt1 = (w1 >> 10) & 15, bd1 = restrict_ ? ((w1 >> 14) & 15) : 0;
            const int kk = w1 >> 18;
            const int nt = c_nth[kk];
            const int le = e - l_inn[li], m1 = hp_div(le, 1.0f / float(nt)), r = le - m1 * nt;
            int i1, j1;
            if (F == 0) { i1 = 0; j1 = m1; }
            else if (F == 2) { i1 = nx1 - 1; j1 = m1; }
            else if (F == 1) { i1 = m1; j1 = 0; }
            else { i1 = m1; j1 = ny1 - 1; }
            A acc = A(0);
            if (!(bd1 && hp_on_bdry(bd1, i1, j1, nx1, ny1))) {
                const T *tr = x + m.link_meta[4 * l] + (i1 * ny1 + j1) * nt1;
                const double *Et = m.Eth_pool + m.link_Eth_off[l] + r;
                for (int a = 0; a < nt1; ++a) acc += A(Et[a * nt]) * A(tr[a]);
            }
            inn[e] = acc;
        }
    }
    __syncthreads();

    // ---- phase 2: per tile of 16 output angles R and group of HP_TG node
    // tiles: the scattering contraction on the matrix cores, then the
    // transport terms of the same outputs in the accumulator layout
    const int MT = (nsp + 15) / 16;
    const int NRT = (na + 15) / 16;
    const int NTG = (MT + HP_TG - 1) / HP_TG;
    typedef typename std::conditional<PREC == RTX_PREC_F64, hp_d4, hp_f4>::type acc4;
    for (int item = wave; item < NRT * NTG; item += HP_WAVES) {
        const int rt = item / NTG, tg = (item - rt * NTG) * HP_TG;
        const int R = rt * 16 + l16;
        const bool rok = R < na;
        // lanes past the last angle read valid addresses; their output
        // column is never stored
        const int Rc = rok ? R : na - 1;
        acc4 acc[HP_TG];
#pragma unroll
        for (int t = 0; t < HP_TG; ++t) acc[t] = acc4{0, 0, 0, 0};
        if (ops & RTX_OP_SCAT) {
            // operand rows: node tg + t of this lane (rows past the last node
            // read the last node; their output rows are never stored)
            int arow[HP_TG];
#pragma unroll
            for (int t = 0; t < HP_TG; ++t) {
                const int s = (tg + t) * 16 + l16;
                arow[t] = (s < nsp ? s : nsp - 1);
            }
            if constexpr (PREC != RTX_PREC_MIXED) {
                typedef typename std::conditional<PREC == RTX_PREC_F64, double, float>::type E;
                const E *G = (PREC == RTX_PREC_F64) ? (const E *)(m.G + m.col_G_off[c])
                                                    : (const E *)(m.G_f + m.col_G_off[c]);
                const E *gcol = G + Rc;
                // G rows past na are clamped to the last row: the copy of x
                // is zero in those columns, so they add nothing
                auto gload = [&](int Ak) -> E { return gcol[size_t(Ak < na ? Ak : na - 1) * na]; };
                const A *xrow[HP_TG];
#pragma unroll
                for (int t = 0; t < HP_TG; ++t) xrow[t] = xs + arow[t] * XS + lg;
                // four K steps of G per load group; the next two groups are
                // in flight while the current one multiplies
                E bc[4], bn[4], bf[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    bc[u] = gload(4 * u + lg);
                    bn[u] = gload(16 + 4 * u + lg);
                }
                for (int kb = 0; kb < na; kb += 16) {
#pragma unroll
                    for (int u = 0; u < 4; ++u) bf[u] = gload(kb + 32 + 4 * u + lg);
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        if (kb + 4 * u < na) {            // uniform across the wave
#pragma unroll
                            for (int t = 0; t < HP_TG; ++t) {
                                const E a = E(xrow[t][kb + 4 * u]);
                                if constexpr (PREC == RTX_PREC_F64)
                                    acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, bc[u], acc[t], 0, 0, 0);
                                else
                                    acc[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, bc[u], acc[t], 0, 0, 0);
                            }
                        }
                    }
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        bc[u] = bn[u];
                        bn[u] = bf[u];
                    }
                }
            } else {
                const uint16_t *gcol = m.G_bf16 + m.col_G_off[c] + Rc;
                const uint16_t *xrow[HP_TG];
#pragma unroll
                for (int t = 0; t < HP_TG; ++t) xrow[t] = xb + arow[t] * XSB + 4 * lg;
                for (int kb = 0; kb < na; kb += 16) {
                    hp_s4 bv;
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        const int Ak = kb + 4 * lg + u;
                        bv[u] = short(gcol[size_t(Ak < na ? Ak : na - 1) * na]);
                    }
#pragma unroll
                    for (int t = 0; t < HP_TG; ++t) {
                        const hp_s4 av = *reinterpret_cast<const hp_s4 *>(xrow[t] + kb);
                        acc[t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av, bv, acc[t], 0, 0, 0);
                    }
                }
            }
        }
        if (!rok) continue;
        // the output angle's cell and constants
        const int ac = acell[R];
        const int kk = ac & 255, r = ac >> 8;
        const int nt = c_nth[kk], bd = c_bd[kk], quad = c_quad[kk], lf = bconv ? c_lf[kk] : 0;
        const int lk0 = c_l0[kk], lk1 = c_l0[kk + 1];
        const A dth = c_dth[kk], ct = cth[R], sn = sth[R], w_r = wr[R];
        const A vol = A(dx * dy * 0.125) * dth * w_r;
        const A cx = A(dy * 0.25) * dth * w_r * ct, cy = A(dx * 0.25) * dth * w_r * sn;
        const A fx = A(dy * 0.25) * dth, fy = A(dx * 0.25) * dth;   // face factors, x / y faces
        T *yc = y + cdof + nsp * c_ao[kk] + r;                      // y[cell + s nt + r]
        const A *xr = xs + R;                                      // column R of the copy
#pragma unroll
        for (int t = 0; t < HP_TG; ++t) {
            if (tg + t >= MT) break;
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                // f64 D row = l/16 + 4 v; f32 D row = 4 (l/16) + v
                const int s = (tg + t) * 16 + ((PREC == RTX_PREC_F64) ? (lg + 4 * v) : (4 * lg + v));
                if (s >= nsp) continue;
                const int pq = n_pq[s];
                const int p = pq & 255, q = pq >> 8;
                const A xv = xr[s * XS];
                A val = A(0);
                if (ops & RTX_OP_MASS) val += vol * w2k[s] * xv;
                if (ops & RTX_OP_ICONV) {
                    A sx = A(0), sy = A(0);
                    const A *lxp = xr + q * XS, *lyp = xr + p * ny * XS;
                    const A *wdx = WDx + p * HP_NM, *wdy = WDy + q * HP_NM;
                    for (int i = 0; i < nx; ++i) sx += wdx[i] * lxp[i * ny * XS];
                    for (int j = 0; j < ny; ++j) sy += wdy[j] * lyp[j * XS];
                    val -= cx * wys[q] * sx + cy * wxs[p] * sy;
                }
                // faces this node lies on (bit F: face F)
                const int onf = ((p == nx - 1) ? 1 : 0) | ((q == ny - 1) ? 2 : 0) |
                                ((p == 0) ? 4 : 0) | ((q == 0) ? 8 : 0);
                if (bconv && quad >= 0 && onf) {
#pragma unroll
                    for (int hh = 0; hh < 2; ++hh) {
                        const int F = (quad + hh) & 3;
                        if (onf & (1 << F)) {
                            const A tf = F == 0 ? ct : (F == 1 ? sn : (F == 2 ? -ct : -sn));
                            val += ((F & 1) ? fy * wxs[p] : fx * wys[q]) * w_r * tf * xv;
                        }
                    }
                }
                if (onf & lf) {
                    // upwind couplings of this cell on the node's faces, in link order
                    for (int li = lk0; li < lk1; ++li) {
                        const int w1 = l_w[li];
                        const int F = w1 & 3, n1 = w1 >> 2;
                        if (!(onf & (1 << F))) continue;
                        const int m0 = (F & 1) ? p : q;
                        const int n0 = (F & 1) ? nx : ny;
                        const double *Es = m.Es_pool + l_es[li] + m0;
                        const A *in_ = inn + l_inn[li] + r;
                        A f = A(0);
                        for (int m1 = 0; m1 < n1; ++m1) f += A(Es[m1 * n0]) * in_[m1 * nt];
                        val += ((F & 1) ? fy : fx) * f;
                    }
                }
                if (ops & RTX_OP_SCAT) val -= sgw[s] * A(acc[t][v]);
                const bool masked = bd && (onf & bd);
                T *yp = yc + s * nt;
                if (accumulate) {
                    if (!masked) *yp = T(A(*yp) + val);
                } else {
                    *yp = T(masked ? A(0) : val);
                }
            }
        }
    }
}

// ---- host side

// Bytes of dynamic LDS a column needs (FP64 or FP32 storage, mixed adds the
// BF16 copy).
static size_t hp_bytes(int nsp, int na, int ninn, int prec) {
    if (prec == RTX_PREC_F64) return size_t(hp_layout<double, RTX_PREC_F64>(nsp, na, ninn).total);
    if (prec == RTX_PREC_F32) return size_t(hp_layout<float, RTX_PREC_F32>(nsp, na, ninn).total);
    return size_t(hp_layout<float, RTX_PREC_MIXED>(nsp, na, ninn).total);
}

// Static LDS of k_hpcol (the tables above), for the capacity check.
constexpr size_t HP_STATIC_BYTES =
    size_t(6) * (HP_MAX_CELLS + 1) * 4 + 8 * HP_MAX_CELLS + (3 * HP_MAX_LINKS + 1) * 4 +
    4 * HP_NM * HP_NM + 8 * (4 * HP_NM * HP_NM + 2 * HP_NM) + 64;

int hpcol_build(rtx_plan *p, const rtx_mesh_desc *d) {
    const int nc = d->ncol;
    p->hp_ok = false;
    std::vector<int32_t> link_inn(std::max(d->nlink, 1), 0), col_inn(nc, 0);
    std::vector<char> fits(nc, 0);
    std::vector<size_t> need[3];
    for (int k = 0; k < 3; ++k) need[k].assign(nc, 0);
    int max_lds = 0;
    RTX_HIP(hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, p->device));
    const size_t cap = size_t(max_lds) > HP_STATIC_BYTES ? size_t(max_lds) - HP_STATIC_BYTES : 0;
    int nfit = 0;
    for (int c = 0; c < nc; ++c) {
        const int k0 = d->col_cell_start[c], k1 = d->col_cell_start[c + 1];
        const int nsp = d->col_nx[c] * d->col_ny[c], na = d->col_nang[c];
        int o = 0;
        for (int l = d->link_start[k0]; l < d->link_start[k1]; ++l) {
            const int s = d->link_src[l], F = d->link_face[l], c1 = d->cell_col[s];
            const int n1 = (F & 1) ? d->col_nx[c1] : d->col_ny[c1];
            link_inn[l] = o;
            o += n1 * d->cell_nth[d->link_target[l]];
        }
        col_inn[c] = o;
        const int nlc = d->link_start[k1] - d->link_start[k0];
        bool ok = (k1 - k0) <= HP_MAX_CELLS && nlc <= HP_MAX_LINKS && nsp <= HP_MT * 16 &&
                  na <= 4096;
        for (int pr = 0; pr < 3; ++pr) {
            need[pr][c] = hp_bytes(nsp, na, o, pr);
            ok = ok && need[pr][c] <= cap;
        }
        fits[c] = ok ? 1 : 0;
        nfit += ok ? 1 : 0;
    }
    // LDS classes: columns whose workgroup (dynamic + static LDS) fits half
    // of a CU's 160 KB run two workgroups per CU, the rest one (FP64 sizes;
    // the FP32 and mixed launches of a class take the class's own maximum)
    const size_t half = 80 * 1024 > HP_STATIC_BYTES ? 80 * 1024 - HP_STATIC_BYTES : 0;
    const size_t lim[HP_NCLS] = {half, cap};
    std::vector<int32_t> lists[HP_NCLS], over;
    for (int c = 0; c < nc; ++c) {
        if (!fits[c]) {
            over.push_back(c);
            continue;
        }
        int cls = HP_NCLS - 1;
        for (int k = 0; k < HP_NCLS; ++k)
            if (need[0][c] <= lim[k]) { cls = k; break; }
        lists[cls].push_back(c);
        for (int pr = 0; pr < 3; ++pr)
            p->hp_cls[cls].bytes[pr] = std::max(p->hp_cls[cls].bytes[pr], need[pr][c]);
    }
    // the columns the kernel cannot take run on the generic kernels
    std::vector<int32_t> gcells;
    for (int c : over)
        for (int k = d->col_cell_start[c]; k < d->col_cell_start[c + 1]; ++k) gcells.push_back(k);
    const std::vector<int32_t> gitems = scatter_items(p, over);
    size_t total = link_inn.size() + col_inn.size() + gcells.size() + gitems.size();
    for (int k = 0; k < HP_NCLS; ++k) total += lists[k].size();
    RTX_HIP(hipMalloc(&p->hp_alloc, total * sizeof(int32_t)));
    int32_t *base = static_cast<int32_t *>(p->hp_alloc);
    size_t o = 0;
    auto put = [&](const std::vector<int32_t> &v) -> const int32_t * {
        if (v.empty()) return nullptr;
        const int32_t *ptr = base + o;
        (void)hipMemcpy(base + o, v.data(), v.size() * sizeof(int32_t), hipMemcpyHostToDevice);
        o += v.size();
        return ptr;
    };
    p->hp.link_inn = put(link_inn);
    p->hp.col_inn = put(col_inn);
    for (int k = 0; k < HP_NCLS; ++k) {
        p->hp_cls[k].cols = put(lists[k]);
        p->hp_cls[k].n = int(lists[k].size());
    }
    p->hp_gen_cells = put(gcells);
    p->hp_n_gen_cells = int(gcells.size());
    p->hp_gen_items = put(gitems);
    p->hp_n_gen_items = int(gitems.size() / 2);
    RTX_HIP(hipGetLastError());
    p->hp_over_cols = over;
    p->hp_ok = nfit > 0;
    // the default apply of a mesh that is not one uniform order bin
    // (rtx_autotune may move it back to the generic kernels)
    p->hp_default = p->hp_ok && !p->uni_ok;
    return RTX_OK;
}

template <typename T, typename A, int PREC>
static int launch_hp_class(const rtx_plan *p, const HpClass &hc, uint32_t ops, const void *dx,
                           void *dy, hipStream_t st, int accumulate) {
    const size_t lds = hc.bytes[PREC == RTX_PREC_F64 ? 0 : (PREC == RTX_PREC_F32 ? 1 : 2)];
    const int per = (hc.n + HP_XCD - 1) / HP_XCD;
    RTX_HIP(hipFuncSetAttribute((const void *)k_hpcol<T, A, PREC>,
                                hipFuncAttributeMaxDynamicSharedMemorySize, int(lds)));
    k_hpcol<T, A, PREC><<<per * HP_XCD, HP_THREADS, lds, st>>>(p->m, p->hp, ops, accumulate,
                                                               hc.cols, hc.n, (const T *)dx,
                                                               (T *)dy);
    RTX_HIP(hipGetLastError());
    return RTX_OK;
}

// The column kernel on every column it takes (largest LDS class first).
int launch_hpcol(rtx_plan *p, uint32_t ops, int prec, const void *dx, void *dy, hipStream_t st,
                 int accumulate) {
    for (int k = HP_NCLS - 1; k >= 0; --k) {
        const HpClass &hc = p->hp_cls[k];
        if (hc.n == 0) continue;
        int rc;
        if (prec == RTX_PREC_F64)
            rc = launch_hp_class<double, double, RTX_PREC_F64>(p, hc, ops, dx, dy, st, accumulate);
        else if (prec == RTX_PREC_F32)
            rc = launch_hp_class<float, float, RTX_PREC_F32>(p, hc, ops, dx, dy, st, accumulate);
        else
            rc = launch_hp_class<float, float, RTX_PREC_MIXED>(p, hc, ops, dx, dy, st, accumulate);
        if (rc) return rc;
    }
    return RTX_OK;
}

}  // namespace rtx
