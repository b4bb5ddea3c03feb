// Matrix-free DG transport operator for uneven-order hp meshes: one fused
// pass per spatial column.
//
// A column's DOFs are contiguous in the reference's global order (columns by
// key, cells by key, then beta = nth ny i + nth j + a inside a cell).  One
// workgroup of 512 threads owns one column of N x N spatial nodes:
//
//   phase 0  the column's cell, angle, node and link tables -> LDS
//   phase 1  x of the column -> LDS, angle-major (row = column angle A, one
//            entry per spatial node; inflow-boundary DOFs zeroed; mixed
//            precision also keeps a node-major BF16 copy for the matrix-core
//            operands), and, for every upwind link, the neighbour's trace
//            contracted with E^theta over the neighbour's angles:
//            inn[m1, r] = sum_a E^th[a, r] tr[m1, a]
//            (boundary_convection_matrix.py:348-524, the angular half), then
//            with E^s over the neighbour's face nodes, summed over the links
//            of each cell face: one value per (cell, face, face node, angle)
//   phase 2  each wavefront takes (tile of 16 output angles R, group of
//            node tiles) items: it accumulates the scattering contraction
//            sum_A x[s, A] G^T[A, R] on the matrix cores
//            (scattering_matrix.py:142-237), then, in the accumulator layout,
//            adds for each of its (s, R) the mass (mass_matrix.py:140-173),
//            interior convection (interior_convection_matrix.py:161-243), the
//            outflow faces (boundary_convection_matrix.py:284-346) and the
//            inflow-face sums of the node's faces, and stores y.
//
// Phases 1 and 2 are compiled per node count N (3..10), so the line reads of
// the derivative terms and the E^s rows have compile-time offsets.  Every
// output has one owner and a fixed order of additions, so results are
// bit-reproducible; x is read from HBM once (plus the neighbours' face
// traces), G once and y written once.  Columns are grouped by the LDS they
// need into two launches, so columns of up to about 70 KB run two
// workgroups per CU.  Arithmetic: FP64 (v_mfma_f64_16x16x4f64), FP32
// (v_mfma_f32_16x16x4f32) or mixed (FP32 transport terms, BF16 scattering
// operands on v_mfma_f32_16x16x16bf16_1k with FP32 accumulation).
#include <algorithm>
#include <type_traits>
#include <utility>

#include "rtx_internal.h"

namespace rtx {

constexpr int HP_THREADS = 512;
constexpr int HP_WAVES = HP_THREADS / 64;
constexpr int HP_NM = RTX_NMAX;
constexpr int HP_MT = (RTX_NMAX * RTX_NMAX + 15) / 16;   // node tiles of 16
constexpr int HP_XCD = 8;
constexpr int HP_TG = 2;          // node tiles per scattering pass (accumulators in registers)
constexpr int HP_NMIN = 3;        // node counts compiled for the column kernel
constexpr int HP_FV = 8;          // inflow-face sums per thread (a column holds <= HP_FV * HP_THREADS)
constexpr int HP_LINK_LOG = 8;    // 2^8 >= HP_MAX_LINKS
constexpr int HP_CELL_LOG = 6;    // 2^6 >= HP_MAX_CELLS
static_assert((1 << HP_LINK_LOG) >= HP_MAX_LINKS && (1 << HP_CELL_LOG) >= HP_MAX_CELLS, "search range");

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

// Largest i in [0, hi) with starts[i] <= t (starts ascending, starts[0] <= t,
// hi <= 2^LOG).  A fixed number of steps, so the searches of an unrolled
// loop's iterations are independent and their LDS reads overlap.
template <int LOG>
__device__ inline int hp_find(const int *starts, int hi, int t) {
    int lo = 0;
#pragma unroll
    for (int s = 1 << (LOG - 1); s > 0; s >>= 1) {
        const int mid = lo + s;
        if (mid < hi && starts[mid] <= t) lo = mid;
    }
    return lo;
}

// Row stride (elements) of a node-major copy: at least na rounded up to the
// K step, and congruent to c modulo M, so that the 16 rows x K lanes of one
// matrix-core operand read fall on distinct banks (BF16 quads, 16 x 2 per
// half-wave: stride / 8 odd, c = 8, M = 16).
__host__ __device__ inline int hp_stride(int na, int kstep, int c, int M) {
    const int n = (na + kstep - 1) / kstep * kstep;
    return n + ((c - n) % M + M) % M;
}

struct HpLayout {           // dynamic LDS of one column, in bytes from the base
    int xs, xb, inn, es, cth, sth, wr, acell, total;
    int S, NA, XSB;         // angle-major row (node) stride, padded angle rows, BF16 stride
};

// Angle-major copy: NA rows (the column's angles, padded to the K step with
// zero rows) of S = nsp rounded up to odd entries; the same region first
// holds the ntr neighbour-trace values of the column's links.  The inn
// region holds nin = max(ninn, nffc) entries: the E^theta-contracted traces,
// then the inflow-face sums written over them; the es region the column's
// nes distinct E^s values.  A: LDS element type.
template <typename A, int PREC>
__host__ __device__ inline HpLayout hp_layout(int nsp, int na, int nin, int ntr, int nes) {
    HpLayout L;
    const int ea = int(sizeof(A));
    const int kstep = (PREC == RTX_PREC_MIXED) ? 16 : 4;
    L.S = nsp | 1;
    L.NA = (na + kstep - 1) / kstep * kstep;
    L.XSB = (PREC == RTX_PREC_MIXED) ? hp_stride(na, 16, 8, 16) : 0;
    int o = 0;
    L.xs = o;
    o += (L.NA * L.S > ntr ? L.NA * L.S : ntr) * ea;
    o = (o + 15) & ~15;
    L.xb = o;
    o += nsp * L.XSB * 2;
    o = (o + 15) & ~15;
    L.inn = o;
    o += nin * ea;
    L.es = o;
    o += nes * ea;
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

// LDS tables of one column (phase 0).
template <typename A>
struct HpShared {
    int c_nth[HP_MAX_CELLS], c_bd[HP_MAX_CELLS], c_lf[HP_MAX_CELLS];
    int c_quad[HP_MAX_CELLS], c_ao[HP_MAX_CELLS + 1], c_l0[HP_MAX_CELLS + 1];
    int c_fo[HP_MAX_CELLS + 1];     // offsets of the cells' inflow-face sums
    A c_dth[HP_MAX_CELLS];
    int l_w[HP_MAX_LINKS], l_inn[HP_MAX_LINKS + 1], l_es[HP_MAX_LINKS];
    // per link: packed source shape word (link_meta[4 l + 1]), source cell's
    // first DOF, E^theta offset, column-local trace offset
    int l_m1[HP_MAX_LINKS], l_src[HP_MAX_LINKS], l_et[HP_MAX_LINKS], l_tr[HP_MAX_LINKS + 1];
    A w2k[HP_NM * HP_NM], sgw[HP_NM * HP_NM];
    A wxs[HP_NM], wys[HP_NM], WDx[HP_NM * HP_NM], WDy[HP_NM * HP_NM];
};

// Phases 1 and 2 of one column of N x N nodes.
template <typename T, typename A, int PREC, int N>
__device__ __forceinline__ void hp_column(const DevMesh &m, uint32_t ops, int accumulate,
                                          HpShared<A> &sh, unsigned char *lds, int c, int k0,
                                          int ncells, int na, int cdof, int L0, int nl,
                                          int ninn, int nffc, int ntr, int es0, int nes,
                                          const double *__restrict__ h_es, double dx,
                                          double dy, const T *__restrict__ x,
                                          T *__restrict__ y) {
    constexpr int NSP = N * N;
    constexpr int MT = (NSP + 15) / 16;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int l16 = lane & 15, lg = lane >> 4;
    const bool restrict_ = (ops & RTX_OP_RESTRICT) != 0;
    const bool bconv = (ops & RTX_OP_BCONV) != 0;
    const int ncd = NSP * na;
    const HpLayout Lo = hp_layout<A, PREC>(NSP, na, ninn > nffc ? ninn : nffc, ntr, nes);
    const int S = Lo.S, XSB = Lo.XSB;
    A *xs = reinterpret_cast<A *>(lds + Lo.xs);
    uint16_t *xb = reinterpret_cast<uint16_t *>(lds + Lo.xb);
    A *inn = reinterpret_cast<A *>(lds + Lo.inn);
    A *es = reinterpret_cast<A *>(lds + Lo.es);
    const A *cth = reinterpret_cast<const A *>(lds + Lo.cth);
    const A *sth = reinterpret_cast<const A *>(lds + Lo.sth);
    const A *wr = reinterpret_cast<const A *>(lds + Lo.wr);
    const int *acell = reinterpret_cast<const int *>(lds + Lo.acell);

    // ---- phase 1.  The first batch of x is loaded into registers; while it
    // is in flight the neighbour traces of the column's links are gathered
    // into the x region (batched loads), then contracted with E^theta into
    // inn; then x is written over them, angle-major and masked.  DOF t of
    // the column lies in the cell of angle t / NSP (cell kk's DOFs start at
    // NSP * c_ao[kk]).
    constexpr int XU = 4;                 // x and trace values per thread and pass
    A vv[XU];
#pragma unroll
    for (int u = 0; u < XU; ++u) {
        const int t = tid + u * HP_THREADS;
        vv[u] = (t < ncd) ? A(x[cdof + t]) : A(0);
    }
    if (bconv) {
        // the column's distinct E^s blocks -> LDS (read after two barriers)
        for (int e = tid; e < nes; e += HP_THREADS) es[e] = A(h_es[es0 + e]);
        A *tr = xs;
        for (int e0 = tid; e0 < ntr; e0 += XU * HP_THREADS) {
            int addr[XU];
#pragma unroll
            for (int u = 0; u < XU; ++u) {
                const int e = e0 + u * HP_THREADS;
                addr[u] = -1;
                if (e < ntr) {
                    const int li = hp_find<HP_LINK_LOG>(sh.l_tr, nl, e);
                    const int w1 = sh.l_m1[li];
                    const int F = w1 & 3, nx1 = (w1 >> 2) & 15, ny1 = (w1 >> 6) & 15;
                    const int nt1 = (w1 >> 10) & 15, bd1 = restrict_ ? ((w1 >> 14) & 15) : 0;
                    const int le = e - sh.l_tr[li];
                    const int m1 = hp_div(le, 1.0f / float(nt1)), a = le - m1 * nt1;
                    int i1, j1;
                    if (F == 0) { i1 = 0; j1 = m1; }
                    else if (F == 2) { i1 = nx1 - 1; j1 = m1; }
                    else if (F == 1) { i1 = m1; j1 = 0; }
                    else { i1 = m1; j1 = ny1 - 1; }
                    if (!(bd1 && hp_on_bdry(bd1, i1, j1, nx1, ny1)))
                        addr[u] = sh.l_src[li] + (i1 * ny1 + j1) * nt1 + a;
                }
            }
            A tv[XU];
#pragma unroll
            for (int u = 0; u < XU; ++u) tv[u] = (addr[u] >= 0) ? A(x[addr[u]]) : A(0);
#pragma unroll
            for (int u = 0; u < XU; ++u) {
                const int e = e0 + u * HP_THREADS;
                if (e < ntr) tr[e] = tv[u];
            }
        }
        __syncthreads();
        // two entries per thread and step, so that their searches and E^theta
        // loads are in flight together
        for (int e0 = tid; e0 < ninn; e0 += 2 * HP_THREADS) {
#pragma unroll
            for (int u = 0; u < 2; ++u) {
                const int e = e0 + u * HP_THREADS;
                if (e < ninn) {
                    const int li = hp_find<HP_LINK_LOG>(sh.l_inn, nl, e);
                    const int w1 = sh.l_m1[li];
                    const int nt1 = (w1 >> 10) & 15, kk = w1 >> 18;
                    const int nt = sh.c_nth[kk];
                    const int le = e - sh.l_inn[li], m1 = hp_div(le, 1.0f / float(nt));
                    const int r = le - m1 * nt;
                    const double *Et = m.Eth_pool + sh.l_et[li] + r;
                    const A *trl = tr + sh.l_tr[li] + m1 * nt1;
                    // the E^theta column's (at most RTX_NMAX) loads in flight
                    // together: unconditional (rows past nt1 reload row
                    // nt1 - 1), so that no conversion or wait sits between
                    // them
                    double ev[HP_NM];
#pragma unroll
                    for (int a = 0; a < HP_NM; ++a) ev[a] = Et[(a < nt1 ? a : nt1 - 1) * nt];
                    // the trace reads likewise unconditional; terms past nt1
                    // are zero times a finite trace value
                    A acc = A(0);
#pragma unroll
                    for (int a = 0; a < HP_NM; ++a) {
                        const A tv = trl[a < nt1 ? a : nt1 - 1];
                        acc += (a < nt1 ? A(ev[a]) : A(0)) * tv;
                    }
                    inn[e] = acc;
                }
            }
        }
        __syncthreads();
        // inflow-face sums.  Entry (kk, h, m0, r): cell kk, the h-th of its
        // faces that carry links (ascending face number F), own face node
        // m0, angle r of the cell:
        //   sum over kk's links on F, in link order, of
        //   (face factor) * sum_m1 E^s[m0, m1] inn[link, m1, r]
        // (the spatial half of boundary_convection_matrix.py:348-524).  They
        // are held in registers over a barrier, then written over inn, which
        // no one reads after them.
        A fv[HP_FV];
#pragma unroll
        for (int u = 0; u < HP_FV; ++u) {
            const int e = tid + u * HP_THREADS;
            fv[u] = A(0);
            if (e < nffc) {
                const int kk = hp_find<HP_CELL_LOG>(sh.c_fo, ncells, e);
                const int nt = sh.c_nth[kk], lf = sh.c_lf[kk];
                const int le = e - sh.c_fo[kk];
                const int row = hp_div(le, 1.0f / float(nt)), r = le - row * nt;
                const int hs = hp_div(row, 1.0f / float(N)), m0 = row - hs * N;
                int F = 0;
                for (int cnt = hs; F < 4; ++F)
                    if ((lf >> F) & 1) {
                        if (cnt == 0) break;
                        --cnt;
                    }
                const A fxy = A(((F & 1) ? dx : dy) * 0.25) * sh.c_dth[kk];
                A sum = A(0);
                for (int li = sh.c_l0[kk]; li < sh.c_l0[kk + 1]; ++li) {
                    const int w1 = sh.l_w[li];
                    if ((w1 & 3) != F) continue;
                    const int n1 = w1 >> 2;
                    const A *Es = es + sh.l_es[li] + m0;
                    const A *in_ = inn + sh.l_inn[li] + r;
                    // the E^s row's and the traces' (at most RTX_NMAX each)
                    // reads in flight together: unconditional, rows past n1
                    // reread row n1 - 1 and add zero times a finite value
                    A ev[HP_NM], iv[HP_NM];
#pragma unroll
                    for (int m1 = 0; m1 < HP_NM; ++m1) {
                        const int mm = m1 < n1 ? m1 : n1 - 1;
                        ev[m1] = Es[mm * N];
                        iv[m1] = in_[mm * nt];
                    }
                    A f = A(0);
#pragma unroll
                    for (int m1 = 0; m1 < HP_NM; ++m1) f += (m1 < n1 ? ev[m1] : A(0)) * iv[m1];
                    sum += fxy * f;
                }
                fv[u] = sum;
            }
        }
        __syncthreads();
#pragma unroll
        for (int u = 0; u < HP_FV; ++u) {
            const int e = tid + u * HP_THREADS;
            if (e < nffc) inn[e] = fv[u];
        }
    }
    for (int t0 = tid; t0 < ncd; t0 += XU * HP_THREADS) {
        if (t0 != tid) {
#pragma unroll
            for (int u = 0; u < XU; ++u) {
                const int t = t0 + u * HP_THREADS;
                vv[u] = (t < ncd) ? A(x[cdof + t]) : A(0);
            }
        }
#pragma unroll
        for (int u = 0; u < XU; ++u) {
            const int t = t0 + u * HP_THREADS;
            if (t >= ncd) break;
            A v = vv[u];
            const int kk = acell[t / NSP] & 255;
            const int nt = sh.c_nth[kk], ao = sh.c_ao[kk];
            const int lt = t - NSP * ao;
            const int s = hp_div(lt, 1.0f / float(nt)), a = lt - s * nt;
            if (sh.c_bd[kk]) {
                const int p = s / N, q = s - p * N;
                if (hp_on_bdry(sh.c_bd[kk], p, q, N, N)) v = A(0);
            }
            xs[(ao + a) * S + s] = v;
            if constexpr (PREC == RTX_PREC_MIXED) xb[s * XSB + ao + a] = hp_bf16(float(v));
        }
    }
    // zero the padding: angle rows [na, NA) and (mixed) columns [na, XSB)
    for (int t = tid; t < (Lo.NA - na) * S; t += HP_THREADS) xs[na * S + t] = A(0);
    if constexpr (PREC == RTX_PREC_MIXED) {
        for (int t = tid; t < NSP * (XSB - na); t += HP_THREADS) {
            const int s = t / (XSB - na), a = na + t % (XSB - na);
            xb[s * XSB + a] = 0;
        }
    }
    __syncthreads();

    // ---- phase 2
    constexpr int NTG = (MT + HP_TG - 1) / HP_TG;
    const int NRT = (na + 15) / 16;
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
            // operand rows: node (tg + t) * 16 + l16 of this lane (rows past
            // the last node read the last node; their outputs are not stored)
            int arow[HP_TG];
#pragma unroll
            for (int t = 0; t < HP_TG; ++t) {
                const int s = (tg + t) * 16 + l16;
                arow[t] = (s < NSP ? s : NSP - 1);
            }
            if constexpr (PREC != RTX_PREC_MIXED) {
                typedef typename std::conditional<PREC == RTX_PREC_F64, double, float>::type E;
                const E *G = (PREC == RTX_PREC_F64) ? (const E *)(m.G + m.col_G_off[c])
                                                    : (const E *)(m.G_f + m.col_G_off[c]);
                const E *gcol = G + Rc;
                // G rows past na are clamped to the last row: the angle rows of
                // the copy are zero there, so they add nothing
                auto gload = [&](int Ak) -> E { return gcol[size_t(Ak < na ? Ak : na - 1) * na]; };
                const A *xcol = xs + lg * S;
                // four K steps of G per load group; the next two groups are
                // in flight while the current one multiplies (three: 20 bytes
                // per lane of register spills in FP64)
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
                            const A *xk = xcol + (kb + 4 * u) * S;
#pragma unroll
                            for (int t = 0; t < HP_TG; ++t) {
                                if ((tg + t) < MT) {
                                    const E a = E(xk[arow[t]]);
                                    if constexpr (PREC == RTX_PREC_F64)
                                        acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, bc[u], acc[t], 0, 0, 0);
                                    else
                                        acc[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, bc[u], acc[t], 0, 0, 0);
                                }
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
                        if ((tg + t) < MT) {
                            const hp_s4 av = *reinterpret_cast<const hp_s4 *>(xrow[t] + kb);
                            acc[t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av, bv, acc[t], 0, 0, 0);
                        }
                    }
                }
            }
        }
        if (!rok) continue;
        // the output angle's cell and constants
        const int ac = acell[R];
        const int kk = ac & 255, r = ac >> 8;
        const int nt = sh.c_nth[kk], bd = sh.c_bd[kk], quad = sh.c_quad[kk];
        const int lf = bconv ? sh.c_lf[kk] : 0;
        const A dth = sh.c_dth[kk], ct = cth[R], sn = sth[R], w_r = wr[R];
        const A vol = A(dx * dy * 0.125) * dth * w_r;
        const A cx = A(dy * 0.25) * dth * w_r * ct, cy = A(dx * 0.25) * dth * w_r * sn;
        const A fx = A(dy * 0.25) * dth, fy = A(dx * 0.25) * dth;   // face factors, x / y faces
        T *yc = y + cdof + NSP * sh.c_ao[kk] + r;                  // y[cell + s nt + r]
        const A *xr = xs + R * S;                                  // angle row R of the copy
#pragma unroll
        for (int t = 0; t < HP_TG; ++t) {
            if (tg + t >= MT) break;
            // one output at a time (the line reads of an output are in
            // flight together; rolling this loop keeps them in registers)
#pragma unroll 1
            for (int v = 0; v < 4; ++v) {
                // f64 D row = l/16 + 4 v; f32 D row = 4 (l/16) + v
                const int s = (tg + t) * 16 + ((PREC == RTX_PREC_F64) ? (lg + 4 * v) : (4 * lg + v));
                if (s >= NSP) continue;
                const int p = s / N, q = s - p * N;
                const A xv = xr[s];
                A val = A(0);
                if (ops & RTX_OP_MASS) val += vol * sh.w2k[s] * xv;
                if (ops & RTX_OP_ICONV) {
                    const A *lx = xr + q, *ly = xr + p * N;
                    const A *wdx = sh.WDx + p * HP_NM, *wdy = sh.WDy + q * HP_NM;
                    A sx = A(0), sy = A(0);
#pragma unroll
                    for (int i = 0; i < N; ++i) {
                        sx += wdx[i] * lx[i * N];
                        sy += wdy[i] * ly[i];
                    }
                    val -= cx * sh.wys[q] * sx + cy * sh.wxs[p] * sy;
                }
                // faces this node lies on (bit F: face F)
                const int onf = ((p == N - 1) ? 1 : 0) | ((q == N - 1) ? 2 : 0) |
                                ((p == 0) ? 4 : 0) | ((q == 0) ? 8 : 0);
                if (bconv && quad >= 0 && onf) {
#pragma unroll
                    for (int hh = 0; hh < 2; ++hh) {
                        const int F = (quad + hh) & 3;
                        if (onf & (1 << F)) {
                            const A tf = F == 0 ? ct : (F == 1 ? sn : (F == 2 ? -ct : -sn));
                            val += ((F & 1) ? fy * sh.wxs[p] : fx * sh.wys[q]) * w_r * tf * xv;
                        }
                    }
                }
                if (onf & lf) {
                    // the cell's inflow-face sums on the node's faces, by face
                    const int fo = sh.c_fo[kk];
                    int slot = 0;
                    // kept rolled: unrolled, the compiler's schedule of this
                    // loop spills phase 2's registers
#pragma unroll 1
                    for (int F = 0; F < 4; ++F) {
                        if ((lf >> F) & 1) {
                            if ((onf >> F) & 1) {
                                const int m0 = (F & 1) ? p : q;
                                val += inn[fo + (slot * N + m0) * nt + r];
                            }
                            ++slot;
                        }
                    }
                }
                if (ops & RTX_OP_SCAT) val -= sh.sgw[s] * A(acc[t][v]);
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

template <typename T, typename A, int PREC>
__global__ __launch_bounds__(HP_THREADS) __attribute__((amdgpu_waves_per_eu(4)))
void k_hpcol(DevMesh m, HpTables h, uint32_t ops, int accumulate,
             const int32_t *__restrict__ cols, int ncols, const T *__restrict__ x,
             T *__restrict__ y) {
    extern __shared__ __align__(16) unsigned char hp_lds[];
    __shared__ HpShared<A> sh;

    // XCD-aware block -> column: block b runs on XCD b % 8, and XCD j takes
    // the j-th contiguous run of the (key-ordered) list, so the neighbour
    // columns whose traces a column reads are mostly cached on the same XCD
    const int per = (ncols + HP_XCD - 1) / HP_XCD;
    const int bi = blockIdx.x % HP_XCD, bj = blockIdx.x / HP_XCD;
    if (bj >= per) return;
    const int ci = bi * per + bj;
    if (ci >= ncols) return;
    const int c = cols[ci];
    const int tid = threadIdx.x;
    // the column's scalars from its header (square columns: ny = nx)
    const int32_t *hd = h.col_hdr + size_t(c) * HP_HDR;
    const int k0 = hd[0], ncells = hd[1], nx = hd[2], ny = nx, nsp = nx * ny;
    const int na = hd[3], cdof = hd[4], L0 = hd[5], nl = hd[6], ga0 = hd[7];
    const int ninn = hd[8], nffc = hd[9], ntr = hd[10], es0 = hd[11], nes = hd[12];
    const int sp0 = hd[13];
    const bool restrict_ = (ops & RTX_OP_RESTRICT) != 0;
    const double x0 = m.col_geom[4 * c + 0], y0 = m.col_geom[4 * c + 1];
    const double dx = m.col_geom[4 * c + 2] - x0, dy = m.col_geom[4 * c + 3] - y0;
    const HpLayout Lo = hp_layout<A, PREC>(nsp, na, ninn > nffc ? ninn : nffc, ntr, nes);
    A *cth = reinterpret_cast<A *>(hp_lds + Lo.cth);
    A *sth = reinterpret_cast<A *>(hp_lds + Lo.sth);
    A *wr = reinterpret_cast<A *>(hp_lds + Lo.wr);
    int *acell = reinterpret_cast<int *>(hp_lds + Lo.acell);

    // ---- phase 0: tables
    if (tid < ncells) {
        const int k = k0 + tid;
        const int nt = m.cell_nth[k], ao = m.cell_ang_off[k];
        sh.c_nth[tid] = nt;
        sh.c_bd[tid] = restrict_ ? m.cell_bdry[k] : 0;
        sh.c_quad[tid] = m.cell_quad[k];
        sh.c_ao[tid] = ao;
        sh.c_l0[tid] = m.link_start[k] - L0;
        sh.c_fo[tid] = h.cell_fo[k];
        sh.c_dth[tid] = A(m.cell_geom[2 * k + 1] - m.cell_geom[2 * k]);
        sh.c_lf[tid] = h.cell_lf[k];
        for (int a = 0; a < nt; ++a) acell[ao + a] = tid | (a << 8);
    }
    if (tid == 0) {
        sh.c_ao[ncells] = na;
        sh.c_l0[ncells] = nl;
        sh.l_inn[nl] = ninn;
        sh.c_fo[ncells] = nffc;
        sh.l_tr[nl] = ntr;
    }
    for (int a = tid; a < na; a += HP_THREADS) {
        cth[a] = A(m.ang_trig[3 * (ga0 + a)]);
        sth[a] = A(m.ang_trig[3 * (ga0 + a) + 1]);
        wr[a] = A(m.ang_trig[3 * (ga0 + a) + 2]);
    }
    for (int li = tid; li < nl; li += HP_THREADS) {
        const int l = L0 + li;
        sh.l_w[li] = h.link_w[l];
        sh.l_inn[li] = h.link_inn[l];
        sh.l_es[li] = h.link_esl[l];
        sh.l_m1[li] = m.link_meta[4 * l + 1];
        sh.l_src[li] = m.link_meta[4 * l];
        sh.l_tr[li] = m.link_meta[4 * l + 2];
        sh.l_et[li] = m.link_Eth_off[l];
    }
    if (tid < nsp) {
        const int i = tid / ny, j = tid - (tid / ny) * ny;
        const double w2 = m.lgl_w[nx * HP_NM + i] * m.lgl_w[ny * HP_NM + j];
        sh.w2k[tid] = A(w2 * m.kappa[sp0 + tid]);
        sh.sgw[tid] = A(dx * dy * m.sigma[sp0 + tid] * w2);
    }
    if (tid < HP_NM * HP_NM) {
        const int pp = tid / HP_NM, ii = tid % HP_NM;
        sh.WDx[tid] = (pp < nx && ii < nx)
                          ? A(m.lgl_w[nx * HP_NM + ii] * m.lgl_D[(size_t(nx) * HP_NM + pp) * HP_NM + ii])
                          : A(0);
        sh.WDy[tid] = (pp < ny && ii < ny)
                          ? A(m.lgl_w[ny * HP_NM + ii] * m.lgl_D[(size_t(ny) * HP_NM + pp) * HP_NM + ii])
                          : A(0);
    }
    if (tid < HP_NM) {
        sh.wxs[tid] = A(tid < nx ? m.lgl_w[nx * HP_NM + tid] : 0.0);
        sh.wys[tid] = A(tid < ny ? m.lgl_w[ny * HP_NM + tid] : 0.0);
    }
    __syncthreads();

    // phases 1 and 2, compiled per node count (square columns; the host
    // admits only those)
    switch (nx) {
#define HP_CASE(NN)                                                                           \
    case NN:                                                                                  \
        hp_column<T, A, PREC, NN>(m, ops, accumulate, sh, hp_lds, c, k0, ncells, na, cdof, L0, \
                                  nl, ninn, nffc, ntr, es0, nes, h.es_pool, dx, dy, x, y);   \
        break;
        HP_CASE(3) HP_CASE(4) HP_CASE(5) HP_CASE(6) HP_CASE(7) HP_CASE(8) HP_CASE(9) HP_CASE(10)
#undef HP_CASE
        default: break;
    }
}

// ---- host side

// Bytes of dynamic LDS a column needs (FP64 or FP32 storage, mixed adds the
// BF16 copy).
static size_t hp_bytes(int nsp, int na, int nin, int ntr, int nes, int prec) {
    if (prec == RTX_PREC_F64)
        return size_t(hp_layout<double, RTX_PREC_F64>(nsp, na, nin, ntr, nes).total);
    if (prec == RTX_PREC_F32)
        return size_t(hp_layout<float, RTX_PREC_F32>(nsp, na, nin, ntr, nes).total);
    return size_t(hp_layout<float, RTX_PREC_MIXED>(nsp, na, nin, ntr, nes).total);
}

// Static LDS of k_hpcol (the tables above), for the capacity check.
constexpr size_t HP_STATIC_BYTES = sizeof(HpShared<double>) + 64;

int hpcol_build(rtx_plan *p, const rtx_mesh_desc *d) {
    const int nc = d->ncol;
    p->hp_ok = false;
    std::vector<int32_t> link_inn(std::max(d->nlink, 1), 0), col_inn(nc, 0);
    std::vector<int32_t> cell_fo(std::max(d->ncell, 1), 0), col_ffc(nc, 0);
    std::vector<int32_t> link_esl(std::max(d->nlink, 1), 0), col_es0(nc, 0), col_nes(nc, 0);
    std::vector<double> es_pool;
    std::vector<int32_t> col_hdr(size_t(nc) * HP_HDR, 0), cell_lf(std::max(d->ncell, 1), 0);
    std::vector<int32_t> link_w(std::max(d->nlink, 1), 0);
    std::vector<std::pair<int64_t, int>> seen;   // (pool offset, staged offset) of one column
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
        // inflow-face sums: per cell, one block of nx x nth per face that
        // carries links
        int fo = 0;
        for (int k = k0; k < k1; ++k) {
            int lf = 0;
            for (int l = d->link_start[k]; l < d->link_start[k + 1]; ++l) {
                const int F = d->link_face[l], c1 = d->cell_col[d->link_src[l]];
                lf |= 1 << F;
                link_w[l] = F | (((F & 1) ? d->col_nx[c1] : d->col_ny[c1]) << 2);
            }
            cell_lf[k] = lf;
            cell_fo[k] = fo;
            fo += __builtin_popcount(lf) * d->col_nx[c] * d->cell_nth[k];
        }
        col_ffc[c] = fo;
        // the column's distinct E^s blocks (n0 x n1 each, keyed by pool
        // offset), copied in link order
        seen.clear();
        col_es0[c] = int(es_pool.size());
        int nes = 0;
        for (int l = d->link_start[k0]; l < d->link_start[k1]; ++l) {
            const int64_t off = d->link_Es_off[l];
            int pos = -1;
            for (const auto &sp : seen)
                if (sp.first == off) { pos = sp.second; break; }
            if (pos < 0) {
                const int s = d->link_src[l], F = d->link_face[l], c1 = d->cell_col[s];
                const int n0 = (F & 1) ? d->col_nx[c] : d->col_ny[c];
                const int n1 = (F & 1) ? d->col_nx[c1] : d->col_ny[c1];
                pos = nes;
                seen.emplace_back(off, pos);
                es_pool.insert(es_pool.end(), d->Es_pool + off, d->Es_pool + off + int64_t(n0) * n1);
                nes += n0 * n1;
            }
            link_esl[l] = pos;
        }
        col_nes[c] = nes;
        const int nlc = d->link_start[k1] - d->link_start[k0];
        int32_t *hc = col_hdr.data() + size_t(c) * HP_HDR;
        hc[0] = k0;
        hc[1] = k1 - k0;
        hc[2] = d->col_nx[c];
        hc[3] = na;
        hc[4] = d->cell_dof_off[k0];
        hc[5] = d->link_start[k0];
        hc[6] = nlc;
        hc[7] = d->col_ang_start[c];
        hc[8] = o;
        hc[9] = fo;
        hc[10] = d->col_ntrace[c];
        hc[11] = col_es0[c];
        hc[12] = nes;
        hc[13] = d->col_sp_off[c];
        // square columns of HP_NMIN..RTX_NMAX nodes (the compiled node counts)
        bool ok = (k1 - k0) <= HP_MAX_CELLS && nlc <= HP_MAX_LINKS &&
                  d->col_nx[c] == d->col_ny[c] && d->col_nx[c] >= HP_NMIN && na <= 4096 &&
                  fo <= HP_FV * HP_THREADS;
        for (int pr = 0; pr < 3; ++pr) {
            need[pr][c] = hp_bytes(nsp, na, std::max(o, fo), d->col_ntrace[c], nes, pr);
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
    size_t total = link_inn.size() + col_inn.size() + cell_fo.size() + col_ffc.size() +
                   link_esl.size() + col_es0.size() + col_nes.size() + col_hdr.size() +
                   cell_lf.size() + link_w.size() + gcells.size() + gitems.size();
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
    p->hp.cell_fo = put(cell_fo);
    p->hp.col_ffc = put(col_ffc);
    p->hp.link_esl = put(link_esl);
    p->hp.col_es0 = put(col_es0);
    p->hp.col_nes = put(col_nes);
    p->hp.col_hdr = put(col_hdr);
    p->hp.cell_lf = put(cell_lf);
    p->hp.link_w = put(link_w);
    if (es_pool.empty()) es_pool.push_back(0.0);
    RTX_HIP(hipMalloc(&p->hp_es, es_pool.size() * sizeof(double)));
    RTX_HIP(hipMemcpy(p->hp_es, es_pool.data(), es_pool.size() * sizeof(double),
                      hipMemcpyHostToDevice));
    p->hp.es_pool = static_cast<const double *>(p->hp_es);
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
    // side stream (default priority) and fork / join events of the class launches
    if (p->hp_ok) {
        RTX_HIP(hipStreamCreateWithFlags(&p->hp_side, hipStreamNonBlocking));
        RTX_HIP(hipEventCreateWithFlags(&p->hp_fork, hipEventDisableTiming));
        RTX_HIP(hipEventCreateWithFlags(&p->hp_join, hipEventDisableTiming));
    }
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

// The column kernel on every column it takes.  The LDS classes write
// disjoint outputs, so the large-column class runs on the caller's stream
// and the others on the plan's side stream at the same time (the side
// stream waits for the work queued before the call, and the caller's stream
// waits for the side stream before returning): the few large-column
// workgroups then share the GPU with the small ones instead of leaving
// CUs idle through their tail.
int launch_hpcol(rtx_plan *p, uint32_t ops, int prec, const void *dx, void *dy, hipStream_t st,
                 int accumulate) {
    auto launch = [&](int k, hipStream_t s) -> int {
        const HpClass &hc = p->hp_cls[k];
        if (hc.n == 0) return RTX_OK;
        if (prec == RTX_PREC_F64)
            return launch_hp_class<double, double, RTX_PREC_F64>(p, hc, ops, dx, dy, s, accumulate);
        if (prec == RTX_PREC_F32)
            return launch_hp_class<float, float, RTX_PREC_F32>(p, hc, ops, dx, dy, s, accumulate);
        return launch_hp_class<float, float, RTX_PREC_MIXED>(p, hc, ops, dx, dy, s, accumulate);
    };
    int busy = 0;
    for (int k = 0; k < HP_NCLS; ++k) busy += p->hp_cls[k].n > 0 ? 1 : 0;
    if (busy < 2 || !p->hp_side) {
        for (int k = HP_NCLS - 1; k >= 0; --k) {
            const int rc = launch(k, st);
            if (rc) return rc;
        }
        return RTX_OK;
    }
    RTX_HIP(hipEventRecord(p->hp_fork, st));
    RTX_HIP(hipStreamWaitEvent(p->hp_side, p->hp_fork, 0));
    int rc = launch(HP_NCLS - 1, st);
    if (rc) return rc;
    for (int k = HP_NCLS - 2; k >= 0; --k) {
        rc = launch(k, p->hp_side);
        if (rc) return rc;
    }
    RTX_HIP(hipEventRecord(p->hp_join, p->hp_side));
    RTX_HIP(hipStreamWaitEvent(st, p->hp_join, 0));
    return RTX_OK;
}

}  // namespace rtx
