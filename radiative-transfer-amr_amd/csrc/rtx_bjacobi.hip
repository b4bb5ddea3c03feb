// Per-cell block-Jacobi preconditioner.
//
// The reference solves with PETSc bjacobi (src/rt/solve.py:93-95; one block
// per MPI rank with an ILU(0) sub-solver) and computes, but discards, a dense
// inverse of each column block (src/rt/preconditioner_matrix.py:259-262).
// Here the blocks are the angular cells: the cell's diagonal block of the
// interior system A_II (mass, interior convection, outflow faces, in-cell
// scattering and any self-coupling link), identity on inflow-boundary DOFs.
//
// Build: one workgroup per cell assembles the block from its structure
// (per-node angular scattering blocks, the diagonal, the x/y derivative
// lines) in LDS (or in a global scratch block when it does not fit), inverts
// it by Gauss-Jordan with partial pivoting (column-parallel elimination, one
// pivot search per wavefront-shuffle reduction), and stores the inverse
// column-major so the apply is a coalesced batched GEMV.
#include <algorithm>

#include "rtx_internal.h"

namespace rtx {

constexpr int NMq = RTX_NMAX;
constexpr int BJ_THREADS = 256;
#ifndef BJ_APPLY8_THREADS
#define BJ_APPLY8_THREADS 256    // threads per block of the FP8 apply
#endif
#ifndef BJ_APPLY16_THREADS
#define BJ_APPLY16_THREADS 512   // threads per block of the packed BF16 apply (measured
                                 // best of 128, 256, 512, 1024 on 125-DOF blocks)
#endif
constexpr int BJ_COLS = 128;     // column threads in the elimination
constexpr int BJ_SPLIT = 128;    // larger cells are split into groups of angles

__device__ inline bool bj_on_bdry(int bd, int i, int j, int nx, int ny) {
    return ((bd & 1) && i == nx - 1) || ((bd & 2) && j == ny - 1) || ((bd & 4) && i == 0) ||
           ((bd & 8) && j == 0);
}

// Assemble the cell's diagonal block into row-major A (n x n).
__device__ void bj_assemble(const DevMesh &m, int k, int restrict_, double *A, const double *cth,
                            const double *sth) {
    const int tid = threadIdx.x;
    const int c = m.cell_col[k];
    const int nx = m.col_nx[c], ny = m.col_ny[c], nth = m.cell_nth[k];
    const int n = nx * ny * nth, nyt = ny * nth, nsp = nx * ny;
    const double dx = m.col_geom[4 * c + 2] - m.col_geom[4 * c + 0];
    const double dy = m.col_geom[4 * c + 3] - m.col_geom[4 * c + 1];
    const double th0 = m.cell_geom[2 * k], dth = m.cell_geom[2 * k + 1] - th0;
    const double *wx = m.lgl_w + nx * NMq, *wy = m.lgl_w + ny * NMq, *wt = m.lg_w + nth * NMq;
    const double *Dx = m.lgl_D + size_t(nx) * NMq * NMq, *Dy = m.lgl_D + size_t(ny) * NMq * NMq;
    const int ao = m.cell_ang_off[k], na = m.col_nang[c];
    const double *G = m.G + m.col_G_off[c];
    const double *kap = m.kappa + m.col_sp_off[c], *sig = m.sigma + m.col_sp_off[c];
    for (int e = tid; e < n * n; e += BJ_THREADS) A[e] = 0.0;
    __syncthreads();
    // in-cell scattering: row (s, r), column (s, a)
    for (int e = tid; e < nsp * nth * nth; e += BJ_THREADS) {
        const int s = e / (nth * nth), r = (e / nth) % nth, a = e % nth;
        const int p = s / ny, q = s % ny;
        A[size_t(s * nth + r) * n + s * nth + a] =
            -dx * dy * wx[p] * wy[q] * sig[s] * G[size_t(ao + a) * na + ao + r];
    }
    __syncthreads();
    // mass and outflow faces on the diagonal
    const int quad = m.cell_quad[k];
    for (int t = tid; t < n; t += BJ_THREADS) {
        const int p = t / nyt, q = (t / nth) % ny, r = t % nth;
        double v = (dx * dy * dth * 0.125) * wx[p] * wy[q] * wt[r] * kap[p * ny + q];
        if (quad >= 0)
            for (int h = 0; h < 2; ++h) {
                const int F = (quad + h) & 3;
                const bool on = (F == 0 && p == nx - 1) || (F == 2 && p == 0) ||
                                (F == 1 && q == ny - 1) || (F == 3 && q == 0);
                if (on) {
                    const double tf = F == 0 ? cth[r] : (F == 1 ? sth[r] : (F == 2 ? -cth[r] : -sth[r]));
                    v += ((F & 1) ? dx : dy) * dth * 0.25 * ((F & 1) ? wx[p] : wy[q]) * wt[r] * tf;
                }
            }
        A[size_t(t) * n + t] += v;
    }
    __syncthreads();
    // -IC, x-derivative: row (p, q, r), column (i, q, r)
    for (int e = tid; e < n * nx; e += BJ_THREADS) {
        const int t = e / nx, i = e % nx;
        const int p = t / nyt, q = (t / nth) % ny, r = t % nth;
        A[size_t(t) * n + i * nyt + q * nth + r] -=
            (dy * dth * 0.25) * wy[q] * wt[r] * cth[r] * wx[i] * Dx[p * NMq + i];
    }
    __syncthreads();
    // -IC, y-derivative: row (p, q, r), column (p, j, r)
    for (int e = tid; e < n * ny; e += BJ_THREADS) {
        const int t = e / ny, j = e % ny;
        const int p = t / nyt, q = (t / nth) % ny, r = t % nth;
        A[size_t(t) * n + p * nyt + j * nth + r] -=
            (dx * dth * 0.25) * wx[p] * wt[r] * sth[r] * wy[j] * Dy[q * NMq + j];
    }
    __syncthreads();
    // a cell coupled to itself through a periodic face
    for (int l = m.link_start[k]; l < m.link_start[k + 1]; ++l) {
        if (m.link_src[l] != k) continue;
        const int F = m.link_face[l];
        const int n0 = (F & 1) ? nx : ny;
        const double cf = ((F & 1) ? dx : dy) * dth * 0.25;
        const double *Es = m.Es_pool + m.link_Es_off[l];
        const double *Et = m.Eth_pool + m.link_Eth_off[l];
        for (int e = tid; e < n0 * nth * n0 * nth; e += BJ_THREADS) {
            const int m0 = e / (nth * n0 * nth), r = (e / (n0 * nth)) % nth;
            const int m1 = (e / nth) % n0, a = e % nth;
            int row, col;
            if (F == 0) { row = (nx - 1) * nyt + m0 * nth + r; col = m1 * nth + a; }
            else if (F == 2) { row = m0 * nth + r; col = (nx - 1) * nyt + m1 * nth + a; }
            else if (F == 1) { row = m0 * nyt + (ny - 1) * nth + r; col = m1 * nyt + a; }
            else { row = m0 * nyt + r; col = m1 * nyt + (ny - 1) * nth + a; }
            A[size_t(row) * n + col] += cf * Es[m1 * n0 + m0] * Et[a * nth + r];
        }
        __syncthreads();
    }
    // identity rows/columns on inflow-boundary DOFs
    const int bd = restrict_ ? m.cell_bdry[k] : 0;
    if (bd) {
        for (int e = tid; e < n * n; e += BJ_THREADS) {
            const int row = e / n, col = e % n;
            const bool rb = bj_on_bdry(bd, row / nyt, (row / nth) % ny, nx, ny);
            const bool cb = bj_on_bdry(bd, col / nyt, (col / nth) % ny, nx, ny);
            if (rb || cb) A[e] = (row == col) ? 1.0 : 0.0;
        }
        __syncthreads();
    }
}

// In-place Gauss-Jordan inversion with partial pivoting of row-major A.
// Returns false on a zero pivot.  colk: n doubles of LDS.
__device__ bool bj_invert(double *A, int n, double *colk, int *perm, int *piv_row) {
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    for (int k = 0; k < n; ++k) {
        if (wave == 0) {
            double best = -1.0;
            int bi = k;
            for (int i = k + lane; i < n; i += 64) {
                const double v = fabs(A[size_t(i) * n + k]);
                if (v > best) { best = v; bi = i; }
            }
            for (int o = 32; o > 0; o >>= 1) {
                const double ov = __shfl_xor(best, o);
                const int oi = __shfl_xor(bi, o);
                if (ov > best || (ov == best && oi < bi)) { best = ov; bi = oi; }
            }
            if (lane == 0) {
                *piv_row = (best > 0.0) ? bi : -1;
                perm[k] = bi;
            }
        }
        __syncthreads();
        const int pr = *piv_row;
        if (pr < 0) return false;
        if (pr != k)
            for (int c = tid; c < n; c += BJ_THREADS) {
                const double t = A[size_t(k) * n + c];
                A[size_t(k) * n + c] = A[size_t(pr) * n + c];
                A[size_t(pr) * n + c] = t;
            }
        __syncthreads();
        for (int i = tid; i < n; i += BJ_THREADS) colk[i] = A[size_t(i) * n + k];
        __syncthreads();
        const double inv = 1.0 / colk[k];
        for (int c = tid; c < n; c += BJ_THREADS)
            A[size_t(k) * n + c] = (c == k) ? inv : A[size_t(k) * n + c] * inv;
        __syncthreads();
        // Rows are processed in register-staged batches of 8 so that the LDS
        // loads of a batch are in flight together (the row updates of one
        // column are independent).
        const int g = tid / BJ_COLS, ng = BJ_THREADS / BJ_COLS;
        for (int c = tid % BJ_COLS; c < n; c += BJ_COLS) {
            const double akc = A[size_t(k) * n + c];
            const bool pivcol = (c == k);
            int i = g;
            for (; i + 7 * ng < n; i += 8 * ng) {
                double v[8], f[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    v[u] = A[size_t(i + u * ng) * n + c];
                    f[u] = colk[i + u * ng];
                }
#pragma unroll
                for (int u = 0; u < 8; ++u)
                    v[u] = (i + u * ng == k) ? v[u] : (pivcol ? -f[u] * inv : v[u] - f[u] * akc);
#pragma unroll
                for (int u = 0; u < 8; ++u) A[size_t(i + u * ng) * n + c] = v[u];
            }
            for (; i < n; i += ng)
                if (i != k) A[size_t(i) * n + c] = pivcol ? -colk[i] * inv : A[size_t(i) * n + c] - colk[i] * akc;
        }
        __syncthreads();
    }
    // undo the row interchanges as column interchanges, in reverse order
    for (int k = n - 1; k >= 0; --k) {
        const int pr = perm[k];
        if (pr != k)
            for (int i = tid; i < n; i += BJ_THREADS) {
                const double t = A[size_t(i) * n + k];
                A[size_t(i) * n + k] = A[size_t(i) * n + pr];
                A[size_t(i) * n + pr] = t;
            }
        __syncthreads();
    }
    return true;
}

// One workgroup per cell.  USE_LDS: the block lives in dynamic LDS; else in
// the global scratch block of the cell.  The inverse goes to out,
// column-major.
template <bool USE_LDS>
__global__ __launch_bounds__(BJ_THREADS) void k_bj_build(DevMesh m, int restrict_,
                                                        const int64_t *inv_off,
                                                        double *__restrict__ out,
                                                        double *__restrict__ scratch,
                                                        int *__restrict__ status) {
    extern __shared__ double lds[];
    __shared__ double cth[NMq], sth[NMq];
    __shared__ double colk[RTX_NMAX * RTX_NMAX * RTX_NMAX];
    __shared__ int perm[RTX_NMAX * RTX_NMAX * RTX_NMAX];
    __shared__ int piv_row;
    const int k = blockIdx.x, tid = threadIdx.x;
    const int c = m.cell_col[k];
    const int nth = m.cell_nth[k];
    const int n = m.col_nx[c] * m.col_ny[c] * nth;
    if (tid < nth) {
        const double th0 = m.cell_geom[2 * k], dth = m.cell_geom[2 * k + 1] - th0;
        double s, cc;
        sincos(th0 + dth * 0.5 * (m.lg_x[nth * NMq + tid] + 1.0), &s, &cc);
        cth[tid] = cc;
        sth[tid] = s;
    }
    __syncthreads();
    double *A = USE_LDS ? lds : scratch + inv_off[k];
    bj_assemble(m, k, restrict_, A, cth, sth);
    const bool ok = bj_invert(A, n, colk, perm, &piv_row);
    if (!ok) {
        if (tid == 0) atomicMax(status, k + 1);
        return;
    }
    double *o = out + inv_off[k];
    for (int col = 0; col < n; ++col)
        for (int row = tid; row < n; row += BJ_THREADS) o[size_t(col) * n + row] = A[size_t(row) * n + col];
}

// ---- register-resident build for blocks of up to 16 RR DOFs.
//
// The block is distributed cyclically over a 16 x TJ thread grid (thread
// (ti, tj) holds rows ti + 16 u, u < RR, and columns tj + TJ v, v < CC; 16 RR
// = TJ CC), each entry evaluated directly from the operator's structure
// (same terms, same summation order as bj_assemble).  Inversion is
// Gauss-Jordan with partial pivoting done as a sequence of sweeps on the
// pivot entries (p_k, k) without moving rows: only the pivot column and the
// pivot row pass through LDS, so a workgroup needs about 3 KB of LDS
// instead of n^2 doubles and several cells run per CU.  After all sweeps
// the register matrix M holds the inverse with permuted indices:
// inv[q_i][p_j] = M[i][j], q = p^-1.
struct BJCell {
    int nx, ny, nth, n, nyt, ao, na, quad, bd;
    double dx, dy, dth;
    const double *wx, *wy, *wt, *Dx, *Dy, *G, *kap, *sig;
};

struct BJSelf {                  // links of a cell to itself (periodic faces), in LDS
    int n;
    int F[4];
    int es[4], et[4];
};

__device__ inline double bj_entry(const DevMesh &m, const BJCell &x, const BJSelf &sl,
                                  const double *cth, const double *sth, int t, int u) {
    const int p = t / x.nyt, q = (t / x.nth) % x.ny, r = t % x.nth;
    const int i = u / x.nyt, j = (u / x.nth) % x.ny, a = u % x.nth;
    if (x.bd) {
        const bool rb = bj_on_bdry(x.bd, p, q, x.nx, x.ny);
        const bool cb = bj_on_bdry(x.bd, i, j, x.nx, x.ny);
        if (rb || cb) return (t == u) ? 1.0 : 0.0;
    }
    double v = 0.0;
    if (p == i && q == j) {
        v = -x.dx * x.dy * x.wx[p] * x.wy[q] * x.sig[p * x.ny + q] *
            x.G[size_t(x.ao + a) * x.na + x.ao + r];
        if (r == a) {
            double d = (x.dx * x.dy * x.dth * 0.125) * x.wx[p] * x.wy[q] * x.wt[r] * x.kap[p * x.ny + q];
            if (x.quad >= 0)
                for (int h = 0; h < 2; ++h) {
                    const int F = (x.quad + h) & 3;
                    const bool on = (F == 0 && p == x.nx - 1) || (F == 2 && p == 0) ||
                                    (F == 1 && q == x.ny - 1) || (F == 3 && q == 0);
                    if (on) {
                        const double tf = F == 0 ? cth[r] : (F == 1 ? sth[r] : (F == 2 ? -cth[r] : -sth[r]));
                        d += ((F & 1) ? x.dx : x.dy) * x.dth * 0.25 * ((F & 1) ? x.wx[p] : x.wy[q]) *
                             x.wt[r] * tf;
                    }
                }
            v += d;
        }
    }
    if (q == j && r == a)
        v -= (x.dy * x.dth * 0.25) * x.wy[q] * x.wt[r] * cth[r] * x.wx[i] * x.Dx[p * NMq + i];
    if (p == i && r == a)
        v -= (x.dx * x.dth * 0.25) * x.wx[p] * x.wt[r] * sth[r] * x.wy[j] * x.Dy[q * NMq + j];
    for (int l = 0; l < sl.n; ++l) {
        const int F = sl.F[l];
        int m0, m1;
        bool hit;
        if (F == 0) { hit = p == x.nx - 1 && i == 0; m0 = q; m1 = j; }
        else if (F == 2) { hit = p == 0 && i == x.nx - 1; m0 = q; m1 = j; }
        else if (F == 1) { hit = q == x.ny - 1 && j == 0; m0 = p; m1 = i; }
        else { hit = q == 0 && j == x.ny - 1; m0 = p; m1 = i; }
        if (hit) {
            const int n0 = (F & 1) ? x.nx : x.ny;
            v += ((F & 1) ? x.dx : x.dy) * x.dth * 0.25 * m.Es_pool[sl.es[l] + m1 * n0 + m0] *
                 m.Eth_pool[sl.et[l] + a * x.nth + r];
        }
    }
    return v;
}

// One workgroup per block, the block's index from the size-class list bl
// (blocks of at most 16 RR DOFs).  T: the elimination's arithmetic and the
// inverse's storage (FP64, or FP32 for the reduced-precision copies alone;
// the entries are assembled in FP64 either way).
template <int RR, int CC, int TJ, typename T = double>
__global__ __launch_bounds__(16 * TJ)
__attribute__((amdgpu_waves_per_eu(TJ == 24 ? 5 : (TJ == 32 && sizeof(T) == 4 ? 6 : 1))))
void k_bj_build_reg(DevMesh m, int restrict_,
                                                         const int64_t *inv_off,
                                                         const int32_t *__restrict__ blk,
                                                         const int32_t *__restrict__ bl,
                                                         T *__restrict__ out,
                                                         int *__restrict__ status) {
    constexpr int NR = 16 * RR;
    static_assert(NR == TJ * CC, "square register tiles");
    static_assert(NR <= 128, "pivot rows fit the key's 7 low bits");
    __shared__ double cth[NMq], sth[NMq];
    __shared__ T colk[2][NR];
    __shared__ T prow[NR];
    // pivot key of each step (double-buffered): the bits of |value| with the
    // low 7 mantissa bits replaced by 127 - row, so the largest key is the
    // largest value (to 45 mantissa bits), the lowest row among ties
    __shared__ unsigned long long red_key[2];
    __shared__ int perm[NR], qinv[NR];
    __shared__ int used[NR];
    __shared__ BJSelf sl;
    const int b = bl[blockIdx.x], tid = threadIdx.x, ti = tid / TJ, tj = tid % TJ;
    const int k = blk[3 * b], a0 = blk[3 * b + 1], g = blk[3 * b + 2];
    BJCell x;
    const int c = m.cell_col[k];
    x.nx = m.col_nx[c];
    x.ny = m.col_ny[c];
    x.nth = m.cell_nth[k];
    x.n = x.nx * x.ny * x.nth;
    x.nyt = x.ny * x.nth;
    x.dx = m.col_geom[4 * c + 2] - m.col_geom[4 * c + 0];
    x.dy = m.col_geom[4 * c + 3] - m.col_geom[4 * c + 1];
    const double th0 = m.cell_geom[2 * k];
    x.dth = m.cell_geom[2 * k + 1] - th0;
    x.wx = m.lgl_w + x.nx * NMq;
    x.wy = m.lgl_w + x.ny * NMq;
    x.wt = m.lg_w + x.nth * NMq;
    x.Dx = m.lgl_D + size_t(x.nx) * NMq * NMq;
    x.Dy = m.lgl_D + size_t(x.ny) * NMq * NMq;
    x.ao = m.cell_ang_off[k];
    x.na = m.col_nang[c];
    x.G = m.G + m.col_G_off[c];
    x.kap = m.kappa + m.col_sp_off[c];
    x.sig = m.sigma + m.col_sp_off[c];
    x.quad = m.cell_quad[k];
    x.bd = restrict_ ? m.cell_bdry[k] : 0;
    if (tid == 0) {
        int ns = 0;
        for (int l = m.link_start[k]; l < m.link_start[k + 1] && ns < 4; ++l)
            if (m.link_src[l] == k) {
                sl.F[ns] = m.link_face[l];
                sl.es[ns] = m.link_Es_off[l];
                sl.et[ns] = m.link_Eth_off[l];
                ++ns;
            }
        sl.n = ns;
    }
    if (tid < x.nth) {
        double s, cc;
        sincos(th0 + x.dth * 0.5 * (m.lg_x[x.nth * NMq + tid] + 1.0), &s, &cc);
        cth[tid] = cc;
        sth[tid] = s;
    }
    if (tid < NR) used[tid] = 0;
    if (tid < 2) red_key[tid] = 0ull;
    // the block: angles a0 .. a0 + g - 1 of the cell at every spatial node;
    // block index t <-> cell DOF (t / g) nth + a0 + t % g
    const int n = (x.n / x.nth) * g;
    __syncthreads();
    T A[RR][CC];
#pragma unroll
    for (int u = 0; u < RR; ++u)
#pragma unroll
        for (int v = 0; v < CC; ++v) {
            const int row = ti + 16 * u, col = tj + TJ * v;
            A[u][v] = (row < n && col < n)
                          ? T(bj_entry(m, x, sl, cth, sth, (row / g) * x.nth + a0 + row % g,
                                       (col / g) * x.nth + a0 + col % g))
                          : T(0);
        }
    // Every step is written without per-element branches (selects around
    // one FMA per element), so the wave runs straight-line code.
    for (int kk = 0; kk < n; ++kk) {
        const int buf = kk & 1, vk = kk / TJ;
        if (tj == kk % TJ) {           // the 16 owners of column kk
            double best = -1.0;
            int bi = n;
#pragma unroll
            for (int u = 0; u < RR; ++u) {
                T val = A[u][0];
#pragma unroll
                for (int v = 1; v < CC; ++v) val = (v == vk) ? A[u][v] : val;
                const int row = ti + 16 * u;
                const bool ok = row < n;
                if (ok) colk[buf][row] = val;
                const double av = fabs(double(val));
                const bool better = ok && !used[ok ? row : 0] && (av > best || (av == best && row < bi));
                best = better ? av : best;
                bi = better ? row : bi;
            }
            // no candidate (or a zero column): key 0, which fails below
            const unsigned long long key =
                (bi < n && best > 0.0)
                    ? ((static_cast<unsigned long long>(__double_as_longlong(best)) & ~0x7Full) |
                       static_cast<unsigned long long>(127 - bi))
                    : 0ull;
            atomicMax(&red_key[buf], key);
        }
        __syncthreads();
        const unsigned long long key = red_key[buf];
        const double best = __longlong_as_double(static_cast<long long>(key & ~0x7Full));
        const int pr = 127 - int(key & 0x7Full);
        // the other buffer is next read after this step's second barrier
        if (tid == 0) red_key[buf ^ 1] = 0ull;
        if (!(best > 0.0) || pr >= n) {
            if (tid == 0) atomicMax(status, k + 1);
            return;                    // uniform across the workgroup
        }
        if (ti == (pr & 15)) {
            const int up = pr >> 4;
#pragma unroll
            for (int v = 0; v < CC; ++v) {
                T pval = A[0][v];
#pragma unroll
                for (int u = 1; u < RR; ++u) pval = (u == up) ? A[u][v] : pval;
                const int col = tj + TJ * v;
                if (col < n) prow[col] = pval;
            }
        }
        if (tid == 0) {
            used[pr] = 1;
            perm[kk] = pr;
        }
        __syncthreads();
        // in-place Gauss-Jordan: pivot row -> row / p with 1 / p in column
        // kk; other rows -> row - c_r (pivot row / p) with -c_r / p in column
        // kk.  With the pivot row's kk entry taken as 1 and the other rows'
        // kk entries as 0 this is one FMA per element.
        const T invp = T(1) / colk[buf][pr];
        T pv[CC];
#pragma unroll
        for (int v = 0; v < CC; ++v) {
            const int col = tj + TJ * v;
            const T pr_v = (col == kk) ? T(1) : ((col < n) ? prow[col] : T(0));
            pv[v] = pr_v * invp;
        }
#pragma unroll
        for (int u = 0; u < RR; ++u) {
            const int row = ti + 16 * u;
            const bool isp = row == pr;
            const T ck = (row < n) ? colk[buf][row] : T(0);
#pragma unroll
            for (int v = 0; v < CC; ++v) {
                const int col = tj + TJ * v;
                const T base = (col == kk) ? T(0) : A[u][v];
                const T upd = fma(-ck, pv[v], base);
                A[u][v] = isp ? pv[v] : upd;
            }
        }
    }
    __syncthreads();
    if (tid < n) qinv[perm[tid]] = tid;
    __syncthreads();
    T *o = out + inv_off[b];
#pragma unroll
    for (int u = 0; u < RR; ++u)
#pragma unroll
        for (int v = 0; v < CC; ++v) {
            const int row = ti + 16 * u, col = tj + TJ * v;
            if (row < n && col < n) o[size_t(perm[col]) * n + qinv[row]] = A[u][v];
        }
}

__global__ void k_to_f32_bj(const double *__restrict__ s, float *__restrict__ d, int64_t n) {
    int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    for (; i < n; i += int64_t(gridDim.x) * blockDim.x) d[i] = float(s[i]);
}

// z_K = inv_K r_K, inv column-major.  One workgroup per cell; the columns
// are split across the four wavefronts' partial sums when n is small.
// BF16 storage of an inverse entry, widened on load.
template <typename T, typename M>
__global__ __launch_bounds__(BJ_THREADS) void k_bj_apply(DevMesh m, const int64_t *inv_off,
                                                        const int4 *__restrict__ apd,
                                                        const M *__restrict__ inv,
                                                        const T *__restrict__ r,
                                                        T *__restrict__ z) {
    __shared__ double rs[RTX_NMAX * RTX_NMAX * RTX_NMAX];
    __shared__ double part[BJ_THREADS];
    const int b = blockIdx.x, tid = threadIdx.x;
    const int4 d = apd[b];                // n, first DOF, cell angles, angles g
    const int n = d.x, nth = d.z, g = d.w;
    // block index t <-> cell DOF (t / g) nth + a0 + t % g (identity for a whole cell)
    const int off = d.y;
    for (int i = tid; i < n; i += BJ_THREADS) rs[i] = double(r[off + (i / g) * nth + i % g]);
    __syncthreads();
    const M *B = inv + inv_off[b];
    if (n <= BJ_THREADS / 2) {
        // two column halves per output row
        const int half = tid / (BJ_THREADS / 2), i = tid % (BJ_THREADS / 2);
        double acc = 0.0;
        if (i < n) {
            const int j0 = half * ((n + 1) / 2), j1 = half ? n : (n + 1) / 2;
#pragma unroll 4
            for (int j = j0; j < j1; ++j) acc += double(B[size_t(j) * n + i]) * rs[j];
        }
        part[tid] = acc;
        __syncthreads();
        if (tid < n) z[off + (tid / g) * nth + tid % g] = T(part[tid] + part[tid + BJ_THREADS / 2]);
    } else {
        for (int i = tid; i < n; i += BJ_THREADS) {
            double acc = 0.0;
#pragma unroll 4
            for (int j = 0; j < n; ++j) acc += double(B[size_t(j) * n + i]) * rs[j];
            z[off + (i / g) * nth + i % g] = T(acc);
        }
    }
}

__device__ inline uint16_t bj_to_bf16(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);
    return uint16_t(u >> 16);
}

__device__ inline int bj_block_n(const DevMesh &m, const int32_t *blk, int b) {
    const int k = blk[3 * b], g = blk[3 * b + 2];
    const int c = m.cell_col[k];
    return m.col_nx[c] * m.col_ny[c] * g;
}

// Packed BF16 copy: column j of block b at off16[b] + j * ldp, ldp = n
// rounded up to 8 (padding rows zero), so every column starts on 16 bytes.
__global__ __launch_bounds__(BJ_THREADS) void k_bj_pack_bf16(DevMesh m, const int64_t *off,
                                                             const int64_t *off16,
                                                             const int32_t *__restrict__ blk,
                                                             const float *__restrict__ src,
                                                             uint16_t *__restrict__ dst) {
    const int b = blockIdx.x;
    const int n = bj_block_n(m, blk, b), ldp = (n + 7) / 8 * 8;
    const float *S = src + off[b];
    uint16_t *D = dst + off16[b];
    for (int e = threadIdx.x; e < n * ldp; e += BJ_THREADS) {
        const int j = e / ldp, i = e % ldp;
        D[e] = i < n ? bj_to_bf16(S[size_t(j) * n + i]) : uint16_t(0);
    }
}

// z = B r for one block from the packed BF16 inverse: thread (cg, rg) owns
// rows 8 rg .. 8 rg + 7 and columns cg, cg + CG, ...; one 16-byte load per
// column; the CG partial sums of each row are added in LDS.
template <int NT>
__global__ __launch_bounds__(NT) void k_bj_apply_bf16(DevMesh m, const int64_t *off16,
                                                              const int4 *__restrict__ apd,
                                                              const uint16_t *__restrict__ inv,
                                                              const float *__restrict__ r,
                                                              float *__restrict__ z) {
    __shared__ float rs[RTX_NMAX * RTX_NMAX * RTX_NMAX];
    __shared__ float part[NT * 8];
    const int b = blockIdx.x, tid = threadIdx.x;
    const int4 d = apd[b];                // n, first DOF, cell angles, angles g
    const int n = d.x, nth = d.z, g = d.w;
    const int ldp = (n + 7) / 8 * 8, RG = ldp / 8, CG = NT / RG;
    const int off = d.y;
    const int rg = tid % RG, cg = tid / RG;
    const uint4 *B = reinterpret_cast<const uint4 *>(inv + off16[b]) + rg;
    const int ldq = ldp / 8;
    // the first 8 columns of this thread are fetched before the wait on r,
    // so the inverse loads and the r gather are in flight together
    constexpr int PF = 2048 / NT;
    uint4 q[PF];
#pragma unroll
    for (int u = 0; u < PF; ++u) {
        const int j = cg + u * CG;
        q[u] = (cg < CG && j < n) ? B[size_t(j) * ldq] : make_uint4(0u, 0u, 0u, 0u);
    }
    for (int i = tid; i < n; i += NT) rs[i] = r[off + (i / g) * nth + i % g];
    __syncthreads();
    float acc[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    auto fma8 = [&](const uint4 &v, float rj) {
        const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            acc[2 * u] += __uint_as_float(w[u] << 16) * rj;
            acc[2 * u + 1] += __uint_as_float(w[u] & 0xffff0000u) * rj;
        }
    };
    if (cg < CG) {
#pragma unroll
        for (int u = 0; u < PF; ++u) {
            const int j = cg + u * CG;
            if (j < n) fma8(q[u], rs[j]);
        }
#pragma unroll 4
        for (int j = cg + PF * CG; j < n; j += CG) fma8(B[size_t(j) * ldq], rs[j]);
#pragma unroll
        for (int u = 0; u < 8; ++u) part[cg * ldp + rg * 8 + u] = acc[u];
    }
    __syncthreads();
    for (int i = tid; i < n; i += NT) {
        float s = 0.f;
        for (int q = 0; q < CG; ++q) s += part[q * ldp + i];
        z[off + (i / g) * nth + i % g] = s;
    }
}

// FP8 e4m3 (bias 7, largest finite 448, no infinities): nearest code,
// ties to even; |v| above 448 saturates.
__device__ inline uint8_t bj_to_e4m3(float v) {
    const uint8_t sgn = v < 0.f ? 0x80 : 0;
    const float a = fabsf(v);
    if (!(a > 0.f)) return sgn;
    if (a >= 448.f) return sgn | 0x7E;
    int e;
    (void)frexpf(a, &e);                       // a = f 2^e, f in [0.5, 1)
    int E = e - 1;                             // a = (1 + m / 8) 2^E for normals
    if (E < -6) {                              // subnormals: m 2^-9
        const int mq = int(rintf(a * 512.f));
        return sgn | uint8_t(mq >= 8 ? 8 : mq);
    }
    int mq = int(rintf((ldexpf(a, -E) - 1.f) * 8.f));
    if (mq == 8) {
        mq = 0;
        ++E;
    }
    const int eb = E + 7;
    if (eb > 15 || (eb == 15 && mq > 6)) return sgn | 0x7E;
    return sgn | uint8_t((eb << 3) | mq);
}

// Column scales of one block for the coded copies: scale_j = max_i |B_ij| /
// div (1 for an all-zero column) into sc[j] and its reciprocal into inv_s[j].
// One wavefront per column: coalesced loads of the column and a shuffle
// reduction (the maximum does not depend on the order).
__device__ inline void bj_col_scales(const float *__restrict__ S, int n, float div,
                                     float *__restrict__ sc, float *inv_s) {
    static_assert(BJ_THREADS % 64 == 0, "whole wavefronts");
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int j = wave; j < n; j += BJ_THREADS / 64) {
        float mx = 0.f;
        for (int i = lane; i < n; i += 64) mx = fmaxf(mx, fabsf(S[size_t(j) * n + i]));
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) mx = fmaxf(mx, __shfl_xor(mx, o));
        if (lane == 0) {
            const float v = mx > 0.f ? mx / div : 1.f;
            sc[j] = v;
            inv_s[j] = 1.f / v;
        }
    }
}

// FP8 copy of block b: column j of codes at off8[b] + j * ldp (ldp = n
// rounded up to 16, padding zero), scale_j = max_i |B_ij| / 448 at
// scales[soff[b] + j]; code = e4m3(B_ij / scale_j).
__global__ __launch_bounds__(BJ_THREADS) void k_bj_pack_fp8(DevMesh m, const int64_t *off,
                                                            const int64_t *off8,
                                                            const int64_t *soff,
                                                            const int32_t *__restrict__ blk,
                                                            const float *__restrict__ src,
                                                            uint8_t *__restrict__ dst,
                                                            float *__restrict__ scales) {
    __shared__ float inv_s[RTX_NMAX * RTX_NMAX * RTX_NMAX];
    const int b = blockIdx.x;
    const int n = bj_block_n(m, blk, b), ldp = (n + 15) / 16 * 16;
    const float *S = src + off[b];
    uint8_t *D = dst + off8[b];
    bj_col_scales(S, n, 448.f, scales + soff[b], inv_s);
    __syncthreads();
    for (int e = threadIdx.x; e < n * ldp; e += BJ_THREADS) {
        const int j = e / ldp, i = e % ldp;
        D[e] = i < n ? bj_to_e4m3(S[size_t(j) * n + i] * inv_s[j]) : uint8_t(0);
    }
}

// FP4 e2m1 (OCP: bias 1, one mantissa bit, largest 6, no infinities or
// NaN; magnitudes 0, 0.5, 1, 1.5, 2, 3, 4, 6 for codes 0..7): nearest code,
// ties to the even code; |v| above 6 saturates.
__device__ inline uint32_t bj_to_e2m1(float v) {
    const uint32_t sgn = v < 0.f ? 8u : 0u;
    const float a = fabsf(v);
    // midpoints between consecutive magnitudes; at a midpoint the even code
    // (0, 2, 4, 6) wins
    uint32_t c;
    if (a <= 0.25f) c = 0;
    else if (a < 0.75f) c = 1;
    else if (a <= 1.25f) c = 2;
    else if (a < 1.75f) c = 3;
    else if (a <= 2.5f) c = 4;
    else if (a < 3.5f) c = 5;
    else if (a <= 5.f) c = 6;
    else c = 7;
    return sgn | c;
}

// FP4 copy of block b: the FP8 copy's layout at half the bytes (column j's
// codes at off8[b] / 2 + j * ldp / 2, row i in byte i / 2, the even row in
// the low nibble), scale_j = max_i |B_ij| / 6 at scales[soff[b] + j];
// code = e2m1(B_ij / scale_j).
__global__ __launch_bounds__(BJ_THREADS) void k_bj_pack_fp4(DevMesh m, const int64_t *off,
                                                            const int64_t *off8,
                                                            const int64_t *soff,
                                                            const int32_t *__restrict__ blk,
                                                            const float *__restrict__ src,
                                                            uint8_t *__restrict__ dst,
                                                            float *__restrict__ scales) {
    __shared__ float inv_s[RTX_NMAX * RTX_NMAX * RTX_NMAX];
    const int b = blockIdx.x;
    const int n = bj_block_n(m, blk, b), ldp = (n + 15) / 16 * 16, hp = ldp / 2;
    const float *S = src + off[b];
    uint8_t *D = dst + off8[b] / 2;
    bj_col_scales(S, n, 6.f, scales + soff[b], inv_s);
    __syncthreads();
    for (int e = threadIdx.x; e < n * hp; e += BJ_THREADS) {
        const int j = e / hp, i = 2 * (e % hp);
        const uint32_t lo = i < n ? bj_to_e2m1(S[size_t(j) * n + i] * inv_s[j]) : 0u;
        const uint32_t hi = i + 1 < n ? bj_to_e2m1(S[size_t(j) * n + i + 1] * inv_s[j]) : 0u;
        D[e] = uint8_t(lo | (hi << 4));
    }
}

// Sixteen consecutive codes of one column (rows 16 rg .. 16 rg + 15),
// decoded to FP32 by the gfx950 hardware conversions and multiplied into
// the row sums: acc[v] += code_v * rj.  FP8 e4m3: one 16-byte load, two
// codes per conversion; FP4 e2m1: one 8-byte load, two codes per
// conversion (scale 1).
template <int BITS>
struct BjCodes;
template <>
struct BjCodes<8> {
    typedef uint4 V;
    static constexpr int SHIFT = 0;    // byte offset of a block = off8[b] >> SHIFT
    __device__ static V zero() { return make_uint4(0u, 0u, 0u, 0u); }
    __device__ static void fma16(const V &v, float rj, float (&acc)[16]) {
        typedef float f2 __attribute__((ext_vector_type(2)));
        const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const f2 lo = __builtin_amdgcn_cvt_pk_f32_fp8(int(w[u]), false);
            const f2 hi = __builtin_amdgcn_cvt_pk_f32_fp8(int(w[u]), true);
            acc[4 * u + 0] += lo.x * rj;
            acc[4 * u + 1] += lo.y * rj;
            acc[4 * u + 2] += hi.x * rj;
            acc[4 * u + 3] += hi.y * rj;
        }
    }
};
template <>
struct BjCodes<4> {
    typedef uint2 V;
    static constexpr int SHIFT = 1;
    __device__ static V zero() { return make_uint2(0u, 0u); }
    __device__ static void fma16(const V &v, float rj, float (&acc)[16]) {
        typedef float f2 __attribute__((ext_vector_type(2)));
        const uint32_t w[2] = {v.x, v.y};
#pragma unroll
        for (int u = 0; u < 2; ++u) {
            // byte s of word u holds rows 8 u + 2 s (low nibble) and + 1
            const f2 p0 = __builtin_amdgcn_cvt_scalef32_pk_f32_fp4(w[u], 1.f, 0);
            const f2 p1 = __builtin_amdgcn_cvt_scalef32_pk_f32_fp4(w[u], 1.f, 1);
            const f2 p2 = __builtin_amdgcn_cvt_scalef32_pk_f32_fp4(w[u], 1.f, 2);
            const f2 p3 = __builtin_amdgcn_cvt_scalef32_pk_f32_fp4(w[u], 1.f, 3);
            acc[8 * u + 0] += p0.x * rj;
            acc[8 * u + 1] += p0.y * rj;
            acc[8 * u + 2] += p1.x * rj;
            acc[8 * u + 3] += p1.y * rj;
            acc[8 * u + 4] += p2.x * rj;
            acc[8 * u + 5] += p2.y * rj;
            acc[8 * u + 6] += p3.x * rj;
            acc[8 * u + 7] += p3.y * rj;
        }
    }
};

// z = B r from the FP8 (BITS 8) or FP4 (BITS 4) pool: thread (cg, rg) owns
// rows 16 rg .. 16 rg + 15 and columns cg, cg + CG, ...; one load of sixteen
// codes per column, decoded by the hardware conversion; the column scales
// are folded into r.
template <int NT, typename RT = float, int BITS = 8>
__global__ __launch_bounds__(NT) void k_bj_apply_fp8(DevMesh m, const int64_t *off8,
                                                     const int64_t *soff,
                                                     const int4 *__restrict__ apd,
                                                     const uint8_t *__restrict__ inv,
                                                     const float *__restrict__ scales,
                                                     const RT *__restrict__ r,
                                                     float *__restrict__ z) {
    __shared__ float rs[RTX_NMAX * RTX_NMAX * RTX_NMAX];
    __shared__ float part[NT * 16];
    const int b = blockIdx.x, tid = threadIdx.x;
    const int4 d = apd[b];                // n, first DOF, cell angles, angles g
    const int n = d.x, nth = d.z, g = d.w;
    const int ldp = (n + 15) / 16 * 16, RG = ldp / 16, CG = NT / RG;
    const int off = d.y;
    const int rg = tid % RG, cg = tid / RG;
    typedef BjCodes<BITS> Codes;
    typedef typename Codes::V CV;
    const CV *B = reinterpret_cast<const CV *>(inv + (off8[b] >> Codes::SHIFT)) + rg;
    const int ldq = ldp / 16;
    constexpr int PF = 4;
    CV q[PF];
#pragma unroll
    for (int u = 0; u < PF; ++u) {
        const int j = cg + u * CG;
        q[u] = (cg < CG && j < n) ? B[size_t(j) * ldq] : Codes::zero();
    }
    const float *sc = scales + soff[b];
    for (int i = tid; i < n; i += NT) rs[i] = float(r[off + (i / g) * nth + i % g]) * sc[i];
    __syncthreads();
    float acc[16];
#pragma unroll
    for (int u = 0; u < 16; ++u) acc[u] = 0.f;
    auto fma16 = [&](const CV &v, float rj) { Codes::fma16(v, rj, acc); };
    if (cg < CG) {
#pragma unroll
        for (int u = 0; u < PF; ++u) {
            const int j = cg + u * CG;
            if (j < n) fma16(q[u], rs[j]);
        }
        for (int j = cg + PF * CG; j < n; j += CG) fma16(B[size_t(j) * ldq], rs[j]);
#pragma unroll
        for (int u = 0; u < 16; ++u) part[cg * ldp + rg * 16 + u] = acc[u];
    }
    __syncthreads();
    for (int i = tid; i < n; i += NT) {
        float s = 0.f;
        for (int qq = 0; qq < CG; ++qq) s += part[qq * ldp + i];
        z[off + (i / g) * nth + i % g] = s;
    }
}

// z = B r from the FP8 (or FP4) pool, one wavefront per block of at most 128 rows
// (the register-built blocks), four blocks per workgroup: lane (rg, cg),
// rg = lane mod RG (RG = ldp / 16 row groups of 16), owns rows
// 16 rg .. 16 rg + 15 and columns cg, cg + CG, ... (CG = 64 / RG, at most 16
// columns per lane), all its 16-byte code loads and r entries in flight
// together; the CG partial sums of a row are added in a per-wave LDS tile.
// A quarter of the threads per block of k_bj_apply_fp8 and no block-wide
// barrier, so more blocks' loads are in flight per CU.  LOOP: the grid holds
// at most seven workgroups per CU and each wavefront applies blocks b,
// b + W, ... (W wavefronts in the grid), the next block's descriptor loaded
// while the current block is applied; taken for pools that fit the 256 MB
// last-level cache (FP4 on 16,384 blocks of 125: 26.9 against 28.5 us).
// Streaming from HBM, one block per wavefront was faster (FP4 on 65,536
// blocks of 125: 110 against 120 us).
constexpr int BJ_W_MAXN = 128;
constexpr int BJ_W_KMAX = 16;      // columns per lane: ceil(n / CG) <= 16 for n <= 128
constexpr int BJ_W_WAVES = 4;
template <typename RT = float, int BITS = 8, bool LOOP = false>
__global__ __launch_bounds__(64 * BJ_W_WAVES) void k_bj_apply_fp8_w(
    DevMesh m, const int64_t *off8, const int64_t *soff, const int4 *__restrict__ apd, int nblk,
    const uint8_t *__restrict__ inv, const float *__restrict__ scales,
    const RT *__restrict__ r, float *__restrict__ z) {
    __shared__ float part[BJ_W_WAVES][BJ_W_MAXN * 8];   // CG x ldp <= 1024 per wave
    __shared__ float rsh[BJ_W_WAVES][BJ_W_MAXN];        // r_j s_j of the block
    const int wave = LOOP ? __builtin_amdgcn_readfirstlane(threadIdx.x >> 6) : threadIdx.x >> 6;
    const int lane = threadIdx.x & 63;
    const int nw = LOOP ? gridDim.x * BJ_W_WAVES : nblk;
    int b = blockIdx.x * BJ_W_WAVES + wave;
    if (b >= nblk) return;                               // whole waves; no block barrier below
    int4 dcur = apd[b];
    int64_t o8cur = off8[b], socur = soff[b];
    for (; b < nblk; b += nw) {
        // the next block's descriptor in flight while this one is applied
        int4 dnx = dcur;
        int64_t o8nx = o8cur, sonx = socur;
        if constexpr (LOOP) {
            const int bn = b + nw < nblk ? b + nw : b;
            dnx = apd[bn];
            o8nx = off8[bn];
            sonx = soff[bn];
        }
        const int4 d = dcur;                  // n, first DOF, cell angles, angles g
        const int n = d.x, nth = d.z, g = d.w;
        const int ldp = (n + 15) / 16 * 16, RG = ldp / 16, CG = 64 / RG;
        const int rg = lane % RG, cg = lane / RG;
        const int off = d.y;
        typedef BjCodes<BITS> Codes;
        typedef typename Codes::V CV;
        const CV *B = reinterpret_cast<const CV *>(inv + (o8cur >> Codes::SHIFT)) + rg;
        const int ldq = ldp / 16;
        const float *sc = scales + socur;
        // r_j s_j once per block entry (two per lane)
        float *rs = rsh[wave];
        for (int i = lane; i < n; i += 64) rs[i] = float(r[off + (i / g) * nth + i % g]) * sc[i];
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        float acc[16];
#pragma unroll
        for (int v = 0; v < 16; ++v) acc[v] = 0.f;
        // FP8: two batches of eight columns, each batch's loads in flight
        // together (all sixteen at once: 132 VGPRs, three waves per SIMD,
        // slower); FP4: the half-size loads of all sixteen columns at once
        constexpr int BATCH = BITS == 4 ? BJ_W_KMAX : BJ_W_KMAX / 2;
#pragma unroll 1
        for (int u0 = 0; u0 < BJ_W_KMAX; u0 += BATCH) {
            CV q[BATCH];
            float rj[BATCH];
#pragma unroll
            for (int u = 0; u < BATCH; ++u) {
                const int j = cg + (u0 + u) * CG;
                const bool ok = cg < CG && j < n;
                q[u] = ok ? B[size_t(j) * ldq] : Codes::zero();
                rj[u] = ok ? rs[j] : 0.f;
            }
#pragma unroll
            for (int u = 0; u < BATCH; ++u) Codes::fma16(q[u], rj[u], acc);
        }
        float *pw = part[wave];
        if (cg < CG) {
#pragma unroll
            for (int v = 0; v < 16; ++v) pw[cg * ldp + rg * 16 + v] = acc[v];
        }
        // the wave's own LDS writes, read back by other lanes of the same wave
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        for (int i = lane; i < n; i += 64) {
            float s = 0.f;
            for (int qq = 0; qq < CG; ++qq) s += pw[qq * ldp + i];
            z[off + (i / g) * nth + i % g] = s;
        }
        if constexpr (!LOOP) break;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        dcur = dnx;
        o8cur = o8nx;
        socur = sonx;
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
    std::vector<int32_t> dof0(p->m.ncell);
    RTX_HIP(hipMemcpy(dof0.data(), p->m.cell_dof_off, dof0.size() * 4, hipMemcpyDeviceToHost));
    const int split = (p->bj_method != 0) ? 1 : 0;
    std::vector<int32_t> blk, apd, cls[4];
    std::vector<int64_t> off(1, 0), off16(1, 0), off8(1, 0), soff(1, 0);
    int32_t maxb = 0;
    for (int k = 0; k < p->m.ncell; ++k) {
        const int nsp = nx[cc[k]] * ny[cc[k]], nth = nt[k];
        const int g = (split && nsp * nth > BJ_SPLIT) ? std::max(1, BJ_SPLIT / nsp) : nth;
        for (int a0 = 0; a0 < nth; a0 += g) {
            const int gg = std::min(g, nth - a0);
            const int64_t n = int64_t(nsp) * gg;
            blk.push_back(k);
            blk.push_back(a0);
            blk.push_back(gg);
            apd.push_back(int32_t(n));
            apd.push_back(dof0[k] + a0);
            apd.push_back(nth);
            apd.push_back(gg);
            const int q = n <= 32 ? 0 : n <= 64 ? 1 : n <= 96 ? 2 : 3;
            cls[q].push_back(int32_t(blk.size() / 3 - 1));
            off.push_back(off.back() + n * n);
            off16.push_back(off16.back() + n * ((n + 7) / 8 * 8));
            off8.push_back(off8.back() + n * ((n + 15) / 16 * 16));
            soff.push_back(soff.back() + n);
            maxb = std::max<int32_t>(maxb, int32_t(n));
        }
    }
    if (p->bj_split != split || !p->bj_off) {
        if (p->bj_off) RTX_HIP(hipFree(p->bj_off));
        if (p->bj_off16) RTX_HIP(hipFree(p->bj_off16));
        if (p->bj_off8) RTX_HIP(hipFree(p->bj_off8));
        if (p->bj_soff) RTX_HIP(hipFree(p->bj_soff));
        if (p->bj_blk) RTX_HIP(hipFree(p->bj_blk));
        if (p->bj_apd) RTX_HIP(hipFree(p->bj_apd));
        if (p->bj_cls) RTX_HIP(hipFree(p->bj_cls));
        p->bj_off = p->bj_off16 = p->bj_off8 = p->bj_soff = nullptr;
        p->bj_blk = p->bj_apd = p->bj_cls = nullptr;
        RTX_HIP(hipMalloc(&p->bj_off, off.size() * sizeof(int64_t)));
        RTX_HIP(hipMemcpy(p->bj_off, off.data(), off.size() * sizeof(int64_t),
                          hipMemcpyHostToDevice));
        RTX_HIP(hipMalloc(&p->bj_off16, off16.size() * sizeof(int64_t)));
        RTX_HIP(hipMemcpy(p->bj_off16, off16.data(), off16.size() * sizeof(int64_t),
                          hipMemcpyHostToDevice));
        RTX_HIP(hipMalloc(&p->bj_off8, off8.size() * sizeof(int64_t)));
        RTX_HIP(hipMemcpy(p->bj_off8, off8.data(), off8.size() * sizeof(int64_t),
                          hipMemcpyHostToDevice));
        RTX_HIP(hipMalloc(&p->bj_soff, soff.size() * sizeof(int64_t)));
        RTX_HIP(hipMemcpy(p->bj_soff, soff.data(), soff.size() * sizeof(int64_t),
                          hipMemcpyHostToDevice));
        RTX_HIP(hipMalloc(&p->bj_blk, blk.size() * sizeof(int32_t)));
        RTX_HIP(hipMemcpy(p->bj_blk, blk.data(), blk.size() * sizeof(int32_t),
                          hipMemcpyHostToDevice));
        RTX_HIP(hipMalloc(&p->bj_apd, apd.size() * sizeof(int32_t)));
        RTX_HIP(hipMemcpy(p->bj_apd, apd.data(), apd.size() * sizeof(int32_t),
                          hipMemcpyHostToDevice));
        std::vector<int32_t> all;
        for (int q = 0; q < 4; ++q) {
            p->bj_cls_start[q] = int32_t(all.size());
            all.insert(all.end(), cls[q].begin(), cls[q].end());
        }
        p->bj_cls_start[4] = int32_t(all.size());
        RTX_HIP(hipMalloc(&p->bj_cls, std::max<size_t>(all.size(), 1) * sizeof(int32_t)));
        if (!all.empty())
            RTX_HIP(hipMemcpy(p->bj_cls, all.data(), all.size() * sizeof(int32_t),
                              hipMemcpyHostToDevice));
        p->bj_split = split;
        p->bj_ready = false;
    }
    p->bj_nblk = int32_t(blk.size() / 3);
    p->bj_max_block = maxb;
    p->bj_entries = off.back();
    p->bj_entries16 = off16.back();
    p->bj_codes8 = (off8.back() + 15) / 16 * 16;
    p->bj_bytes8 = p->bj_codes8 + soff.back() * int64_t(sizeof(float));
    p->bj_codes4 = (off8.back() / 2 + 15) / 16 * 16;
    p->bj_bytes4 = p->bj_codes4 + soff.back() * int64_t(sizeof(float));
    *n_entries = off.back();
    return RTX_OK;
}

extern "C" int rtx_bjacobi_size_fp8(rtx_plan *p, int64_t *n_bytes) {
    RTX_CHECK(p && n_bytes, "rtx_bjacobi_size_fp8: null argument");
    if (!p->bj_off8) {
        int64_t n;
        const int rc = rtx_bjacobi_size(p, &n);
        if (rc) return rc;
    }
    *n_bytes = p->bj_bytes8;
    return RTX_OK;
}

extern "C" int rtx_bjacobi_pack_fp8(rtx_plan *p, const float *d_inv32, uint8_t *d_inv8,
                                    void *stream) {
    RTX_CHECK(p && d_inv32 && d_inv8, "rtx_bjacobi_pack_fp8: null argument");
    RTX_CHECK(p->bj_off8 && p->bj_nblk > 0, "rtx_bjacobi_pack_fp8: call rtx_bjacobi_size first");
    k_bj_pack_fp8<<<p->bj_nblk, BJ_THREADS, 0, static_cast<hipStream_t>(stream)>>>(
        p->m, p->bj_off, p->bj_off8, p->bj_soff, p->bj_blk, d_inv32, d_inv8,
        reinterpret_cast<float *>(d_inv8 + p->bj_codes8));
    RTX_HIP(hipGetLastError());
    return RTX_OK;
}

extern "C" int rtx_bjacobi_size_fp4(rtx_plan *p, int64_t *n_bytes) {
    RTX_CHECK(p && n_bytes, "rtx_bjacobi_size_fp4: null argument");
    if (!p->bj_off8) {
        int64_t n;
        const int rc = rtx_bjacobi_size(p, &n);
        if (rc) return rc;
    }
    *n_bytes = p->bj_bytes4;
    return RTX_OK;
}

extern "C" int rtx_bjacobi_pack_fp4(rtx_plan *p, const float *d_inv32, uint8_t *d_inv4,
                                    void *stream) {
    RTX_CHECK(p && d_inv32 && d_inv4, "rtx_bjacobi_pack_fp4: null argument");
    RTX_CHECK(p->bj_off8 && p->bj_nblk > 0, "rtx_bjacobi_pack_fp4: call rtx_bjacobi_size first");
    k_bj_pack_fp4<<<p->bj_nblk, BJ_THREADS, 0, static_cast<hipStream_t>(stream)>>>(
        p->m, p->bj_off, p->bj_off8, p->bj_soff, p->bj_blk, d_inv32, d_inv4,
        reinterpret_cast<float *>(d_inv4 + p->bj_codes4));
    RTX_HIP(hipGetLastError());
    return RTX_OK;
}

/* Entries (uint16) of the packed BF16 inverse pool. */
extern "C" int rtx_bjacobi_size_bf16(rtx_plan *p, int64_t *n_entries) {
    RTX_CHECK(p && n_entries, "rtx_bjacobi_size_bf16: null argument");
    if (!p->bj_off16) {
        int64_t n;
        const int rc = rtx_bjacobi_size(p, &n);
        if (rc) return rc;
    }
    *n_entries = p->bj_entries16;
    return RTX_OK;
}

/* Round the FP32 inverses to BF16 (nearest even) into the packed layout. */
extern "C" int rtx_bjacobi_pack_bf16(rtx_plan *p, const float *d_inv32, uint16_t *d_inv16,
                                     void *stream) {
    RTX_CHECK(p && d_inv32 && d_inv16, "rtx_bjacobi_pack_bf16: null argument");
    RTX_CHECK(p->bj_off16 && p->bj_nblk > 0, "rtx_bjacobi_pack_bf16: call rtx_bjacobi_size first");
    k_bj_pack_bf16<<<p->bj_nblk, BJ_THREADS, 0, static_cast<hipStream_t>(stream)>>>(
        p->m, p->bj_off, p->bj_off16, p->bj_blk, d_inv32, d_inv16);
    RTX_HIP(hipGetLastError());
    return RTX_OK;
}

extern "C" int rtx_bjacobi_factor(rtx_plan *p, int restrict_interior, double *d_inv,
                                  float *d_inv_f32, double *d_scratch, void *stream) {
    RTX_CHECK(p && (d_inv || d_inv_f32), "rtx_bjacobi_factor: null argument");
    if (!p->has_coeffs) {
        set_error("rtx_bjacobi_factor: coefficients not set");
        return RTX_ERR_STATE;
    }
    if (!p->bj_off || p->bj_split != ((p->bj_method != 0) ? 1 : 0)) {
        int64_t n;
        int rc = rtx_bjacobi_size(p, &n);
        if (rc) return rc;
    }
    hipStream_t st = static_cast<hipStream_t>(stream);
    const int64_t nmax = p->bj_max_block;
    const size_t lds = size_t(nmax) * nmax * sizeof(double);
    int dev = 0, max_lds = 0;
    RTX_HIP(hipGetDevice(&dev));
    RTX_HIP(hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
    const size_t static_lds = 2 * NMq * 8 + RTX_NMAX * RTX_NMAX * RTX_NMAX * 12 + 16;
    int *d_status = nullptr;
    RTX_HIP(hipMallocAsync((void **)&d_status, sizeof(int), st));
    RTX_HIP(hipMemsetAsync(d_status, 0, sizeof(int), st));
    const bool reg = p->bj_method != 0 && nmax <= 128;
    // no FP64 inverse wanted: the register build eliminates in FP32 straight
    // into d_inv_f32 (the FP32, BF16, FP8 and FP4 copies all start there)
    const bool f32_only = d_inv == nullptr;
    if (f32_only && !reg) {
        (void)hipFreeAsync(d_status, st);
        set_error("rtx_bjacobi_factor: an FP32-only factorisation needs the register build "
                  "(blocks of up to 128 DOFs, method -1 or 1); blocks of %lld DOFs",
                  (long long)nmax);
        return RTX_ERR_ARG;
    }
    if (reg) {
        // each size class on the smallest register grid that holds it:
        // fewer padded elements per elimination step, more blocks per CU
        const int32_t *cs = p->bj_cls, *c0 = p->bj_cls_start;
        auto cnt = [&](int q) { return c0[q + 1] - c0[q]; };
        auto run = [&](auto *out) {
            using T = typename std::remove_pointer<decltype(out)>::type;
            if (cnt(0))
                k_bj_build_reg<2, 2, 16, T><<<cnt(0), 256, 0, st>>>(
                    p->m, restrict_interior, p->bj_off, p->bj_blk, cs + c0[0], out, d_status);
            if (cnt(1))
                k_bj_build_reg<4, 4, 16, T><<<cnt(1), 256, 0, st>>>(
                    p->m, restrict_interior, p->bj_off, p->bj_blk, cs + c0[1], out, d_status);
            if (cnt(2))
                k_bj_build_reg<6, 4, 24, T><<<cnt(2), 384, 0, st>>>(
                    p->m, restrict_interior, p->bj_off, p->bj_blk, cs + c0[2], out, d_status);
            if (cnt(3))
                k_bj_build_reg<8, 4, 32, T><<<cnt(3), 512, 0, st>>>(
                    p->m, restrict_interior, p->bj_off, p->bj_blk, cs + c0[3], out, d_status);
        };
        if (f32_only)
            run(d_inv_f32);
        else
            run(d_inv);
    } else if (p->bj_method == 1) {
        (void)hipFreeAsync(d_status, st);
        set_error("rtx_bjacobi_factor: the register build handles blocks of up to 128 DOFs, "
                  "not %lld", (long long)nmax);
        return RTX_ERR_ARG;
    } else if (lds + static_lds <= size_t(max_lds)) {
        RTX_HIP(hipFuncSetAttribute((const void *)k_bj_build<true>,
                                    hipFuncAttributeMaxDynamicSharedMemorySize, int(lds)));
        k_bj_build<true><<<p->m.ncell, BJ_THREADS, lds, st>>>(p->m, restrict_interior, p->bj_off,
                                                            d_inv, nullptr, d_status);
    } else {
        if (d_scratch == nullptr) {
            (void)hipFreeAsync(d_status, st);
            set_error("rtx_bjacobi_factor: blocks of %lld DOFs do not fit LDS; a scratch buffer "
                      "of the same size as the inverses is required",
                      (long long)nmax);
            return RTX_ERR_ARG;
        }
        k_bj_build<false><<<p->m.ncell, BJ_THREADS, 0, st>>>(p->m, restrict_interior, p->bj_off,
                                                           d_inv, d_scratch, d_status);
    }
    RTX_HIP(hipGetLastError());
    if (d_inv_f32 && !f32_only)
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

extern "C" int rtx_bjacobi_set_method(rtx_plan *p, int method) {
    RTX_CHECK(p && method >= -1 && method <= 1, "rtx_bjacobi_set_method: method must be -1, 0 or 1");
    p->bj_method = method;
    return RTX_OK;
}

// The FP8 / FP4 applies: the wavefront-per-block kernel for blocks of at
// most BJ_W_MAXN rows, else the slab kernel; r of type RT, z FP32.
template <int BITS, typename RT>
static void launch_bj_coded(const rtx_plan *p, const void *d_inv, const RT *r, float *z,
                            hipStream_t st) {
    const uint8_t *codes = static_cast<const uint8_t *>(d_inv);
    const float *scales =
        reinterpret_cast<const float *>(codes + (BITS == 4 ? p->bj_codes4 : p->bj_codes8));
    const int4 *apd = reinterpret_cast<const int4 *>(p->bj_apd);
    const int nwg = (p->bj_nblk + BJ_W_WAVES - 1) / BJ_W_WAVES;
    if (p->bj_max_block <= BJ_W_MAXN && (BITS == 4 ? p->bj_bytes4 : p->bj_bytes8) <= (int64_t(256) << 20))
        k_bj_apply_fp8_w<RT, BITS, true><<<std::min(nwg, std::max(7 * p->n_cu, 1)),
                                           64 * BJ_W_WAVES, 0, st>>>(
            p->m, p->bj_off8, p->bj_soff, apd, p->bj_nblk, codes, scales, r, z);
    else if (p->bj_max_block <= BJ_W_MAXN)
        k_bj_apply_fp8_w<RT, BITS, false><<<nwg, 64 * BJ_W_WAVES, 0, st>>>(
            p->m, p->bj_off8, p->bj_soff, apd, p->bj_nblk, codes, scales, r, z);
    else
        k_bj_apply_fp8<BJ_APPLY8_THREADS, RT, BITS><<<p->bj_nblk, BJ_APPLY8_THREADS, 0, st>>>(
            p->m, p->bj_off8, p->bj_soff, apd, codes, scales, r, z);
}

extern "C" int rtx_bjacobi_apply(rtx_plan *p, int precision, const void *d_inv, const void *d_r,
                                 void *d_z, void *stream) {
    RTX_CHECK(p && d_inv && d_r && d_z, "rtx_bjacobi_apply: null argument");
    RTX_CHECK(d_r != d_z, "rtx_bjacobi_apply: r and z must not alias");
    if (!p->bj_ready) {
        set_error("rtx_bjacobi_apply: call rtx_bjacobi_factor first");
        return RTX_ERR_STATE;
    }
    const int4 *apd = reinterpret_cast<const int4 *>(p->bj_apd);
    hipStream_t st = static_cast<hipStream_t>(stream);
    if (precision == RTX_PREC_F64)
        k_bj_apply<double, double><<<p->bj_nblk, BJ_THREADS, 0, st>>>(
            p->m, p->bj_off, apd, (const double *)d_inv, (const double *)d_r, (double *)d_z);
    else if (precision == RTX_PREC_F32)
        k_bj_apply<float, float><<<p->bj_nblk, BJ_THREADS, 0, st>>>(
            p->m, p->bj_off, apd, (const float *)d_inv, (const float *)d_r, (float *)d_z);
    else if (precision == RTX_PREC_MIXED)
        k_bj_apply_bf16<BJ_APPLY16_THREADS><<<p->bj_nblk, BJ_APPLY16_THREADS, 0, st>>>(
            p->m, p->bj_off16, apd, (const uint16_t *)d_inv, (const float *)d_r,
            (float *)d_z);
    else if (precision == RTX_PREC_FP8)
        launch_bj_coded<8, float>(p, d_inv, (const float *)d_r, (float *)d_z, st);
    else if (precision == RTX_PREC_FP4)
        launch_bj_coded<4, float>(p, d_inv, (const float *)d_r, (float *)d_z, st);
    else {
        set_error("rtx_bjacobi_apply: unsupported precision %d", precision);
        return RTX_ERR_ARG;
    }
    RTX_HIP(hipGetLastError());
    return RTX_OK;
}

/* z = M r with the FP8 (RTX_PREC_FP8) or FP4 (RTX_PREC_FP4) inverses, r in
   FP16 (IEEE binary16) and z in FP32: the preconditioner of GMRES on an
   FP16 basis.  Same inverses and arithmetic as rtx_bjacobi_apply; only the
   input type differs. */
extern "C" int rtx_bjacobi_apply_h(rtx_plan *p, int precision, const void *d_inv, const void *d_r,
                                   void *d_z, void *stream) {
    RTX_CHECK(p && d_inv && d_r && d_z, "rtx_bjacobi_apply_h: null argument");
    RTX_CHECK(d_r != d_z, "rtx_bjacobi_apply_h: r and z must not alias");
    RTX_CHECK(precision == RTX_PREC_FP8 || precision == RTX_PREC_FP4,
              "rtx_bjacobi_apply_h: precision must be RTX_PREC_FP8 or RTX_PREC_FP4, not %d",
              precision);
    if (!p->bj_ready) {
        set_error("rtx_bjacobi_apply_h: call rtx_bjacobi_factor first");
        return RTX_ERR_STATE;
    }
    RTX_CHECK(p->bj_off8, "rtx_bjacobi_apply_h: call rtx_bjacobi_size first");
    hipStream_t st = static_cast<hipStream_t>(stream);
    if (precision == RTX_PREC_FP8)
        launch_bj_coded<8, _Float16>(p, d_inv, (const _Float16 *)d_r, (float *)d_z, st);
    else
        launch_bj_coded<4, _Float16>(p, d_inv, (const _Float16 *)d_r, (float *)d_z, st);
    RTX_HIP(hipGetLastError());
    return RTX_OK;
}

extern "C" int rtx_bjacobi_apply_fp8_h(rtx_plan *p, const void *d_inv, const void *d_r, void *d_z,
                                       void *stream) {
    return rtx_bjacobi_apply_h(p, RTX_PREC_FP8, d_inv, d_r, d_z, stream);
}
