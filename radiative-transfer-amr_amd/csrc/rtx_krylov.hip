// Krylov building blocks for the device-resident GMRES.
//
// The reference hands the assembled interior system to PETSc KSP
// (src/rt/solve.py:81-100: GMRES(50), which orthogonalises with classical
// Gram-Schmidt).  Here one Arnoldi step (classical Gram-Schmidt with a
// device-side "twice is enough" re-orthogonalisation test, then
// normalisation) is a fixed sequence of fused FP64 kernels over the tall
// basis V; reductions are two-phase and deterministic (per-block partial
// sums, then one reduction per output).
#include <mutex>
#include <type_traits>

#include "rtx_internal.h"

namespace rtx {

constexpr int KT = 256;          // threads per block
constexpr int KCH = 8;           // basis vectors per mdot block (gridDim.y chunks)
constexpr int KNB = 1024;        // row blocks of the tall-skinny kernels
constexpr int KMAXV = 256;       // largest basis handled by one Arnoldi step
constexpr double REORTH = 0.5;   // re-orthogonalise when |w'|^2 < 0.5 |w|^2
// state block after the partials (doubles): [0] |w|^2, [1] |w'|^2, [4] flag (int)
constexpr int WORK_TAIL = 16;

template <int N>
__device__ inline void block_sum(double (&v)[N], double *sh) {
    // sh: KT * N doubles; result in sh[j * KT].  (Cross-lane shuffles of
    // the doubles first, then one LDS stage, measured slower: 18.7 against
    // 15.6 us for the FP16-basis dot products on 2,048,000 rows.)
    const int tid = threadIdx.x;
#pragma unroll
    for (int j = 0; j < N; ++j) sh[j * KT + tid] = v[j];
    __syncthreads();
    for (int s = KT / 2; s > 0; s >>= 1) {
        if (tid < s) {
#pragma unroll
            for (int j = 0; j < N; ++j) sh[j * KT + tid] += sh[j * KT + tid + s];
        }
        __syncthreads();
    }
}

// Four consecutive values as doubles (16-byte loads; the caller checks the
// alignment).
__device__ inline void ld4(const float *p, double (&o)[4]) {
    const float4 f = *reinterpret_cast<const float4 *>(p);
    o[0] = f.x; o[1] = f.y; o[2] = f.z; o[3] = f.w;
}
__device__ inline void ld4(const double *p, double (&o)[4]) {
    const double2 a = *reinterpret_cast<const double2 *>(p);
    const double2 b = *reinterpret_cast<const double2 *>(p + 2);
    o[0] = a.x; o[1] = a.y; o[2] = b.x; o[3] = b.y;
}
// FP16 basis vectors: four consecutive values in one 8-byte load / store
typedef _Float16 kh4 __attribute__((ext_vector_type(4)));
__device__ inline void ld4(const _Float16 *p, double (&o)[4]) {
    const kh4 v = *reinterpret_cast<const kh4 *>(p);
    o[0] = double(v[0]); o[1] = double(v[1]); o[2] = double(v[2]); o[3] = double(v[3]);
}
__device__ inline void st4(_Float16 *p, const _Float16 (&v)[4]) {
    *reinterpret_cast<kh4 *>(p) = kh4{v[0], v[1], v[2], v[3]};
}
// a double rounded to a basis storage type (FP16 through FP32)
template <typename T>
__device__ inline T to_store(double x) {
    if constexpr (std::is_same<T, _Float16>::value) return _Float16(float(x));
    else return T(x);
}
__device__ inline void st4(float *p, const float (&v)[4]) {
    *reinterpret_cast<float4 *>(p) = make_float4(v[0], v[1], v[2], v[3]);
}
__device__ inline void st4(double *p, const double (&v)[4]) {
    *reinterpret_cast<double2 *>(p) = make_double2(v[0], v[1]);
    *reinterpret_cast<double2 *>(p + 2) = make_double2(v[2], v[3]);
}

// Row range of block b; with VEC the range starts on a multiple of 4.
template <bool VEC>
__device__ inline void row_range(int64_t n, int b, int nb, int64_t &i0, int64_t &i1) {
    int64_t per = (n + nb - 1) / nb;
    if (VEC) per = (per + 3) & ~int64_t(3);
    i0 = int64_t(b) * per;
    i1 = (i0 + per < n) ? i0 + per : n;
    if (i0 > n) i0 = n;
}

// partial[j * nb + b] = sum over block b's rows of V_j[i] w[i], j < nv;
// if with_ww, partial[nv * nb + b] = sum w[i]^2 (computed by chunk 0).
template <typename T, bool VEC, typename TW = T>
__global__ __launch_bounds__(KT) void k_mdot(const T *__restrict__ V, int64_t ld, int nv,
                                             const TW *__restrict__ w, int64_t n,
                                             double *__restrict__ partial, int with_ww,
                                             const int *__restrict__ skip) {
    if (skip && *skip) return;
    __shared__ double sh[KT * (KCH + 1)];
    const int b = blockIdx.x, nb = gridDim.x;
    const int j0 = blockIdx.y * KCH;
    int64_t i0, i1;
    row_range<VEC>(n, b, nb, i0, i1);
    double acc[KCH + 1];
#pragma unroll
    for (int j = 0; j <= KCH; ++j) acc[j] = 0.0;
    const int64_t nq = VEC ? (i1 - i0) / 4 : 0;
    for (int64_t q = threadIdx.x; q < nq; q += KT) {
        const int64_t i = i0 + 4 * q;
        double wv[4];
        ld4(w + i, wv);
#pragma unroll
        for (int j = 0; j < KCH; ++j)
            if (j0 + j < nv) {
                double vv[4];
                ld4(V + int64_t(j0 + j) * ld + i, vv);
                acc[j] += vv[0] * wv[0] + vv[1] * wv[1] + vv[2] * wv[2] + vv[3] * wv[3];
            }
        acc[KCH] += wv[0] * wv[0] + wv[1] * wv[1] + wv[2] * wv[2] + wv[3] * wv[3];
    }
    for (int64_t i = i0 + 4 * nq + threadIdx.x; i < i1; i += KT) {
        const double wi = double(w[i]);
#pragma unroll
        for (int j = 0; j < KCH; ++j)
            if (j0 + j < nv) acc[j] += double(V[int64_t(j0 + j) * ld + i]) * wi;
        acc[KCH] += wi * wi;
    }
    block_sum<KCH + 1>(acc, sh);
    if (threadIdx.x == 0) {
        for (int j = 0; j < KCH; ++j)
            if (j0 + j < nv) partial[int64_t(j0 + j) * nb + b] = sh[j * KT];
        if (with_ww && blockIdx.y == 0) partial[int64_t(nv) * nb + b] = sh[KCH * KT];
    }
}

// out[j] (+)= sum_b partial[j * nb + b]; one block per j.
__global__ __launch_bounds__(KT) void k_reduce(const double *__restrict__ partial, int nb,
                                               double *__restrict__ out, int accumulate,
                                               const int *__restrict__ skip) {
    if (skip && *skip) return;
    __shared__ double sh[KT];
    const int j = blockIdx.x;
    double v[1] = {0.0};
    for (int b = threadIdx.x; b < nb; b += KT) v[0] += partial[int64_t(j) * nb + b];
    block_sum<1>(v, sh);
    if (threadIdx.x == 0) out[j] = accumulate ? out[j] + sh[0] : sh[0];
}

// w -= V h (nv vectors); partial[b] = sum over block b of the new w^2
// (accumulated in FP64 for either storage type).
template <typename T, bool VEC>
__global__ __launch_bounds__(KT) void k_update(const T *__restrict__ V, int64_t ld, int nv,
                                               const double *__restrict__ h,
                                               T *__restrict__ w, int64_t n,
                                               double *__restrict__ partial,
                                               const int *__restrict__ skip) {
    if (skip && *skip) return;
    __shared__ double hs[KMAXV];
    __shared__ double sh[KT];
    for (int j = threadIdx.x; j < nv; j += KT) hs[j] = h[j];
    __syncthreads();
    const int b = blockIdx.x, nb = gridDim.x;
    int64_t i0, i1;
    row_range<VEC>(n, b, nb, i0, i1);
    double ss[1] = {0.0};
    const int64_t nq = VEC ? (i1 - i0) / 4 : 0;
    for (int64_t q = threadIdx.x; q < nq; q += KT) {
        const int64_t i = i0 + 4 * q;
        double acc[4];
        ld4(w + i, acc);
        for (int j = 0; j < nv; ++j) {
            double vv[4];
            ld4(V + int64_t(j) * ld + i, vv);
            const double hj = hs[j];
            acc[0] -= vv[0] * hj; acc[1] -= vv[1] * hj;
            acc[2] -= vv[2] * hj; acc[3] -= vv[3] * hj;
        }
        T out[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            out[u] = T(acc[u]);
            ss[0] += double(out[u]) * double(out[u]);
        }
        st4(w + i, out);
    }
    for (int64_t i = i0 + 4 * nq + threadIdx.x; i < i1; i += KT) {
        double acc = double(w[i]);
        int j = 0;
        for (; j + 4 <= nv; j += 4) {
            acc -= double(V[int64_t(j) * ld + i]) * hs[j] +
                   double(V[int64_t(j + 1) * ld + i]) * hs[j + 1] +
                   double(V[int64_t(j + 2) * ld + i]) * hs[j + 2] +
                   double(V[int64_t(j + 3) * ld + i]) * hs[j + 3];
        }
        for (; j < nv; ++j) acc -= double(V[int64_t(j) * ld + i]) * hs[j];
        const T wn = T(acc);
        w[i] = wn;
        ss[0] += double(wn) * double(wn);
    }
    block_sum<1>(ss, sh);
    if (threadIdx.x == 0) partial[b] = sh[0];
}

// One Arnoldi update for an FP32 basis with the norm taken from the dot
// products: hd[0..nv-1] = V^T w, hd[nv] = |w|^2 (FP64 sums of FP32 data),
// |w'|^2 = |w|^2 - sum h_j^2 (every block computes it, in the same order),
// v_out = (w - V h) / |w'| written directly (w itself is left as it was);
// a relative norm below 1e-6 counts as a breakdown (v_out = 0, h = 0).
// Block 0 writes hcol[0..nv] and, when state is given, folds the column
// into the Givens factorisation (k_normalize_givens's recurrence).
template <typename T, bool VEC, typename TW = T>
__global__ __launch_bounds__(KT) void k_update_pyth(const T *__restrict__ V, int64_t ld, int nv,
                                                    const double *__restrict__ hd,
                                                    const TW *__restrict__ w,
                                                    T *__restrict__ v_out, int64_t n,
                                                    double *__restrict__ hcol, int j, int m,
                                                    double *state) {
    __shared__ double hs[KMAXV + 1];
    __shared__ double beta_sh;
    for (int q = threadIdx.x; q <= nv; q += KT) hs[q] = hd[q];
    __syncthreads();
    if (threadIdx.x == 0) {
        double hh = 0.0;
        for (int q = 0; q < nv; ++q) hh += hs[q] * hs[q];
        const double r2 = hs[nv] - hh;
        beta_sh = (r2 > 1e-12 * hs[nv]) ? sqrt(r2) : 0.0;
    }
    __syncthreads();
    const double beta = beta_sh;
    const double sc = beta > 0.0 ? 1.0 / beta : 0.0;
    const int b = blockIdx.x, nb = gridDim.x;
    int64_t i0, i1;
    row_range<VEC>(n, b, nb, i0, i1);
    const int64_t nq = VEC ? (i1 - i0) / 4 : 0;
    for (int64_t q = threadIdx.x; q < nq; q += KT) {
        const int64_t i = i0 + 4 * q;
        double acc[4];
        ld4(w + i, acc);
        for (int jj = 0; jj < nv; ++jj) {
            double vv[4];
            ld4(V + int64_t(jj) * ld + i, vv);
            const double hj = hs[jj];
            acc[0] -= vv[0] * hj; acc[1] -= vv[1] * hj;
            acc[2] -= vv[2] * hj; acc[3] -= vv[3] * hj;
        }
        T out[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) out[u] = to_store<T>(acc[u] * sc);
        st4(v_out + i, out);
    }
    for (int64_t i = i0 + 4 * nq + threadIdx.x; i < i1; i += KT) {
        double acc = double(w[i]);
        for (int jj = 0; jj < nv; ++jj) acc -= double(V[int64_t(jj) * ld + i]) * hs[jj];
        v_out[i] = to_store<T>(acc * sc);
    }
    if (b != 0 || threadIdx.x != 0) return;
    for (int q = 0; q < nv; ++q) hcol[q] = hs[q];
    hcol[nv] = beta;
    if (!state) return;
    double *cs = state, *sn = state + m, *g = state + 2 * m;
    double *Rj = state + 3 * m + 1 + int64_t(j) * m;
    double *Hj = state + 3 * m + 1 + int64_t(m) * m + int64_t(j) * (m + 1);
    double *est = state + 3 * m + 1 + int64_t(m) * m + int64_t(m + 1) * m;
    for (int t = 0; t <= j; ++t) Hj[t] = hs[t];
    Hj[j + 1] = beta;
    double carry = hs[0];
    for (int t = 0; t < j; ++t) {
        const double bb = hs[t + 1];
        Rj[t] = cs[t] * carry + sn[t] * bb;
        carry = -sn[t] * carry + cs[t] * bb;
    }
    const double den = hypot(carry, beta);
    const double c = den != 0.0 ? carry / den : 1.0;
    const double sv = den != 0.0 ? beta / den : 0.0;
    cs[j] = c;
    sn[j] = sv;
    Rj[j] = den;
    const double gj = g[j];
    g[j + 1] = -sv * gj;
    g[j] = c * gj;
    est[j] = fabs(g[j + 1]);
}

// state layout (doubles): [0] |w|^2 before projection, [1] |w'|^2 after;
// flag (int) = 1 when the second pass is skipped.
__global__ void k_reorth_test(const double *state, int *flag) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *flag = (state[1] >= REORTH * state[0]) ? 1 : 0;
}

__global__ void k_set_flag(int *flag, int v) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *flag = v;
}

// v_out = w / sqrt(nrm2); hcol[nv] = sqrt(nrm2).  nrm2 comes from state.
template <typename T>
__global__ void k_normalize(const T *__restrict__ w, T *__restrict__ v_out, int64_t n,
                            const double *__restrict__ nrm2, double *__restrict__ hcol, int nv) {
    const double beta = sqrt(*nrm2);
    const double s = beta > 0.0 ? 1.0 / beta : 0.0;
    int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    for (; i < n; i += int64_t(gridDim.x) * blockDim.x) v_out[i] = T(double(w[i]) * s);
    if (blockIdx.x == 0 && threadIdx.x == 0) hcol[nv] = beta;
}

// x += sum_j y_j V_j  (y on the device)
template <typename TV, typename TX>
__global__ __launch_bounds__(KT) void k_combine(const TV *__restrict__ V, int64_t ld, int nv,
                                                const double *__restrict__ y,
                                                TX *__restrict__ x, int64_t n, int accumulate) {
    __shared__ double ys[KMAXV];
    for (int j = threadIdx.x; j < nv; j += KT) ys[j] = y[j];
    __syncthreads();
    int64_t i = int64_t(blockIdx.x) * KT + threadIdx.x;
    for (; i < n; i += int64_t(gridDim.x) * KT) {
        double acc = accumulate ? double(x[i]) : 0.0;
        for (int j = 0; j < nv; ++j) acc += double(V[int64_t(j) * ld + i]) * ys[j];
        x[i] = to_store<TX>(acc);
    }
}

// y = a x + b y with x and y of different storage types
template <typename TX, typename TY>
__global__ void k_axpby_mixed(int64_t n, double a, const TX *__restrict__ x, double b,
                              TY *__restrict__ y) {
    int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    for (; i < n; i += int64_t(gridDim.x) * blockDim.x)
        y[i] = to_store<TY>(a * double(x[i]) + (b == 0.0 ? 0.0 : b * double(y[i])));
}

// y = a x + b y (device scalars optional)
template <typename T>
__global__ void k_axpby(int64_t n, double a, const T *__restrict__ x, double b,
                        T *__restrict__ y) {
    int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    for (; i < n; i += int64_t(gridDim.x) * blockDim.x)
        y[i] = T(a * double(x[i]) + (b == 0.0 ? 0.0 : b * double(y[i])));
}

// ---- least-squares state of one GMRES restart cycle of m steps, on the
// device (doubles): cs[m], sn[m], g[m + 1], R[m * m] (column j at R + j m),
// H[(m + 1) * m] (raw Hessenberg, column j at H + j (m + 1)), est[m].
struct GState {
    double *cs, *sn, *g, *R, *H, *est;
    __device__ GState(double *s, int m)
        : cs(s), sn(s + m), g(s + 2 * m), R(s + 3 * m + 1), H(s + 3 * m + 1 + m * m),
          est(s + 3 * m + 1 + m * m + (m + 1) * m) {}
};

__global__ void k_gmres_begin(int m, double beta, double *state) {
    GState s(state, m);
    for (int i = threadIdx.x; i <= m; i += blockDim.x) s.g[i] = (i == 0) ? beta : 0.0;
}

// Fold Hessenberg column j (hcol[0..j+1]) into the QR factorisation: apply
// the previous rotations, form rotation j, update g; est[j] = |g[j + 1]|
// is the residual norm of the minimiser over the first j + 1 directions.
__global__ void k_gmres_givens(int j, int m, const double *__restrict__ hcol, double *state) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    GState s(state, m);
    double *Rj = s.R + int64_t(j) * m;
    double *Hj = s.H + int64_t(j) * (m + 1);
    for (int i = 0; i <= j + 1; ++i) Hj[i] = hcol[i];
    for (int i = 0; i <= j; ++i) Rj[i] = hcol[i];
    const double hj1 = hcol[j + 1];
    for (int i = 0; i < j; ++i) {
        const double a = Rj[i], b = Rj[i + 1];
        Rj[i] = s.cs[i] * a + s.sn[i] * b;
        Rj[i + 1] = -s.sn[i] * a + s.cs[i] * b;
    }
    const double den = hypot(Rj[j], hj1);
    const double c = den != 0.0 ? Rj[j] / den : 1.0;
    const double sn = den != 0.0 ? hj1 / den : 0.0;
    s.cs[j] = c;
    s.sn[j] = sn;
    Rj[j] = den;
    const double gj = s.g[j];
    s.g[j + 1] = -sn * gj;
    s.g[j] = c * gj;
    s.est[j] = fabs(s.g[j + 1]);
}

// v_out = w / |w| (as k_normalize) and, in block 0, the Givens update of
// column j (as k_gmres_givens) with the rotations staged in LDS, so the
// serial recurrence runs on LDS values while the other blocks normalise.
template <typename T>
__global__ __launch_bounds__(KT) void k_normalize_givens(const T *__restrict__ w,
                                                         T *__restrict__ v_out, int64_t n,
                                                         const double *__restrict__ nrm2,
                                                         double *__restrict__ hcol, int nv,
                                                         int j, int m, double *state) {
    const double beta = sqrt(*nrm2);
    const double s = beta > 0.0 ? 1.0 / beta : 0.0;
    int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    for (; i < n; i += int64_t(gridDim.x) * blockDim.x) v_out[i] = T(double(w[i]) * s);
    if (blockIdx.x != 0) return;
    __shared__ double hc[KMAXV + 1], cs[KMAXV], sn[KMAXV];
    GState g(state, m);
    const int tid = threadIdx.x;
    for (int t = tid; t <= j; t += KT) hc[t] = hcol[t];
    if (tid == 0) hc[j + 1] = beta;
    for (int t = tid; t < j; t += KT) {
        cs[t] = g.cs[t];
        sn[t] = g.sn[t];
    }
    __syncthreads();
    double *Hj = g.H + int64_t(j) * (m + 1);
    for (int t = tid; t <= j + 1; t += KT) Hj[t] = hc[t];
    __syncthreads();                 // hc is rotated in place below
    if (tid == 0) {
        hcol[nv] = beta;
        double carry = hc[0];
        for (int t = 0; t < j; ++t) {
            const double b = hc[t + 1];
            hc[t] = cs[t] * carry + sn[t] * b;
            carry = -sn[t] * carry + cs[t] * b;
        }
        const double hj1 = beta;
        const double den = hypot(carry, hj1);
        const double c = den != 0.0 ? carry / den : 1.0;
        const double sv = den != 0.0 ? hj1 / den : 0.0;
        g.cs[j] = c;
        g.sn[j] = sv;
        hc[j] = den;
        const double gj = g.g[j];
        g.g[j + 1] = -sv * gj;
        g.g[j] = c * gj;
        g.est[j] = fabs(g.g[j + 1]);
    }
    __syncthreads();
    double *Rj = g.R + int64_t(j) * m;
    for (int t = tid; t <= j; t += KT) Rj[t] = hc[t];
}

// y = R_k^{-1} g_k (back substitution over the first k columns).
__global__ void k_gmres_solve(int k, int m, const double *state, double *__restrict__ y) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    GState s(const_cast<double *>(state), m);
    for (int i = k - 1; i >= 0; --i) {
        double acc = s.g[i];
        for (int l = i + 1; l < k; ++l) acc -= s.R[int64_t(l) * m + i] * y[l];
        const double d = s.R[int64_t(i) * m + i];
        y[i] = d != 0.0 ? acc / d : 0.0;
    }
}

template <typename S, typename D>
__global__ void k_convert(int64_t n, const S *__restrict__ s, D *__restrict__ d) {
    int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    for (; i < n; i += int64_t(gridDim.x) * blockDim.x) d[i] = D(s[i]);
}

}  // namespace rtx

using namespace rtx;

extern "C" int rtx_arnoldi_workspace(int max_vecs, int64_t *n_doubles) {
    RTX_CHECK(n_doubles && max_vecs >= 1 && max_vecs < KMAXV,
              "rtx_arnoldi_workspace: max_vecs must be in [1, %d)", KMAXV);
    // dot partials ((KMAXV + 1) rows, the last one also holding the second-pass
    // coefficients), norm partials, and a small state block (norms + flag)
    *n_doubles = int64_t(KMAXV + 2) * KNB + WORK_TAIL;
    return RTX_OK;
}

/* One Arnoldi step: orthogonalise w against V_0..V_{nv-1} (columns of length n,
   stride ld) by classical Gram-Schmidt; refine 0: one pass, 1: a second pass
   when the norm dropped by more than the "twice is enough" factor, 2: always
   two passes; normalise into V_nv and write the Hessenberg column
   h_0..h_{nv-1}, h_nv = |w| to d_hcol (device, nv + 1 doubles).  w is
   overwritten.  Everything is enqueued on `stream`; no host sync. */
template <typename T>
static bool aligned16(const T *p) {
    return (reinterpret_cast<uintptr_t>(p) & 15) == 0;
}

template <typename T>
static void launch_mdot(bool vec, dim3 grid, hipStream_t st, const T *V, int64_t ld, int nv,
                        const T *w, int64_t n, double *part, int with_ww, const int *flag) {
    if (vec)
        k_mdot<T, true><<<grid, KT, 0, st>>>(V, ld, nv, w, n, part, with_ww, flag);
    else
        k_mdot<T, false><<<grid, KT, 0, st>>>(V, ld, nv, w, n, part, with_ww, flag);
}

template <typename T>
static void launch_update(bool vec, hipStream_t st, const T *V, int64_t ld, int nv,
                          const double *h, T *w, int64_t n, double *part, const int *flag) {
    if (vec)
        k_update<T, true><<<KNB, KT, 0, st>>>(V, ld, nv, h, w, n, part, flag);
    else
        k_update<T, false><<<KNB, KT, 0, st>>>(V, ld, nv, h, w, n, part, flag);
}

// One Arnoldi step; with d_state, the final normalisation also folds the
// new Hessenberg column j into the least-squares state (rtx_gmres_givens).
template <typename T>
static int arnoldi_step(const T *d_V, int64_t ld, int nv, int64_t n, T *d_w, double *d_hcol,
                        double *d_work, int refine, void *stream, int gj = -1, int gm = 0,
                        double *d_state = nullptr) {
    RTX_CHECK(d_V && d_w && d_hcol && d_work, "rtx_arnoldi_step: null argument");
    RTX_CHECK(refine >= 0 && refine <= 2, "rtx_arnoldi_step: refine must be 0, 1 or 2");
    RTX_CHECK(nv >= 1 && nv < KMAXV - 1, "rtx_arnoldi_step: nv=%d outside [1, %d)", nv,
              KMAXV - 1);
    RTX_CHECK(n >= 1 && ld >= n, "rtx_arnoldi_step: bad n/ld");
    hipStream_t st = static_cast<hipStream_t>(stream);
    double *part = d_work;                                  // (nv + 1) * KNB
    double *part2 = d_work + int64_t(KMAXV + 1) * KNB;      // KNB
    double *state = part2 + KNB;                            // [ww0, ww1, h2...]
    int *flag = reinterpret_cast<int *>(state + 4);
    const int nchunk = (nv + KCH - 1) / KCH;
    T *V_out = const_cast<T *>(d_V) + int64_t(nv) * ld;
    const bool vec = (ld % 4 == 0) && aligned16(d_V) && aligned16(d_w);
    if (refine == 0 && std::is_same<T, float>::value) {
        // FP32 basis (the inner solves of mixed refinement): classical
        // Gram-Schmidt, one pass, with the norm from the dot products, in
        // three launches: dots, their reduction, the update writing the
        // normalised vector (and the Givens update)
        double *hd = part2;                     // nv + 1 reduced values
        launch_mdot<T>(vec, dim3(KNB, nchunk), st, d_V, ld, nv, d_w, n, part, 1, nullptr);
        k_reduce<<<nv + 1, KT, 0, st>>>(part, KNB, hd, 0, nullptr);
        if (vec)
            k_update_pyth<T, true><<<KNB, KT, 0, st>>>(d_V, ld, nv, hd, d_w, V_out, n, d_hcol,
                                                     gj, gm, d_state);
        else
            k_update_pyth<T, false><<<KNB, KT, 0, st>>>(d_V, ld, nv, hd, d_w, V_out, n, d_hcol,
                                                      gj, gm, d_state);
        RTX_HIP(hipGetLastError());
        return RTX_OK;
    }
    auto normalize = [&]() {
        if (d_state)
            k_normalize_givens<T><<<2048, KT, 0, st>>>(d_w, V_out, n, state + 1, d_hcol, nv, gj,
                                                       gm, d_state);
        else
            k_normalize<T><<<2048, KT, 0, st>>>(d_w, V_out, n, state + 1, d_hcol, nv);
    };
    // pass 1: h = V^T w, |w|^2
    launch_mdot<T>(vec, dim3(KNB, nchunk), st, d_V, ld, nv, d_w, n, part, refine == 1 ? 1 : 0,
                   nullptr);
    k_reduce<<<nv, KT, 0, st>>>(part, KNB, d_hcol, 0, nullptr);
    if (refine == 1) k_reduce<<<1, KT, 0, st>>>(part + int64_t(nv) * KNB, KNB, state + 0, 0, nullptr);
    // w -= V h, |w'|^2
    launch_update<T>(vec, st, d_V, ld, nv, d_hcol, d_w, n, part2, nullptr);
    k_reduce<<<1, KT, 0, st>>>(part2, KNB, state + 1, 0, nullptr);
    if (refine == 0) {     // classical Gram-Schmidt, one pass (PETSc's default)
        normalize();
        RTX_HIP(hipGetLastError());
        return RTX_OK;
    }
    if (refine == 1)
        k_reorth_test<<<1, 1, 0, st>>>(state, flag);
    else
        k_set_flag<<<1, 1, 0, st>>>(flag, 0);
    // optional pass 2 (skipped on the device when flag == 1)
    launch_mdot<T>(vec, dim3(KNB, nchunk), st, d_V, ld, nv, d_w, n, part, 0, flag);
    // accumulate the correction into hcol, keep it separately for the update
    double *hcorr = part + int64_t(KMAXV) * KNB;            // nv doubles, after the partials
    k_reduce<<<nv, KT, 0, st>>>(part, KNB, hcorr, 0, flag);
    launch_update<T>(vec, st, d_V, ld, nv, hcorr, d_w, n, part2, flag);
    k_reduce<<<1, KT, 0, st>>>(part2, KNB, state + 1, 0, flag);
    k_reduce<<<nv, KT, 0, st>>>(hcorr, 1, d_hcol, 1, flag);  // hcol += hcorr
    normalize();
    RTX_HIP(hipGetLastError());
    return RTX_OK;
}

// Arnoldi step on an FP16 basis with an FP32 w (the inner solves of mixed
// refinement): classical Gram-Schmidt, one pass, the norm from the dot
// products (accumulated in FP64), in the FP32 basis's three launches; the
// new basis vector is rounded to FP16 once, after the normalisation.
static int arnoldi_step_h(const _Float16 *d_V, int64_t ld, int nv, int64_t n, const float *d_w,
                          double *d_hcol, double *d_work, int refine, void *stream, int gj,
                          int gm, double *d_state) {
    RTX_CHECK(d_V && d_w && d_hcol && d_work, "rtx_gmres_arnoldi: null argument");
    RTX_CHECK(refine == 0, "rtx_gmres_arnoldi: an FP16 basis takes one pass (refine 0)");
    RTX_CHECK(nv >= 1 && nv < KMAXV - 1, "rtx_gmres_arnoldi: nv=%d outside [1, %d)", nv,
              KMAXV - 1);
    RTX_CHECK(n >= 1 && ld >= n, "rtx_gmres_arnoldi: bad n/ld");
    hipStream_t st = static_cast<hipStream_t>(stream);
    double *part = d_work;
    double *hd = d_work + int64_t(KMAXV + 1) * KNB;
    const int nchunk = (nv + KCH - 1) / KCH;
    _Float16 *V_out = const_cast<_Float16 *>(d_V) + int64_t(nv) * ld;
    const bool vec = (ld % 4 == 0) && aligned16(d_V) && aligned16(d_w) &&
                     (reinterpret_cast<uintptr_t>(V_out) & 7) == 0;
    if (vec) {
        k_mdot<_Float16, true, float><<<dim3(KNB, nchunk), KT, 0, st>>>(d_V, ld, nv, d_w, n, part,
                                                                        1, nullptr);
        k_reduce<<<nv + 1, KT, 0, st>>>(part, KNB, hd, 0, nullptr);
        k_update_pyth<_Float16, true, float><<<KNB, KT, 0, st>>>(d_V, ld, nv, hd, d_w, V_out, n,
                                                                d_hcol, gj, gm, d_state);
    } else {
        k_mdot<_Float16, false, float><<<dim3(KNB, nchunk), KT, 0, st>>>(d_V, ld, nv, d_w, n,
                                                                         part, 1, nullptr);
        k_reduce<<<nv + 1, KT, 0, st>>>(part, KNB, hd, 0, nullptr);
        k_update_pyth<_Float16, false, float><<<KNB, KT, 0, st>>>(d_V, ld, nv, hd, d_w, V_out, n,
                                                                 d_hcol, gj, gm, d_state);
    }
    RTX_HIP(hipGetLastError());
    return RTX_OK;
}

/* One Arnoldi step: orthogonalise w against V_0..V_{nv-1} (columns of length n,
   stride ld) by classical Gram-Schmidt; refine 0: one pass, 1: a second pass
   when the norm dropped by more than the "twice is enough" factor, 2: always
   two passes; normalise into V_nv and write the Hessenberg column
   h_0..h_{nv-1}, h_nv = |w| to d_hcol (device, nv + 1 doubles).  w is
   overwritten.  Everything is enqueued on `stream`; no host sync. */
extern "C" int rtx_arnoldi_step(const double *d_V, int64_t ld, int nv, int64_t n, double *d_w,
                                double *d_hcol, double *d_work, int refine, void *stream) {
    return arnoldi_step<double>(d_V, ld, nv, n, d_w, d_hcol, d_work, refine, stream);
}

/* Arnoldi step j of a GMRES cycle of m steps fused with its Givens update:
   rtx_arnoldi_step (FP64 basis, basis_precision 0, FP32, 1, or FP16 with an
   FP32 w, 2, one pass only) on nv = j + 1 vectors, then
   rtx_gmres_givens(j, m, d_hcol, d_state) folded into the final
   normalisation.  No counterpart in the reference (PETSc
   keeps the least-squares problem on the host). */
extern "C" int rtx_gmres_arnoldi(int basis_precision, const void *d_V, int64_t ld, int64_t n,
                                 void *d_w, double *d_hcol, double *d_work, int refine, int j,
                                 int m, double *d_state, void *stream) {
    RTX_CHECK(d_state && m >= 1 && m < KMAXV - 1 && j >= 0 && j < m,
              "rtx_gmres_arnoldi: bad argument (j=%d, m=%d)", j, m);
    RTX_CHECK(basis_precision >= 0 && basis_precision <= 2,
              "rtx_gmres_arnoldi: basis_precision must be 0 (FP64), 1 (FP32) or 2 (FP16)");
    if (basis_precision == 2)
        return arnoldi_step_h(static_cast<const _Float16 *>(d_V), ld, j + 1, n,
                              static_cast<const float *>(d_w), d_hcol, d_work, refine, stream, j,
                              m, d_state);
    if (basis_precision == 0)
        return arnoldi_step<double>(static_cast<const double *>(d_V), ld, j + 1, n,
                                    static_cast<double *>(d_w), d_hcol, d_work, refine, stream,
                                    j, m, d_state);
    return arnoldi_step<float>(static_cast<const float *>(d_V), ld, j + 1, n,
                               static_cast<float *>(d_w), d_hcol, d_work, refine, stream, j, m,
                               d_state);
}

/* ---- rtx_gmres_arnoldi without a workspace argument.  Its reduction
   partials, norms and re-orthogonalisation flag live in a file-scope device
   array instead, so a caller that cannot allocate (code being captured into
   a HIP graph, a serving loop) passes only its inputs and outputs.  The
   array has the workspace layout of rtx_arnoldi_workspace (every step
   writes each partial before it reads it, so nothing needs clearing
   between calls).  One copy per device, held until the process ends;
   calls that overlap on one device share it and race. */
constexpr int64_t STATIC_WORK_DOUBLES = int64_t(KMAXV + 2) * KNB + WORK_TAIL;   // 2.1 MB
static __device__ __align__(256) double scratch[STATIC_WORK_DOUBLES];

constexpr int MAX_DEVICES = 64;
static std::mutex scratch_mutex;
static double *scratch_ptr[MAX_DEVICES];     // device address of `scratch`, per device

// Device address of `scratch` on the current device (looked up once per
// device with that device current); null with the error set on failure.
static double *static_workspace() {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= MAX_DEVICES) {
        set_error("rtx_gmres_arnoldi_static: no usable current device (%d)", dev);
        return nullptr;
    }
    std::lock_guard<std::mutex> lock(scratch_mutex);
    if (!scratch_ptr[dev]) {
        void *ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr) {
            set_error("rtx_gmres_arnoldi_static: hipGetSymbolAddress failed on device %d: %s",
                      dev, hipGetErrorString(e));
            return nullptr;
        }
        scratch_ptr[dev] = static_cast<double *>(ptr);
    }
    return scratch_ptr[dev];
}

extern "C" int rtx_gmres_arnoldi_static(int basis_precision, const void *d_V, int64_t ld,
                                        int64_t n, void *d_w, double *d_hcol, int refine,
                                        int j, int m, double *d_state, void *stream) {
    RTX_CHECK(m >= 1 && m < KMAXV - 1 && j >= 0 && j < m,
              "rtx_gmres_arnoldi_static: cycle length m=%d (step j=%d) outside the static "
              "workspace's limit m < %d",
              m, j, KMAXV - 1);
    double *work = static_workspace();
    if (!work) return RTX_ERR_HIP;
    return rtx_gmres_arnoldi(basis_precision, d_V, ld, n, d_w, d_hcol, work, refine, j, m,
                             d_state, stream);
}

/* The same step on an FP32 basis (dot products, norms and the Hessenberg
   column accumulated in FP64). */
extern "C" int rtx_arnoldi_step_f32(const float *d_V, int64_t ld, int nv, int64_t n, float *d_w,
                                    double *d_hcol, double *d_work, int refine, void *stream) {
    return arnoldi_step<float>(d_V, ld, nv, n, d_w, d_hcol, d_work, refine, stream);
}

extern "C" int rtx_gmres_state_size(int m, int64_t *n_doubles) {
    RTX_CHECK(n_doubles && m >= 1 && m < KMAXV - 1, "rtx_gmres_state_size: m outside [1, %d)",
              KMAXV - 1);
    *n_doubles = 3 * int64_t(m) + 1 + int64_t(m) * m + int64_t(m + 1) * m + m;
    return RTX_OK;
}

extern "C" int rtx_gmres_begin(int m, double beta, double *d_state, void *stream) {
    RTX_CHECK(d_state && m >= 1 && m < KMAXV - 1, "rtx_gmres_begin: bad argument");
    k_gmres_begin<<<1, 256, 0, static_cast<hipStream_t>(stream)>>>(m, beta, d_state);
    RTX_HIP(hipGetLastError());
    return RTX_OK;
}

extern "C" int rtx_gmres_givens(int j, int m, const double *d_hcol, double *d_state,
                                void *stream) {
    RTX_CHECK(d_hcol && d_state && m >= 1 && m < KMAXV - 1 && j >= 0 && j < m,
              "rtx_gmres_givens: bad argument (j=%d, m=%d)", j, m);
    k_gmres_givens<<<1, 64, 0, static_cast<hipStream_t>(stream)>>>(j, m, d_hcol, d_state);
    RTX_HIP(hipGetLastError());
    return RTX_OK;
}

extern "C" int rtx_gmres_solve(int k, int m, const double *d_state, double *d_y, void *stream) {
    RTX_CHECK(d_state && d_y && m >= 1 && m < KMAXV - 1 && k >= 1 && k <= m,
              "rtx_gmres_solve: bad argument (k=%d, m=%d)", k, m);
    k_gmres_solve<<<1, 64, 0, static_cast<hipStream_t>(stream)>>>(k, m, d_state, d_y);
    RTX_HIP(hipGetLastError());
    return RTX_OK;
}

/* d_out[j] = V_j . w for j < nv (device), workspace as for rtx_arnoldi_step. */
extern "C" int rtx_mdot(const double *d_V, int64_t ld, int nv, int64_t n, const double *d_w,
                        double *d_out, double *d_work, void *stream) {
    RTX_CHECK(d_V && d_w && d_out && d_work && nv >= 1 && nv < KMAXV, "rtx_mdot: bad argument");
    hipStream_t st = static_cast<hipStream_t>(stream);
    const int nchunk = (nv + KCH - 1) / KCH;
    // vector loads need every V_j on 16 bytes: any ld for a single vector
    const bool vec = (nv == 1 || ld % 4 == 0) && aligned16(d_V) && aligned16(d_w);
    launch_mdot<double>(vec, dim3(KNB, nchunk), st, d_V, ld, nv, d_w, n, d_work, 0, nullptr);
    k_reduce<<<nv, KT, 0, st>>>(d_work, KNB, d_out, 0, nullptr);
    RTX_HIP(hipGetLastError());
    return RTX_OK;
}

/* d_x (+)= sum_j y_j V_j with y on the device. */
extern "C" int rtx_combine(const double *d_V, int64_t ld, int nv, const double *d_y, int64_t n,
                           double *d_x, int accumulate, void *stream) {
    RTX_CHECK(d_V && d_y && d_x && nv >= 1 && nv < KMAXV, "rtx_combine: bad argument");
    k_combine<double, double><<<2048, KT, 0, static_cast<hipStream_t>(stream)>>>(
        d_V, ld, nv, d_y, d_x, n, accumulate);
    RTX_HIP(hipGetLastError());
    return RTX_OK;
}

/* d_x (+)= sum_j y_j V_j for an FP32 basis; d_x is FP64 (x_precision 0) or
   FP32 (x_precision 1); accumulation in FP64. */
extern "C" int rtx_combine_f32(const float *d_V, int64_t ld, int nv, const double *d_y,
                               int64_t n, void *d_x, int x_precision, int accumulate,
                               void *stream) {
    RTX_CHECK(d_V && d_y && d_x && nv >= 1 && nv < KMAXV && (x_precision == 0 || x_precision == 1),
              "rtx_combine_f32: bad argument");
    hipStream_t st = static_cast<hipStream_t>(stream);
    if (x_precision == 0)
        k_combine<float, double><<<2048, KT, 0, st>>>(d_V, ld, nv, d_y, (double *)d_x, n,
                                                      accumulate);
    else
        k_combine<float, float><<<2048, KT, 0, st>>>(d_V, ld, nv, d_y, (float *)d_x, n,
                                                     accumulate);
    RTX_HIP(hipGetLastError());
    return RTX_OK;
}

/* d_x (+)= sum_j y_j V_j for an FP16 basis (device, IEEE binary16); d_x is
   FP64 (x_precision 0) or FP32 (x_precision 1); accumulation in FP64. */
extern "C" int rtx_combine_f16(const void *d_V, int64_t ld, int nv, const double *d_y,
                               int64_t n, void *d_x, int x_precision, int accumulate,
                               void *stream) {
    RTX_CHECK(d_V && d_y && d_x && nv >= 1 && nv < KMAXV && (x_precision == 0 || x_precision == 1),
              "rtx_combine_f16: bad argument");
    hipStream_t st = static_cast<hipStream_t>(stream);
    const _Float16 *V = static_cast<const _Float16 *>(d_V);
    if (x_precision == 0)
        k_combine<_Float16, double><<<2048, KT, 0, st>>>(V, ld, nv, d_y, (double *)d_x, n,
                                                         accumulate);
    else
        k_combine<_Float16, float><<<2048, KT, 0, st>>>(V, ld, nv, d_y, (float *)d_x, n,
                                                        accumulate);
    RTX_HIP(hipGetLastError());
    return RTX_OK;
}

/* y = a x + b y with x FP32 and y FP64 (direction 0), x FP64 and y FP32
   (direction 1), or x FP64 and y FP16 (direction 2). */
extern "C" int rtx_axpby_mixed(int direction, int64_t n, double a, const void *d_x, double b,
                               void *d_y, void *stream) {
    RTX_CHECK(d_x && d_y && n >= 0 && direction >= 0 && direction <= 2,
              "rtx_axpby_mixed: bad argument");
    hipStream_t st = static_cast<hipStream_t>(stream);
    if (direction == 0)
        k_axpby_mixed<float, double><<<2048, 256, 0, st>>>(n, a, (const float *)d_x, b,
                                                           (double *)d_y);
    else if (direction == 2)
        k_axpby_mixed<double, _Float16><<<2048, 256, 0, st>>>(n, a, (const double *)d_x, b,
                                                              (_Float16 *)d_y);
    else
        k_axpby_mixed<double, float><<<2048, 256, 0, st>>>(n, a, (const double *)d_x, b,
                                                           (float *)d_y);
    RTX_HIP(hipGetLastError());
    return RTX_OK;
}

/* y = a x + b y for FP64 (precision 0) or FP32 (precision 1) vectors. */
extern "C" int rtx_axpby(int precision, int64_t n, double a, const void *d_x, double b,
                         void *d_y, void *stream) {
    RTX_CHECK(d_x && d_y && n >= 0, "rtx_axpby: bad argument");
    hipStream_t st = static_cast<hipStream_t>(stream);
    if (precision == RTX_PREC_F64)
        k_axpby<double><<<2048, 256, 0, st>>>(n, a, (const double *)d_x, b, (double *)d_y);
    else
        k_axpby<float><<<2048, 256, 0, st>>>(n, a, (const float *)d_x, b, (float *)d_y);
    RTX_HIP(hipGetLastError());
    return RTX_OK;
}

/* Precision conversion FP64 <-> FP32 (direction 0: f64 -> f32, 1: f32 -> f64). */
extern "C" int rtx_convert(int direction, int64_t n, const void *d_src, void *d_dst,
                           void *stream) {
    RTX_CHECK(d_src && d_dst && n >= 0, "rtx_convert: bad argument");
    hipStream_t st = static_cast<hipStream_t>(stream);
    if (direction == 0)
        k_convert<double, float><<<2048, 256, 0, st>>>(n, (const double *)d_src, (float *)d_dst);
    else
        k_convert<float, double><<<2048, 256, 0, st>>>(n, (const float *)d_src, (double *)d_dst);
    RTX_HIP(hipGetLastError());
    return RTX_OK;
}
