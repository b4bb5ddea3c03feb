// Learned coarse scalar-flux correction: coarse-grid transfer, MLP training
// (FP32 MFMA GEMMs, Adam) and BF16 inference kernels.
//
// The reference solve has no counterpart (its preconditioners are PETSc
// bjacobi / kaczmarz, src/rt/solve.py:86-95); this is the two-level part of
// the build's solver.  A column-wise restriction turns a residual into three
// angular moments per coarse cell (int r, int r cos th, int r sin th over the
// column), a small MLP maps the normalised moments (plus coefficient
// channels) to a coarse isotropic correction, and a prolongation writes that
// correction back to every interior DOF of each column.
#include "rtx_internal.h"

namespace rtx {

namespace {

constexpr int GT = 256;            // GEMM threads (4 waves as 2 x 2)
constexpr int GM = 64, GN = 64;    // GEMM tile
constexpr int GK = 16;             // K per LDS stage
constexpr int SPLIT_MAX = 64;

typedef float m_flt4 __attribute__((ext_vector_type(4)));

enum { EPI_NONE = 0, EPI_BIAS = 1, EPI_MASK = 2 };

struct GemmArgs {
    int M, N, K;
    const float *A;
    int64_t sam, sak;
    const float *B;
    int64_t sbk, sbn;
    float *C;
    int64_t ldc;
    int epi;
    const float *bias;
    float *H;          // EPI_BIAS: relu(C) copy (ld = ldc), optional
    const float *Z;    // EPI_MASK: C *= (Z > 0)
    int64_t ldz;
    float *part;       // split-K partials [splits][M][N], or null
    int kchunk;        // K range per split
};

__device__ inline float epilogue(const GemmArgs &g, int m, int n, float v) {
    if (g.epi == EPI_BIAS) v += g.bias[n];
    else if (g.epi == EPI_MASK) v = (g.Z[int64_t(m) * g.ldz + n] > 0.f) ? v : 0.f;
    return v;
}

__device__ inline void store_out(const GemmArgs &g, int m, int n, float v) {
    v = epilogue(g, m, n, v);
    g.C[int64_t(m) * g.ldc + n] = v;
    if (g.epi == EPI_BIAS && g.H) g.H[int64_t(m) * g.ldc + n] = v > 0.f ? v : 0.f;
}

// C[m, n] = sum_k A[m, k] B[k, n] over the split's K range, A and B read
// through arbitrary strides (covers X W^T, dZ^T H and dZ W without copies).
__global__ __launch_bounds__(GT) void k_gemm(GemmArgs g) {
    __shared__ float As[GK][GM + 4];
    __shared__ float Bs[GK][GN + 4];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave >> 1, wn = wave & 1;
    const int m0 = blockIdx.y * GM, n0 = blockIdx.x * GN;
    const int kbeg = blockIdx.z * g.kchunk;
    const int kend = min(g.K, kbeg + g.kchunk);
    const int l16 = lane & 15, lg = lane >> 4;
    m_flt4 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = m_flt4{0.f, 0.f, 0.f, 0.f};
    const bool a_kfast = (g.sak == 1), b_nfast = (g.sbn == 1);
    for (int k0 = kbeg; k0 < kend; k0 += GK) {
#pragma unroll
        for (int u = 0; u < (GM * GK) / GT; ++u) {
            const int e = tid + u * GT;
            const int mm = a_kfast ? e / GK : e % GM;
            const int kk = a_kfast ? e % GK : e / GM;
            const int gm = m0 + mm, gk = k0 + kk;
            As[kk][mm] = (gm < g.M && gk < kend) ? g.A[int64_t(gm) * g.sam + int64_t(gk) * g.sak] : 0.f;
        }
#pragma unroll
        for (int u = 0; u < (GN * GK) / GT; ++u) {
            const int e = tid + u * GT;
            const int nn = b_nfast ? e % GN : e / GK;
            const int kk = b_nfast ? e / GN : e % GK;
            const int gn = n0 + nn, gk = k0 + kk;
            Bs[kk][nn] = (gn < g.N && gk < kend) ? g.B[int64_t(gk) * g.sbk + int64_t(gn) * g.sbn] : 0.f;
        }
        __syncthreads();
#pragma unroll
        for (int ks = 0; ks < GK / 4; ++ks) {
            const int kr = ks * 4 + lg;
            float a[2], b[2];
#pragma unroll
            for (int i = 0; i < 2; ++i) a[i] = As[kr][wm * 32 + i * 16 + l16];
#pragma unroll
            for (int j = 0; j < 2; ++j) b[j] = Bs[kr][wn * 32 + j * 16 + l16];
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[i], b[j], acc[i][j], 0, 0, 0);
        }
        __syncthreads();
    }
    // D row 4 (l / 16) + v, column l % 16 of each 16 x 16 sub-tile
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int m = m0 + wm * 32 + i * 16 + 4 * lg + v;
                const int n = n0 + wn * 32 + j * 16 + l16;
                if (m < g.M && n < g.N) {
                    if (g.part)
                        g.part[(int64_t(blockIdx.z) * g.M + m) * g.N + n] = acc[i][j][v];
                    else
                        store_out(g, m, n, acc[i][j][v]);
                }
            }
}

__global__ void k_gemm_reduce(GemmArgs g, int splits) {
    const int64_t MN = int64_t(g.M) * g.N;
    for (int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; e < MN;
         e += int64_t(gridDim.x) * blockDim.x) {
        float s = 0.f;
        for (int z = 0; z < splits; ++z) s += g.part[int64_t(z) * MN + e];
        store_out(g, int(e / g.N), int(e % g.N), s);
    }
}

// out[n] = sum_m A[m, n] (bias gradients; M is the batch).
__global__ void k_colsum(int M, int N, const float *__restrict__ A, float *__restrict__ out) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    float s = 0.f;
    for (int m = 0; m < M; ++m) s += A[int64_t(m) * N + n];
    out[n] = s;
}

__device__ inline float bf16_to_f32(uint16_t h) { return __uint_as_float(uint32_t(h) << 16); }

__device__ inline uint16_t f32_to_bf16(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);
    return uint16_t(u >> 16);
}

// y[n] = act(sum_k W[n, k] x[k] + b[n]) with W in BF16 (row stride ldw),
// x and y FP32; TPR threads per row.
template <int TPR>
__global__ __launch_bounds__(256) void k_gemv_bf16(int N, int K, const uint16_t *__restrict__ W,
                                                   int64_t ldw, const float *__restrict__ x,
                                                   const float *__restrict__ bias, int relu,
                                                   float *__restrict__ y) {
    constexpr int RPB = 256 / TPR;
    __shared__ float red[256];
    const int tid = threadIdx.x;
    const int row = blockIdx.x * RPB + tid / TPR;
    const int lane = tid % TPR;
    float s = 0.f;
    if (row < N) {
        const uint16_t *w = W + int64_t(row) * ldw;
        for (int k = lane; k < K; k += TPR) s += bf16_to_f32(w[k]) * x[k];
    }
    red[tid] = s;
    __syncthreads();
    for (int st = TPR / 2; st > 0; st >>= 1) {
        if (lane < st) red[tid] += red[tid + st];
        __syncthreads();
    }
    if (lane == 0 && row < N) {
        float v = red[tid] + (bias ? bias[row] : 0.f);
        y[row] = (relu && v < 0.f) ? 0.f : v;
    }
}

__global__ void k_adam(int64_t n, float *__restrict__ p, const float *__restrict__ gr,
                       float *__restrict__ m, float *__restrict__ v, float lr, float b1, float b2,
                       float eps, float wd, float c1, float c2) {
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += int64_t(gridDim.x) * blockDim.x) {
        const float g = gr[i];
        const float mi = b1 * m[i] + (1.f - b1) * g;
        const float vi = b2 * v[i] + (1.f - b2) * g * g;
        m[i] = mi;
        v[i] = vi;
        p[i] -= lr * ((mi / c1) / (sqrtf(vi / c2) + eps) + wd * p[i]);
    }
}

// dout = scale (out - y); loss += sum (out - y)^2 (block partials, one
// atomic per block).
__global__ __launch_bounds__(256) void k_mse_grad(int64_t n, float scale,
                                                  const float *__restrict__ out,
                                                  const float *__restrict__ y,
                                                  float *__restrict__ dout,
                                                  double *__restrict__ loss) {
    __shared__ double red[256];
    double s = 0.0;
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += int64_t(gridDim.x) * blockDim.x) {
        const float d = out[i] - y[i];
        dout[i] = scale * d;
        s += double(d) * d;
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if (threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x == 0 && loss) atomicAdd(loss, red[0]);
}

__global__ void k_f32_to_bf16(int64_t n, const float *__restrict__ s, uint16_t *__restrict__ d) {
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += int64_t(gridDim.x) * blockDim.x)
        d[i] = f32_to_bf16(s[i]);
}

__global__ void k_fill_f32(int64_t n, float *__restrict__ d, float v) {
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += int64_t(gridDim.x) * blockDim.x)
        d[i] = v;
}

// ---------------------------------------------------------------- coarse grid
__device__ inline bool c_on_bdry(int bd, int i, int j, int nx, int ny) {
    return ((bd & 1) && i == nx - 1) || ((bd & 2) && j == ny - 1) || ((bd & 4) && i == 0) ||
           ((bd & 8) && j == 0);
}

template <int NV>
__device__ inline void cblock_sum(double (&v)[NV], double *sh) {
    const int tid = threadIdx.x;
#pragma unroll
    for (int q = 0; q < NV; ++q) sh[q * 256 + tid] = v[q];
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (tid < s) {
#pragma unroll
            for (int q = 0; q < NV; ++q) sh[q * 256 + tid] += sh[q * 256 + tid + s];
        }
        __syncthreads();
    }
}

// mode 0: out[q * nc + g] += fcol * sum_{beta in column} v_beta t_q(th_beta),
//         t = (1, cos th, sin th)   (moments of a weak-form residual)
// mode 1: out[g] += fcell * (sum w_beta v_beta) / (sum w_beta)
//         (quadrature-weighted column average of nodal values)
template <typename T>
__global__ __launch_bounds__(256) void k_restrict(DevMesh m, const int32_t *__restrict__ cstart,
                                                  const int32_t *__restrict__ ccell,
                                                  const double *__restrict__ fcell,
                                                  const double *__restrict__ fcol, int nc,
                                                  int mode, const T *__restrict__ v,
                                                  float *__restrict__ out) {
    __shared__ double sh[3 * 256];
    const int c = blockIdx.x;
    const int nx = m.col_nx[c], ny = m.col_ny[c];
    const double dx = m.col_geom[4 * c + 2] - m.col_geom[4 * c + 0];
    const double dy = m.col_geom[4 * c + 3] - m.col_geom[4 * c + 1];
    const int ga0 = m.col_ang_start[c];
    double acc[3] = {0.0, 0.0, 0.0};
    for (int k = m.col_cell_start[c]; k < m.col_cell_start[c + 1]; ++k) {
        const int nth = m.cell_nth[k], off = m.cell_dof_off[k];
        const int a0 = ga0 + m.cell_ang_off[k];
        const double dth = m.cell_geom[2 * k + 1] - m.cell_geom[2 * k];
        const int n = nx * ny * nth;
        for (int t = threadIdx.x; t < n; t += 256) {
            const int a = t % nth;
            const double val = double(v[off + t]);
            if (mode == 0) {
                acc[0] += val;
                acc[1] += val * m.ang_trig[3 * (a0 + a)];
                acc[2] += val * m.ang_trig[3 * (a0 + a) + 1];
            } else {
                const int i = t / (ny * nth), j = (t / nth) % ny;
                const double w = dx * dy * dth * 0.125 * m.lgl_w[nx * RTX_NMAX + i] *
                                 m.lgl_w[ny * RTX_NMAX + j] * m.ang_trig[3 * (a0 + a) + 2];
                acc[0] += w * val;
                acc[1] += w;
            }
        }
    }
    cblock_sum<3>(acc, sh);
    if (threadIdx.x == 0) {
        for (int e = cstart[c]; e < cstart[c + 1]; ++e) {
            const int g = ccell[e];
            if (mode == 0) {
                for (int q = 0; q < 3; ++q) atomicAdd(&out[q * nc + g], float(fcol[e] * sh[q * 256]));
            } else if (sh[256] > 0.0) {
                atomicAdd(&out[g], float(fcell[e] * sh[0] / sh[256]));
            }
        }
    }
}

// z_beta = s * sum_e fcol_e phi[cell_e] on interior DOFs of each column,
// 0 on inflow-boundary DOFs (accumulate: z += instead of z =).
template <typename T>
__global__ __launch_bounds__(256) void k_prolong(DevMesh m, const int32_t *__restrict__ cstart,
                                                 const int32_t *__restrict__ ccell,
                                                 const double *__restrict__ fcol,
                                                 const float *__restrict__ phi,
                                                 const float *__restrict__ scale, int accumulate,
                                                 T *__restrict__ z) {
    __shared__ double uc;
    const int c = blockIdx.x;
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int e = cstart[c]; e < cstart[c + 1]; ++e) s += fcol[e] * double(phi[ccell[e]]);
        uc = s * (scale ? double(*scale) : 1.0);
    }
    __syncthreads();
    const T u = T(uc);
    const int nx = m.col_nx[c], ny = m.col_ny[c];
    for (int k = m.col_cell_start[c]; k < m.col_cell_start[c + 1]; ++k) {
        const int nth = m.cell_nth[k], off = m.cell_dof_off[k], bd = m.cell_bdry[k];
        const int n = nx * ny * nth;
        for (int t = threadIdx.x; t < n; t += 256) {
            const int i = t / (ny * nth), j = (t / nth) % ny;
            const T val = (bd && c_on_bdry(bd, i, j, nx, ny)) ? T(0) : u;
            z[off + t] = accumulate ? T(z[off + t] + val) : val;
        }
    }
}

// s = |m| (1 when m = 0); x = m / s.  One block.
__global__ __launch_bounds__(1024) void k_scale_features(int64_t n, const float *__restrict__ mom,
                                                         float *__restrict__ x,
                                                         float *__restrict__ scale) {
    __shared__ double sh[1024];
    double s = 0.0;
    for (int64_t i = threadIdx.x; i < n; i += 1024) s += double(mom[i]) * mom[i];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int st = 512; st > 0; st >>= 1) {
        if (threadIdx.x < st) sh[threadIdx.x] += sh[threadIdx.x + st];
        __syncthreads();
    }
    const double nrm = sqrt(sh[0]);
    const float sc = nrm > 0.0 ? float(nrm) : 1.f;
    for (int64_t i = threadIdx.x; i < n; i += 1024) x[i] = mom[i] / sc;
    if (threadIdx.x == 0) *scale = sc;
}

// partial[b] = (sum r q, sum q q, sum r r) over block b; then alpha = max(0, rq / qq).
constexpr int DNB = 512;

template <typename T>
__global__ __launch_bounds__(256) void k_dot3(int64_t n, const T *__restrict__ r,
                                              const T *__restrict__ q, double *__restrict__ part) {
    __shared__ double sh[3 * 256];
    double acc[3] = {0.0, 0.0, 0.0};
    for (int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < n; i += int64_t(gridDim.x) * 256) {
        const double qi = double(q[i]), ri = double(r[i]);
        acc[0] += ri * qi;
        acc[1] += qi * qi;
        acc[2] += ri * ri;
    }
    cblock_sum<3>(acc, sh);
    if (threadIdx.x == 0) {
        part[3 * blockIdx.x] = sh[0];
        part[3 * blockIdx.x + 1] = sh[256];
        part[3 * blockIdx.x + 2] = sh[512];
    }
}

__global__ __launch_bounds__(256) void k_alpha(int nb, const double *__restrict__ part,
                                               double *__restrict__ alpha) {
    __shared__ double sh[3 * 256];
    double acc[3] = {0.0, 0.0, 0.0};
    for (int b = threadIdx.x; b < nb; b += 256) {
        acc[0] += part[3 * b];
        acc[1] += part[3 * b + 1];
        acc[2] += part[3 * b + 2];
    }
    cblock_sum<3>(acc, sh);
    if (threadIdx.x == 0) {
        const double rq = sh[0], qq = sh[256], rr = sh[512];
        double a = (qq > 0.0) ? rq / qq : 0.0;
        if (!(a > 0.0)) a = 0.0;     // also clears NaN
        alpha[0] = a;
        alpha[1] = rq;
        alpha[2] = qq;
        alpha[3] = rr;
    }
}

// alpha from two coarse moment vectors (one workgroup): the step that
// minimises |R r - alpha R q|, i.e. the restricted residual after the coarse
// correction (alpha = 1 for an exact Galerkin coarse solve).
__global__ __launch_bounds__(256) void k_coarse_alpha(int64_t n, const float *__restrict__ mr,
                                                      const float *__restrict__ mq,
                                                      double *__restrict__ alpha) {
    __shared__ double sh[3 * 256];
    double acc[3] = {0.0, 0.0, 0.0};
    for (int64_t i = threadIdx.x; i < n; i += 256) {
        const double a = mr[i], b = mq[i];
        acc[0] += a * b;
        acc[1] += b * b;
        acc[2] += a * a;
    }
    cblock_sum<3>(acc, sh);
    if (threadIdx.x == 0) {
        const double rq = sh[0], qq = sh[256], rr = sh[512];
        double a = (qq > 0.0) ? rq / qq : 0.0;
        if (!(a > 0.0)) a = 0.0;
        alpha[0] = a;
        alpha[1] = rq;
        alpha[2] = qq;
        alpha[3] = rr;
    }
}

template <typename T>
__global__ void k_axpy_dev(int64_t n, const double *__restrict__ alpha, double sign,
                           const T *__restrict__ x, T *__restrict__ y) {
    const T a = T(sign * alpha[0]);
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += int64_t(gridDim.x) * blockDim.x)
        y[i] += a * x[i];
}

int grid_for(int64_t n) {
    int64_t b = (n + 255) / 256;
    return int(b < 2048 ? (b > 0 ? b : 1) : 2048);
}

int gemm_splits(int M, int N, int K) {
    const int tiles = ((M + GM - 1) / GM) * ((N + GN - 1) / GN);
    if (tiles >= 256 || K <= 4 * GK) return 1;
    int s = (256 + tiles - 1) / tiles;
    const int maxs = (K + GK - 1) / GK;
    if (s > maxs) s = maxs;
    if (s > SPLIT_MAX) s = SPLIT_MAX;
    return s < 1 ? 1 : s;
}

}  // namespace

}  // namespace rtx

using namespace rtx;

extern "C" int rtx_coarse_setup(rtx_plan *p, int32_t n_coarse, int32_t nmap,
                                const int32_t *col_start, const int32_t *cell,
                                const double *fcell, const double *fcol) {
    RTX_CHECK(p && col_start && cell && fcell && fcol, "rtx_coarse_setup: null argument");
    RTX_CHECK(n_coarse > 0 && nmap > 0, "rtx_coarse_setup: empty coarse map");
    const int ncol = p->m.ncol;
    RTX_CHECK(col_start[0] == 0 && col_start[ncol] == nmap,
              "rtx_coarse_setup: col_start must run from 0 to nmap");
    for (int c = 0; c < ncol; ++c)
        RTX_CHECK(col_start[c + 1] >= col_start[c], "rtx_coarse_setup: col_start not monotone");
    for (int e = 0; e < nmap; ++e)
        RTX_CHECK(cell[e] >= 0 && cell[e] < n_coarse, "rtx_coarse_setup: coarse cell %d out of range",
                  cell[e]);
    RTX_HIP(hipSetDevice(p->device));
    if (p->coarse) {
        RTX_HIP(hipFree(p->coarse));
        p->coarse = nullptr;
    }
    const size_t b_start = size_t(ncol + 1) * 4, b_cell = size_t(nmap) * 4;
    const size_t o_cell = (b_start + 255) / 256 * 256;
    const size_t o_fcell = (o_cell + b_cell + 255) / 256 * 256;
    const size_t o_fcol = o_fcell + (size_t(nmap) * 8 + 255) / 256 * 256;
    const size_t total = o_fcol + size_t(nmap) * 8;
    void *buf = nullptr;
    RTX_HIP(hipMalloc(&buf, total));
    char *b = static_cast<char *>(buf);
    RTX_HIP(hipMemcpy(b, col_start, b_start, hipMemcpyHostToDevice));
    RTX_HIP(hipMemcpy(b + o_cell, cell, b_cell, hipMemcpyHostToDevice));
    RTX_HIP(hipMemcpy(b + o_fcell, fcell, size_t(nmap) * 8, hipMemcpyHostToDevice));
    RTX_HIP(hipMemcpy(b + o_fcol, fcol, size_t(nmap) * 8, hipMemcpyHostToDevice));
    p->coarse = buf;
    p->coarse_n = n_coarse;
    p->coarse_nmap = nmap;
    p->cmap_start = reinterpret_cast<const int32_t *>(b);
    p->cmap_cell = reinterpret_cast<const int32_t *>(b + o_cell);
    p->cmap_fcell = reinterpret_cast<const double *>(b + o_fcell);
    p->cmap_fcol = reinterpret_cast<const double *>(b + o_fcol);
    return RTX_OK;
}

extern "C" int rtx_coarse_restrict(rtx_plan *p, int mode, int precision, const void *d_v,
                                   float *d_out, void *stream) {
    RTX_CHECK(p && d_v && d_out, "rtx_coarse_restrict: null argument");
    RTX_CHECK(p->coarse, "rtx_coarse_restrict: no coarse map (rtx_coarse_setup)");
    RTX_CHECK(mode == 0 || mode == 1, "rtx_coarse_restrict: mode must be 0 or 1");
    hipStream_t st = static_cast<hipStream_t>(stream);
    const int64_t nout = (mode == 0 ? 3 : 1) * int64_t(p->coarse_n);
    k_fill_f32<<<grid_for(nout), 256, 0, st>>>(nout, d_out, 0.f);
    if (precision == RTX_PREC_F64)
        k_restrict<double><<<p->m.ncol, 256, 0, st>>>(p->m, p->cmap_start, p->cmap_cell,
                                                      p->cmap_fcell, p->cmap_fcol, p->coarse_n,
                                                      mode, static_cast<const double *>(d_v), d_out);
    else
        k_restrict<float><<<p->m.ncol, 256, 0, st>>>(p->m, p->cmap_start, p->cmap_cell,
                                                     p->cmap_fcell, p->cmap_fcol, p->coarse_n,
                                                     mode, static_cast<const float *>(d_v), d_out);
    RTX_HIP(hipGetLastError());
    return RTX_OK;
}

extern "C" int rtx_coarse_prolong(rtx_plan *p, int precision, const float *d_phi,
                                  const float *d_scale, int accumulate, void *d_z, void *stream) {
    RTX_CHECK(p && d_phi && d_z, "rtx_coarse_prolong: null argument");
    RTX_CHECK(p->coarse, "rtx_coarse_prolong: no coarse map (rtx_coarse_setup)");
    hipStream_t st = static_cast<hipStream_t>(stream);
    if (precision == RTX_PREC_F64)
        k_prolong<double><<<p->m.ncol, 256, 0, st>>>(p->m, p->cmap_start, p->cmap_cell, p->cmap_fcol,
                                                     d_phi, d_scale, accumulate,
                                                     static_cast<double *>(d_z));
    else
        k_prolong<float><<<p->m.ncol, 256, 0, st>>>(p->m, p->cmap_start, p->cmap_cell, p->cmap_fcol,
                                                    d_phi, d_scale, accumulate,
                                                    static_cast<float *>(d_z));
    RTX_HIP(hipGetLastError());
    return RTX_OK;
}

extern "C" int rtx_scale_features(int64_t n, const float *d_m, float *d_x, float *d_scale,
                                  void *stream) {
    RTX_CHECK(n > 0 && d_m && d_x && d_scale, "rtx_scale_features: bad argument");
    k_scale_features<<<1, 1024, 0, static_cast<hipStream_t>(stream)>>>(n, d_m, d_x, d_scale);
    RTX_HIP(hipGetLastError());
    return RTX_OK;
}

extern "C" int rtx_mlp_gemm_workspace(int M, int N, int K, int64_t *n_floats) {
    RTX_CHECK(M > 0 && N > 0 && K > 0 && n_floats, "rtx_mlp_gemm_workspace: bad argument");
    const int s = gemm_splits(M, N, K);
    *n_floats = s > 1 ? int64_t(s) * M * N : 0;
    return RTX_OK;
}

extern "C" int rtx_mlp_gemm(int M, int N, int K, const float *d_A, int64_t sam, int64_t sak,
                            const float *d_B, int64_t sbk, int64_t sbn, float *d_C, int64_t ldc,
                            int epilogue, const float *d_bias, float *d_H, const float *d_Z,
                            int64_t ldz, float *d_work, int64_t work_floats, void *stream) {
    RTX_CHECK(M > 0 && N > 0 && K > 0, "rtx_mlp_gemm: empty shape %d x %d x %d", M, N, K);
    RTX_CHECK(d_A && d_B && d_C, "rtx_mlp_gemm: null operand");
    RTX_CHECK(ldc >= N, "rtx_mlp_gemm: ldc %lld < N %d", (long long)ldc, N);
    RTX_CHECK(epilogue >= EPI_NONE && epilogue <= EPI_MASK, "rtx_mlp_gemm: bad epilogue");
    RTX_CHECK(epilogue != EPI_BIAS || d_bias, "rtx_mlp_gemm: bias epilogue needs d_bias");
    RTX_CHECK(epilogue != EPI_MASK || (d_Z && ldz >= N), "rtx_mlp_gemm: mask epilogue needs d_Z");
    hipStream_t st = static_cast<hipStream_t>(stream);
    GemmArgs g{M, N, K, d_A, sam, sak, d_B, sbk, sbn, d_C, ldc, epilogue, d_bias, d_H, d_Z, ldz,
               nullptr, K};
    int s = gemm_splits(M, N, K);
    if (s > 1 && (!d_work || work_floats < int64_t(s) * M * N)) s = 1;
    if (s > 1) {
        int kc = (K + s - 1) / s;
        kc = (kc + GK - 1) / GK * GK;
        s = (K + kc - 1) / kc;
        g.kchunk = kc;
        g.part = d_work;
    }
    dim3 grid((N + GN - 1) / GN, (M + GM - 1) / GM, s);
    k_gemm<<<grid, GT, 0, st>>>(g);
    if (s > 1) k_gemm_reduce<<<grid_for(int64_t(M) * N), 256, 0, st>>>(g, s);
    RTX_HIP(hipGetLastError());
    return RTX_OK;
}

extern "C" int rtx_mlp_colsum(int M, int N, const float *d_A, float *d_out, void *stream) {
    RTX_CHECK(M > 0 && N > 0 && d_A && d_out, "rtx_mlp_colsum: bad argument");
    k_colsum<<<(N + 255) / 256, 256, 0, static_cast<hipStream_t>(stream)>>>(M, N, d_A, d_out);
    RTX_HIP(hipGetLastError());
    return RTX_OK;
}

extern "C" int rtx_mlp_gemv_bf16(int N, int K, const uint16_t *d_W, int64_t ldw, const float *d_x,
                                 const float *d_bias, int relu, float *d_y, void *stream) {
    RTX_CHECK(N > 0 && K > 0 && ldw >= K && d_W && d_x && d_y, "rtx_mlp_gemv_bf16: bad argument");
    hipStream_t st = static_cast<hipStream_t>(stream);
    if (K >= 2048)
        k_gemv_bf16<256><<<N, 256, 0, st>>>(N, K, d_W, ldw, d_x, d_bias, relu, d_y);
    else
        k_gemv_bf16<64><<<(N + 3) / 4, 256, 0, st>>>(N, K, d_W, ldw, d_x, d_bias, relu, d_y);
    RTX_HIP(hipGetLastError());
    return RTX_OK;
}

extern "C" int rtx_mlp_to_bf16(int64_t n, const float *d_src, uint16_t *d_dst, void *stream) {
    RTX_CHECK(n > 0 && d_src && d_dst, "rtx_mlp_to_bf16: bad argument");
    k_f32_to_bf16<<<grid_for(n), 256, 0, static_cast<hipStream_t>(stream)>>>(n, d_src, d_dst);
    RTX_HIP(hipGetLastError());
    return RTX_OK;
}

extern "C" int rtx_adam_step(int64_t n, float *d_p, const float *d_g, float *d_m, float *d_v,
                             float lr, float beta1, float beta2, float eps, float weight_decay,
                             int step, void *stream) {
    RTX_CHECK(n > 0 && d_p && d_g && d_m && d_v && step >= 1, "rtx_adam_step: bad argument");
    const float c1 = 1.f - powf(beta1, float(step)), c2 = 1.f - powf(beta2, float(step));
    k_adam<<<grid_for(n), 256, 0, static_cast<hipStream_t>(stream)>>>(n, d_p, d_g, d_m, d_v, lr, beta1,
                                                                     beta2, eps, weight_decay, c1, c2);
    RTX_HIP(hipGetLastError());
    return RTX_OK;
}

extern "C" int rtx_mse_grad(int64_t n, float scale, const float *d_out, const float *d_y,
                            float *d_dout, double *d_loss, void *stream) {
    RTX_CHECK(n > 0 && d_out && d_y && d_dout, "rtx_mse_grad: bad argument");
    k_mse_grad<<<grid_for(n), 256, 0, static_cast<hipStream_t>(stream)>>>(n, scale, d_out, d_y,
                                                                         d_dout, d_loss);
    RTX_HIP(hipGetLastError());
    return RTX_OK;
}

extern "C" int rtx_line_search(int precision, int64_t n, const void *d_r, const void *d_q,
                               double *d_alpha, double *d_work, void *stream) {
    RTX_CHECK(n > 0 && d_r && d_q && d_alpha && d_work, "rtx_line_search: bad argument");
    hipStream_t st = static_cast<hipStream_t>(stream);
    if (precision == RTX_PREC_F64)
        k_dot3<double><<<DNB, 256, 0, st>>>(n, static_cast<const double *>(d_r),
                                            static_cast<const double *>(d_q), d_work);
    else
        k_dot3<float><<<DNB, 256, 0, st>>>(n, static_cast<const float *>(d_r),
                                           static_cast<const float *>(d_q), d_work);
    k_alpha<<<1, 256, 0, st>>>(DNB, d_work, d_alpha);
    RTX_HIP(hipGetLastError());
    return RTX_OK;
}

extern "C" int rtx_coarse_line_search(int64_t n, const float *d_mr, const float *d_mq,
                                      double *d_alpha, void *stream) {
    RTX_CHECK(n > 0 && d_mr && d_mq && d_alpha, "rtx_coarse_line_search: bad argument");
    k_coarse_alpha<<<1, 256, 0, static_cast<hipStream_t>(stream)>>>(n, d_mr, d_mq, d_alpha);
    RTX_HIP(hipGetLastError());
    return RTX_OK;
}

extern "C" int rtx_line_search_workspace(int64_t *n_doubles) {
    RTX_CHECK(n_doubles, "rtx_line_search_workspace: null argument");
    *n_doubles = 3 * DNB;
    return RTX_OK;
}

extern "C" int rtx_axpy_dev(int precision, int64_t n, const double *d_alpha, double sign,
                            const void *d_x, void *d_y, void *stream) {
    RTX_CHECK(n > 0 && d_alpha && d_x && d_y, "rtx_axpy_dev: bad argument");
    hipStream_t st = static_cast<hipStream_t>(stream);
    if (precision == RTX_PREC_F64)
        k_axpy_dev<double><<<grid_for(n), 256, 0, st>>>(n, d_alpha, sign,
                                                        static_cast<const double *>(d_x),
                                                        static_cast<double *>(d_y));
    else
        k_axpy_dev<float><<<grid_for(n), 256, 0, st>>>(n, d_alpha, sign,
                                                       static_cast<const float *>(d_x),
                                                       static_cast<float *>(d_y));
    RTX_HIP(hipGetLastError());
    return RTX_OK;
}
