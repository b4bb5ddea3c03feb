// Entry evaluation for the assembled-matrix interface.
//
// rt.Problem keeps the reference's assembled methods (mass_matrix,
// scattering_matrix, interior_convection_matrix, boundary_convection_matrix,
// preconditioner_matrix of src/rt/*.py), returning scipy CSR matrices.  The
// sparsity pattern is integer work done on the host; every value is
// evaluated here from the same plan tables the matrix-free kernels use, with
// the reference's drop thresholds (|S|, |P| <= 1e-14 and near-zero sigma
// weights between different cells, scattering_matrix.py:114,194,205;
// |BC| <= 1e-15, boundary_convection_matrix.py:163).
#include "rtx_internal.h"

namespace rtx {

constexpr int NMa = RTX_NMAX;

struct CellGeo {
    int c, nx, ny, nth;
    double dx, dy, th0, dth;
};

__device__ inline CellGeo cell_geo(const DevMesh &m, int k) {
    CellGeo g;
    g.c = m.cell_col[k];
    g.nx = m.col_nx[g.c];
    g.ny = m.col_ny[g.c];
    g.nth = m.cell_nth[k];
    g.dx = m.col_geom[4 * g.c + 2] - m.col_geom[4 * g.c + 0];
    g.dy = m.col_geom[4 * g.c + 3] - m.col_geom[4 * g.c + 1];
    g.th0 = m.cell_geom[2 * k];
    g.dth = m.cell_geom[2 * k + 1] - g.th0;
    return g;
}

__device__ inline double node_theta(const DevMesh &m, const CellGeo &g, int r) {
    return g.th0 + g.dth * 0.5 * (m.lg_x[g.nth * NMa + r] + 1.0);
}

__device__ double mass_entry(const DevMesh &m, int k, int l) {
    CellGeo g = cell_geo(m, k);
    const int p = l / (g.ny * g.nth), q = (l / g.nth) % g.ny, r = l % g.nth;
    return (g.dx * g.dy * g.dth * 0.125) * m.lgl_w[g.nx * NMa + p] * m.lgl_w[g.ny * NMa + q] *
           m.lg_w[g.nth * NMa + r] * m.kappa[m.col_sp_off[g.c] + p * g.ny + q];
}

// returned scattering matrix (-S) entry between (k_r, l_r) and (k_c, l_c);
// *small is set when the reference would drop it
__device__ double scat_entry(const DevMesh &m, int kr, int lr, int kc, int lc, bool *skip) {
    CellGeo g = cell_geo(m, kr);
    const int nthc = m.cell_nth[kc];
    const int s = lr / g.nth, r = lr % g.nth, a = lc % nthc;
    const int p = s / g.ny, q = s % g.ny;
    const int na = m.col_nang[g.c];
    const int R = m.cell_ang_off[kr] + r, A = m.cell_ang_off[kc] + a;
    const double wws = m.lgl_w[g.nx * NMa + p] * m.lgl_w[g.ny * NMa + q] *
                       m.sigma[m.col_sp_off[g.c] + s];
    *skip = (kr != kc) && !(fabs(wws) > 1.0e-14);
    return -g.dx * g.dy * wws * m.G[m.col_G_off[g.c] + size_t(A) * na + R];
}

__device__ double ic_entry(const DevMesh &m, int k, int lr, int lc) {
    CellGeo g = cell_geo(m, k);
    const int nyt = g.ny * g.nth;
    const int p = lr / nyt, q = (lr / g.nth) % g.ny, r = lr % g.nth;
    const int i = lc / nyt, j = (lc / g.nth) % g.ny, a = lc % g.nth;
    if (r != a) return 0.0;
    double s, c;
    sincos(node_theta(m, g, r), &s, &c);
    const double *wx = m.lgl_w + g.nx * NMa, *wy = m.lgl_w + g.ny * NMa, *wt = m.lg_w + g.nth * NMa;
    double v = 0.0;
    if (p == i)   // y-derivative term, interior_convection_matrix.py:178-193
        v += (g.dx * g.dth * 0.25) * wx[i] * wy[j] * wt[a] * s *
             m.lgl_D[(size_t(g.ny) * NMa + q) * NMa + j];
    if (q == j)   // x-derivative term, interior_convection_matrix.py:222-237
        v += (g.dy * g.dth * 0.25) * wx[i] * wy[j] * wt[a] * c *
             m.lgl_D[(size_t(g.nx) * NMa + p) * NMa + i];
    return v;
}

__device__ double bc_entry(const DevMesh &m, int kr, int lr, int kc, int lc, int link) {
    CellGeo g = cell_geo(m, kr);
    const int nyt = g.ny * g.nth;
    const int p = lr / nyt, q = (lr / g.nth) % g.ny, r = lr % g.nth;
    if (link < 0) {   // outflow own-trace terms on the faces of the quadrant
        const int quad = m.cell_quad[kr];
        if (quad < 0 || lr != lc || kr != kc) return 0.0;
        double s, c;
        sincos(node_theta(m, g, r), &s, &c);
        double v = 0.0;
        for (int h = 0; h < 2; ++h) {
            const int F = (quad + h) & 3;
            bool on = (F == 0 && p == g.nx - 1) || (F == 2 && p == 0) ||
                      (F == 1 && q == g.ny - 1) || (F == 3 && q == 0);
            if (!on) continue;
            const double tf = F == 0 ? c : (F == 1 ? s : (F == 2 ? -c : -s));
            const double val = ((F & 1) ? g.dx : g.dy) * g.dth * 0.25 *
                               ((F & 1) ? m.lgl_w[g.nx * NMa + p] : m.lgl_w[g.ny * NMa + q]) *
                               m.lg_w[g.nth * NMa + r] * tf;
            if (fabs(val) > 1.0e-15) v += val;
        }
        return v;
    }
    const int F = m.link_face[link];
    const int c1 = m.cell_col[kc];
    const int ny1 = m.col_ny[c1], nt1 = m.cell_nth[kc];
    const int i1 = lc / (ny1 * nt1), j1 = (lc / nt1) % ny1, a = lc % nt1;
    const int m0 = (F & 1) ? p : q, m1 = (F & 1) ? i1 : j1;
    const int n0 = (F & 1) ? g.nx : g.ny;
    const double val = ((F & 1) ? g.dx : g.dy) * g.dth * 0.25 *
                       m.Es_pool[m.link_Es_off[link] + m1 * n0 + m0] *
                       m.Eth_pool[m.link_Eth_off[link] + a * g.nth + r];
    return fabs(val) > 1.0e-15 ? val : 0.0;
}

// part: 0 M, 1 -S, 2 IC, 3 BC, 4 P = M - S
__global__ void k_entries(DevMesh m, int part, int64_t n, const int32_t *__restrict__ kr,
                          const int32_t *__restrict__ lr, const int32_t *__restrict__ kc,
                          const int32_t *__restrict__ lc, const int32_t *__restrict__ link,
                          double *__restrict__ val) {
    int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    for (; e < n; e += int64_t(gridDim.x) * blockDim.x) {
        double v = 0.0;
        const int a = kr[e], b = lr[e], c = kc[e], d = lc[e];
        if (part == 0) {
            v = mass_entry(m, a, b);
        } else if (part == 1 || part == 4) {
            bool skip = false;
            v = scat_entry(m, a, b, c, d, &skip);
            if (part == 4 && a == c && b == d) v += mass_entry(m, a, b);
            if (skip || !(fabs(v) > 1.0e-14)) v = 0.0;
        } else if (part == 2) {
            v = ic_entry(m, a, b, d);
        } else {
            v = bc_entry(m, a, b, c, d, link ? link[e] : -1);
        }
        val[e] = v;
    }
}

}  // namespace rtx

using namespace rtx;

/* Values of operator `part` (0 M, 1 -S, 2 IC, 3 BC, 4 M - S) at the n given
   (row cell, row local DOF, column cell, column local DOF) positions; for BC
   link[e] names the upwind link or -1 for the outflow own-trace term. */
extern "C" int rtx_assemble_entries(rtx_plan *p, int part, int64_t n, const int32_t *d_kr,
                                    const int32_t *d_lr, const int32_t *d_kc, const int32_t *d_lc,
                                    const int32_t *d_link, double *d_val, void *stream) {
    RTX_CHECK(p && d_kr && d_lr && d_kc && d_lc && d_val, "rtx_assemble_entries: null argument");
    RTX_CHECK(part >= 0 && part <= 4, "rtx_assemble_entries: part %d", part);
    RTX_CHECK(part != 3 || d_link, "rtx_assemble_entries: BC needs link indices");
    if (!p->has_coeffs) {
        set_error("rtx_assemble_entries: coefficients not set");
        return RTX_ERR_STATE;
    }
    if (n == 0) return RTX_OK;
    k_entries<<<4096, 256, 0, static_cast<hipStream_t>(stream)>>>(p->m, part, n, d_kr, d_lr,
                                                                 d_kc, d_lc, d_link, d_val);
    RTX_HIP(hipGetLastError());
    return RTX_OK;
}
