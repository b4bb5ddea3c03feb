// Plan creation: copy the flattened hp mesh and coefficient tables to HBM.
#include <algorithm>
#include <cstdarg>
#include <cstddef>
#include <cstring>
#include <vector>

#include "rtx_internal.h"

namespace rtx {

static thread_local char g_err[1024] = "";

void set_error(const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

namespace {

struct Packer {
    struct Item {
        const void *src;
        size_t bytes;
        size_t off;
    };
    std::vector<Item> items;
    size_t total = 0;
    size_t add(const void *src, size_t bytes) {
        size_t off = total;
        items.push_back({src, bytes, off});
        total += (bytes + 255) & ~size_t(255);
        return off;
    }
};

}  // namespace
}  // namespace rtx

using namespace rtx;

extern "C" const char *rtx_last_error(void) { return g_err; }

extern "C" int rtx_mesh_desc_layout(int64_t *o, int n) {
#define RTX_OFF(f) int64_t(offsetof(rtx_mesh_desc, f))
    const int64_t v[] = {
        RTX_OFF(ncol), RTX_OFF(ncell), RTX_OFF(nlink), RTX_OFF(ndof),
        RTX_OFF(col_cell_start), RTX_OFF(col_nx), RTX_OFF(col_ny), RTX_OFF(col_sp_off),
        RTX_OFF(col_nang), RTX_OFF(col_G_off), RTX_OFF(col_geom),
        RTX_OFF(cell_dof_off), RTX_OFF(cell_col), RTX_OFF(cell_nth), RTX_OFF(cell_quad),
        RTX_OFF(cell_ang_off), RTX_OFF(cell_bdry), RTX_OFF(cell_geom),
        RTX_OFF(link_start), RTX_OFF(link_src), RTX_OFF(link_face), RTX_OFF(link_Es_off),
        RTX_OFF(link_Eth_off), RTX_OFF(link_target),
        RTX_OFF(col_ang_start), RTX_OFF(ang_pack), RTX_OFF(ang_trig), RTX_OFF(link_meta),
        RTX_OFF(col_ntrace), RTX_OFF(col_nitems), RTX_OFF(nang),
        RTX_OFF(nEs), RTX_OFF(nEth), RTX_OFF(Es_pool), RTX_OFF(Eth_pool),
        RTX_OFF(lgl_x), RTX_OFF(lgl_w), RTX_OFF(lg_x), RTX_OFF(lg_w), RTX_OFF(lgl_D),
        int64_t(sizeof(rtx_mesh_desc))};
#undef RTX_OFF
    const int count = int(sizeof(v) / sizeof(v[0]));
    for (int i = 0; i < count && i < n && o; ++i) o[i] = v[i];
    return count;
}

extern "C" int rtx_plan_create(const rtx_mesh_desc *d, rtx_plan **out) {
    RTX_CHECK(d != nullptr && out != nullptr, "rtx_plan_create: null argument");
    RTX_CHECK(d->ncol > 0 && d->ncell > 0 && d->nlink >= 0,
              "rtx_plan_create: empty mesh (ncol=%d ncell=%d)", d->ncol, d->ncell);
    RTX_CHECK(d->ndof > 0 && d->ndof < (int64_t(1) << 31),
              "rtx_plan_create: ndof=%lld outside [1, 2^31)", (long long)d->ndof);
    const int nc = d->ncol, nk = d->ncell, nl = d->nlink;
    // Validate the tables the kernels index with.
    int32_t max_nang = 0, max_cell = 0, max_nsp = 0, max_col_dofs = 0, max_col_cells = 0;
    for (int c = 0; c < nc; ++c) {
        int nx = d->col_nx[c], ny = d->col_ny[c];
        RTX_CHECK(nx >= 1 && nx <= RTX_NMAX && ny >= 1 && ny <= RTX_NMAX,
                  "column %d: orders (%d, %d) outside [1, %d]", c, nx, ny, RTX_NMAX);
        int k0 = d->col_cell_start[c], k1 = d->col_cell_start[c + 1];
        RTX_CHECK(k0 >= 0 && k1 > k0 && k1 <= nk, "column %d: bad cell range", c);
        int nang = 0;
        for (int k = k0; k < k1; ++k) {
            RTX_CHECK(d->cell_col[k] == c, "cell %d: owner mismatch", k);
            int nth = d->cell_nth[k];
            RTX_CHECK(nth >= 1 && nth <= RTX_NMAX, "cell %d: nth=%d outside [1, %d]", k,
                      nth, RTX_NMAX);
            RTX_CHECK(d->cell_ang_off[k] == nang, "cell %d: angular offset", k);
            nang += nth;
            int64_t cd = int64_t(nx) * ny * nth;
            RTX_CHECK(d->cell_dof_off[k] >= 0 && d->cell_dof_off[k] + cd <= d->ndof,
                      "cell %d: DOF range outside the vector", k);
            if (cd > max_cell) max_cell = int32_t(cd);
        }
        RTX_CHECK(d->col_nang[c] == nang, "column %d: nang mismatch", c);
        for (int k = k0 + 1; k < k1; ++k)
            RTX_CHECK(d->cell_dof_off[k] == d->cell_dof_off[k - 1] + nx * ny * d->cell_nth[k - 1],
                      "column %d: cell DOF ranges are not contiguous", c);
        if (nang > max_nang) max_nang = nang;
        if (nx * ny * nang > max_col_dofs) max_col_dofs = nx * ny * nang;
        if (k1 - k0 > max_col_cells) max_col_cells = k1 - k0;
        if (nx * ny > max_nsp) max_nsp = nx * ny;
    }
    RTX_CHECK(d->col_cell_start[0] == 0 && d->col_cell_start[nc] == nk,
              "cell ranges do not cover the cells");
    RTX_CHECK(d->link_start[0] == 0 && d->link_start[nk] == nl, "bad link_start");
    for (int k = 0; k < nk; ++k)
        RTX_CHECK(d->link_start[k] <= d->link_start[k + 1], "link_start not sorted");
    std::vector<int32_t> col_links(nc, 0), col_trace(nc, 0);
    for (int k = 0; k < nk; ++k) {
        const int c0 = d->cell_col[k];
        for (int l = d->link_start[k]; l < d->link_start[k + 1]; ++l) {
            int s = d->link_src[l], F = d->link_face[l];
            RTX_CHECK(d->link_target[l] == k, "link %d: target %d is not cell %d", l,
                      d->link_target[l], k);
            RTX_CHECK(s >= 0 && s < nk && F >= 0 && F < 4, "link %d: bad source/face", l);
            const int c1 = d->cell_col[s];
            // E^s is n1 x n0 (trace lengths), E^theta is nth1 x nth0
            int64_t n0 = (F % 2 == 0) ? d->col_ny[c0] : d->col_nx[c0];
            int64_t n1 = (F % 2 == 0) ? d->col_ny[c1] : d->col_nx[c1];
            int64_t es = d->link_Es_off[l], et = d->link_Eth_off[l];
            RTX_CHECK(es >= 0 && es + n0 * n1 <= d->nEs, "link %d: E^s outside its pool", l);
            RTX_CHECK(et >= 0 && et + int64_t(d->cell_nth[k]) * d->cell_nth[s] <= d->nEth,
                      "link %d: E^theta outside its pool", l);
            col_links[c0] += 1;
            col_trace[c0] += int32_t(n1) * d->cell_nth[s];
        }
    }
    int32_t max_col_links = 0, max_col_trace = 0;
    for (int c = 0; c < nc; ++c) {
        if (col_links[c] > max_col_links) max_col_links = col_links[c];
        if (col_trace[c] > max_col_trace) max_col_trace = col_trace[c];
    }
    // The packed fused-kernel tables must agree with the primary tables.
    {
        int64_t g = 0;
        for (int c = 0; c < nc; ++c) {
            const int k0 = d->col_cell_start[c], k1 = d->col_cell_start[c + 1];
            const int nsp = d->col_nx[c] * d->col_ny[c];
            RTX_CHECK(d->col_ang_start[c] == g, "column %d: col_ang_start", c);
            for (int k = k0; k < k1; ++k)
                for (int a = 0; a < d->cell_nth[k]; ++a, ++g) {
                    const int32_t want = (d->cell_dof_off[k] - d->cell_dof_off[k0]) |
                                         (d->cell_nth[k] << 20) | (a << 25);
                    RTX_CHECK(g < d->nang && d->ang_pack[g] == want, "angle %lld: ang_pack",
                              (long long)g);
                }
            (void)nsp;
            int tr = 0, it = 0;
            for (int l = d->link_start[k0]; l < d->link_start[k1]; ++l) {
                const int s = d->link_src[l], F = d->link_face[l], k = d->link_target[l];
                const int c1 = d->cell_col[s];
                const int32_t w1 = F | (d->col_nx[c1] << 2) | (d->col_ny[c1] << 6) |
                                   (d->cell_nth[s] << 10) | (d->cell_bdry[s] << 14) |
                                   ((k - k0) << 18);
                const int32_t *lm = d->link_meta + 4 * int64_t(l);
                RTX_CHECK(lm[0] == d->cell_dof_off[s] && lm[1] == w1 && lm[2] == tr && lm[3] == it,
                          "link %d: link_meta", l);
                tr += ((F & 1) ? d->col_nx[c1] : d->col_ny[c1]) * d->cell_nth[s];
                it += ((F & 1) ? d->col_nx[c] : d->col_ny[c]) * d->cell_nth[k];
            }
            RTX_CHECK(d->col_ntrace[c] == tr && d->col_nitems[c] == it,
                      "column %d: trace/item totals", c);
        }
        RTX_CHECK(d->col_ang_start[nc] == g && d->nang == g, "nang mismatch");
    }

    auto *p = new rtx_plan();
    hipError_t he = hipGetDevice(&p->device);
    if (he == hipSuccess)
        he = hipDeviceGetAttribute(&p->n_cu, hipDeviceAttributeMultiprocessorCount, p->device);
    if (he != hipSuccess) {
        set_error("hipGetDevice: %s", hipGetErrorString(he));
        delete p;
        return RTX_ERR_HIP;
    }
    Packer pk;
    const size_t I = sizeof(int32_t), D = sizeof(double);
    const int NT = (RTX_NMAX + 1) * RTX_NMAX;
    size_t o_ccs = pk.add(d->col_cell_start, (nc + 1) * I);
    size_t o_cnx = pk.add(d->col_nx, nc * I);
    size_t o_cny = pk.add(d->col_ny, nc * I);
    size_t o_csp = pk.add(d->col_sp_off, nc * I);
    size_t o_cna = pk.add(d->col_nang, nc * I);
    size_t o_cgo = pk.add(d->col_G_off, nc * sizeof(int64_t));
    size_t o_cge = pk.add(d->col_geom, 4 * nc * D);
    size_t o_kdo = pk.add(d->cell_dof_off, nk * I);
    size_t o_kco = pk.add(d->cell_col, nk * I);
    size_t o_knt = pk.add(d->cell_nth, nk * I);
    size_t o_kqu = pk.add(d->cell_quad, nk * I);
    size_t o_kao = pk.add(d->cell_ang_off, nk * I);
    size_t o_kbd = pk.add(d->cell_bdry, nk * I);
    size_t o_kge = pk.add(d->cell_geom, 2 * nk * D);
    size_t o_lst = pk.add(d->link_start, (nk + 1) * I);
    size_t o_lsr = pk.add(d->link_src, (nl > 0 ? nl : 1) * I);
    size_t o_lfa = pk.add(d->link_face, (nl > 0 ? nl : 1) * I);
    size_t o_les = pk.add(d->link_Es_off, (nl > 0 ? nl : 1) * I);
    size_t o_let = pk.add(d->link_Eth_off, (nl > 0 ? nl : 1) * I);
    size_t o_ltg = pk.add(d->link_target, (nl > 0 ? nl : 1) * I);
    size_t o_cas = pk.add(d->col_ang_start, (nc + 1) * I);
    size_t o_apk = pk.add(d->ang_pack, d->nang * I);
    size_t o_atr = pk.add(d->ang_trig, 3 * d->nang * D);
    size_t o_lme = pk.add(d->link_meta, 4 * (nl > 0 ? nl : 1) * I);
    size_t o_cnt = pk.add(d->col_ntrace, nc * I);
    size_t o_cni = pk.add(d->col_nitems, nc * I);
    size_t o_esp = pk.add(d->Es_pool, (d->nEs > 0 ? d->nEs : 1) * D);
    size_t o_etp = pk.add(d->Eth_pool, (d->nEth > 0 ? d->nEth : 1) * D);
    size_t o_lx = pk.add(d->lgl_x, NT * D);
    size_t o_lw = pk.add(d->lgl_w, NT * D);
    size_t o_gx = pk.add(d->lg_x, NT * D);
    size_t o_gw = pk.add(d->lg_w, NT * D);
    size_t o_ld = pk.add(d->lgl_D, NT * RTX_NMAX * D);

    he = hipMalloc(&p->arena, pk.total);
    if (he != hipSuccess) {
        set_error("hipMalloc(%zu): %s", pk.total, hipGetErrorString(he));
        delete p;
        return RTX_ERR_HIP;
    }
    p->arena_bytes = pk.total;
    std::vector<char> staging(pk.total, 0);
    for (auto &it : pk.items)
        if (it.src != nullptr && it.bytes > 0) memcpy(staging.data() + it.off, it.src, it.bytes);
    he = hipMemcpy(p->arena, staging.data(), pk.total, hipMemcpyHostToDevice);
    if (he != hipSuccess) {
        set_error("hipMemcpy(plan): %s", hipGetErrorString(he));
        (void)hipFree(p->arena);
        delete p;
        return RTX_ERR_HIP;
    }
    char *b = static_cast<char *>(p->arena);
    DevMesh &m = p->m;
    m.ncol = nc;
    m.ncell = nk;
    m.nlink = nl;
    m.ndof = d->ndof;
    m.col_cell_start = (const int32_t *)(b + o_ccs);
    m.col_nx = (const int32_t *)(b + o_cnx);
    m.col_ny = (const int32_t *)(b + o_cny);
    m.col_sp_off = (const int32_t *)(b + o_csp);
    m.col_nang = (const int32_t *)(b + o_cna);
    m.col_G_off = (const int64_t *)(b + o_cgo);
    m.col_geom = (const double *)(b + o_cge);
    m.cell_dof_off = (const int32_t *)(b + o_kdo);
    m.cell_col = (const int32_t *)(b + o_kco);
    m.cell_nth = (const int32_t *)(b + o_knt);
    m.cell_quad = (const int32_t *)(b + o_kqu);
    m.cell_ang_off = (const int32_t *)(b + o_kao);
    m.cell_bdry = (const int32_t *)(b + o_kbd);
    m.cell_geom = (const double *)(b + o_kge);
    m.link_start = (const int32_t *)(b + o_lst);
    m.link_src = (const int32_t *)(b + o_lsr);
    m.link_face = (const int32_t *)(b + o_lfa);
    m.link_Es_off = (const int32_t *)(b + o_les);
    m.link_Eth_off = (const int32_t *)(b + o_let);
    m.link_target = (const int32_t *)(b + o_ltg);
    m.col_ang_start = (const int32_t *)(b + o_cas);
    m.ang_pack = (const int32_t *)(b + o_apk);
    m.ang_trig = (const double *)(b + o_atr);
    m.link_meta = (const int32_t *)(b + o_lme);
    m.col_ntrace = (const int32_t *)(b + o_cnt);
    m.col_nitems = (const int32_t *)(b + o_cni);
    m.Es_pool = (const double *)(b + o_esp);
    m.Eth_pool = (const double *)(b + o_etp);
    m.lgl_x = (const double *)(b + o_lx);
    m.lgl_w = (const double *)(b + o_lw);
    m.lg_x = (const double *)(b + o_gx);
    m.lg_w = (const double *)(b + o_gw);
    m.lgl_D = (const double *)(b + o_ld);
    p->max_nang = max_nang;
    p->max_cell_ndof = max_cell;
    p->max_nsp = max_nsp;
    p->max_col_dofs = max_col_dofs;
    p->max_col_cells = max_col_cells;
    p->max_col_links = max_col_links;
    p->max_col_trace = max_col_trace;
    p->min_col_cells = max_col_cells;
    {
        // uniform bin detection
        bool ok = true;
        const int N = d->col_nx[0], NTH = d->cell_nth[0];
        const int NC = d->col_cell_start[1] - d->col_cell_start[0];
        for (int c = 0; c < nc && ok; ++c) {
            ok = d->col_nx[c] == N && d->col_ny[c] == N &&
                 d->col_cell_start[c + 1] - d->col_cell_start[c] == NC &&
                 d->col_G_off[c] == d->col_G_off[0];
        }
        for (int k = 0; k < nk && ok; ++k) ok = d->cell_nth[k] == NTH;
        // the uniform kernel computes a column's first cell, DOF, spatial
        // node and angle from the column index
        for (int c = 0; c < nc && ok; ++c)
            ok = d->col_cell_start[c] == c * NC &&
                 int64_t(d->cell_dof_off[c * NC]) == int64_t(c) * N * N * NTH * NC &&
                 d->col_sp_off[c] == c * N * N && d->col_ang_start[c] == c * NTH * NC;
        // every link must couple equal shapes (trace of N x NTH values), and a
        // cell may have at most one link on an x-face (F even) and one on a
        // y-face (F odd): the kernel adds a corner node's two face fluxes in
        // one thread, without atomics
        for (int l = 0; l < nl && ok; ++l) {
            const int s = d->link_src[l];
            ok = d->cell_nth[s] == NTH && d->col_nx[d->cell_col[s]] == N;
        }
        for (int k = 0; k < nk && ok; ++k) {
            int par = 0;
            for (int l = d->link_start[k]; l < d->link_start[k + 1] && ok; ++l) {
                const int bit = 1 << (d->link_face[l] & 1);
                ok = (par & bit) == 0;
                par |= bit;
            }
        }
        p->uni_ok = ok && uniform_shape_compiled(N, NTH, NC);
        p->uni_n = N;
        p->uni_nth = NTH;
        p->uni_ncells = NC;
    }
    for (int c = 0; c < nc; ++c) {
        const int k0 = d->col_cell_start[c], k1 = d->col_cell_start[c + 1], ncl = k1 - k0;
        if (ncl < p->min_col_cells) p->min_col_cells = ncl;
        const int nsp = d->col_nx[c] * d->col_ny[c];
        for (int H : {1, 2, 4}) {
            for (int h = 0; h < H; ++h) {
                const int lo = k0 + (h * ncl) / H, hi = k0 + ((h + 1) * ncl) / H;
                int dofs = 0, tr = 0;
                for (int k = lo; k < hi; ++k) dofs += nsp * d->cell_nth[k];
                const int l0 = d->link_start[lo], l1 = d->link_start[hi];
                for (int l = l0; l < l1; ++l) {
                    const int s = d->link_src[l], F = d->link_face[l], c1 = d->cell_col[s];
                    tr += ((F & 1) ? d->col_nx[c1] : d->col_ny[c1]) * d->cell_nth[s];
                }
                if (dofs > p->max_part_dofs[H]) p->max_part_dofs[H] = dofs;
                if (tr > p->max_part_trace[H]) p->max_part_trace[H] = tr;
                if (l1 - l0 > p->max_part_links[H]) p->max_part_links[H] = l1 - l0;
            }
        }
    }
    // Order bins: every column goes to the fused kernel with the smallest
    // part count H whose part fits 32 KB of LDS (else the smallest within the
    // kernel's cap), or to the generic kernels when no H fits.
    {
        std::vector<int32_t> lists[RTX_NBINS], gcols, gcells;
        const int Hs[3] = {1, 2, 4};
        for (int c = 0; c < nc; ++c) {
            const int k0 = d->col_cell_start[c], k1 = d->col_cell_start[c + 1], ncl = k1 - k0;
            const int nsp = d->col_nx[c] * d->col_ny[c];
            const int na = d->col_nang[c];
            int64_t pd[3] = {0, 0, 0}, pt[3] = {0, 0, 0};
            int pl[3] = {0, 0, 0};
            for (int hi_ = 0; hi_ < 3; ++hi_) {
                const int H = Hs[hi_];
                for (int h = 0; h < H; ++h) {
                    const int lo = k0 + (h * ncl) / H, hi = k0 + ((h + 1) * ncl) / H;
                    int64_t dofs = 0, tr = 0;
                    for (int k = lo; k < hi; ++k) dofs += int64_t(nsp) * d->cell_nth[k];
                    const int l0 = d->link_start[lo], l1 = d->link_start[hi];
                    for (int l = l0; l < l1; ++l) {
                        const int s = d->link_src[l], F = d->link_face[l], c1 = d->cell_col[s];
                        tr += ((F & 1) ? d->col_nx[c1] : d->col_ny[c1]) * d->cell_nth[s];
                    }
                    pd[hi_] = std::max(pd[hi_], dofs);
                    pt[hi_] = std::max(pt[hi_], tr);
                    pl[hi_] = std::max(pl[hi_], l1 - l0);
                }
            }
            int pick = -1, cls = 0;
            for (int pass = 0; pass < 2 && pick < 0; ++pass)
                for (int hi_ = 0; hi_ < 3 && pick < 0; ++hi_) {
                    if (Hs[hi_] > ncl) continue;
                    const size_t b64 = fused_bytes(int64_t(nsp) * na, na, pd[hi_], pt[hi_], 8);
                    if (!fused_column_ok(ncl, nsp, na, pl[hi_], b64)) continue;
                    if (pass == 0 && b64 > 32 * 1024) continue;
                    pick = hi_;
                    cls = pass;
                }
            if (pick < 0) {
                gcols.push_back(c);
                for (int k = k0; k < k1; ++k) gcells.push_back(k);
                continue;
            }
            FusedBin &fb = p->fbin[3 * cls + pick];
            lists[3 * cls + pick].push_back(c);
            for (int hi_ = 0; hi_ < 3; ++hi_) {
                const int64_t cd = int64_t(nsp) * na;
                fb.bytes[0][hi_] = std::max(fb.bytes[0][hi_], fused_bytes(cd, na, pd[hi_], pt[hi_], 8));
                fb.bytes[1][hi_] = std::max(fb.bytes[1][hi_], fused_bytes(cd, na, pd[hi_], pt[hi_], 4));
                fb.part_links[hi_] = std::max(fb.part_links[hi_], pl[hi_]);
            }
            fb.min_cells = fb.n == 0 ? ncl : std::min(fb.min_cells, ncl);
            fb.n += 1;
            fb.H = Hs[pick];
        }
        p->h_col_cell_start.assign(d->col_cell_start, d->col_cell_start + nc + 1);
        p->h_col_nsp.resize(nc);
        p->h_col_nang.assign(d->col_nang, d->col_nang + nc);
        for (int c = 0; c < nc; ++c) p->h_col_nsp[c] = d->col_nx[c] * d->col_ny[c];
        for (int b = 0; b < RTX_NBINS; ++b) p->h_bin_cols[b] = lists[b];
        p->h_over_cols = gcols;
        std::vector<int32_t> all_cols(nc);
        for (int c = 0; c < nc; ++c) all_cols[c] = c;
        const std::vector<int32_t> sc_all = scatter_items(p, all_cols);
        std::vector<int32_t> bcells[RTX_NBINS], bitems[RTX_NBINS];
        for (int b = 0; b < RTX_NBINS; ++b) {
            for (int c : lists[b])
                for (int k = d->col_cell_start[c]; k < d->col_cell_start[c + 1]; ++k)
                    bcells[b].push_back(k);
            bitems[b] = scatter_items(p, lists[b]);
        }
        size_t total = sc_all.size();
        for (int b = 0; b < RTX_NBINS; ++b)
            total += lists[b].size() + bcells[b].size() + bitems[b].size();
        if (total > 0) {
            RTX_HIP(hipMalloc(&p->bins, total * sizeof(int32_t)));
            int32_t *base = static_cast<int32_t *>(p->bins);
            size_t o = 0;
            auto put = [&](const std::vector<int32_t> &v) -> const int32_t * {
                if (v.empty()) return nullptr;
                (void)hipMemcpy(base + o, v.data(), v.size() * sizeof(int32_t),
                                hipMemcpyHostToDevice);
                const int32_t *ptr = base + o;
                o += v.size();
                return ptr;
            };
            for (int b = 0; b < RTX_NBINS; ++b) {
                FusedBin &fb = p->fbin[b];
                fb.cols = put(lists[b]);
                fb.cells = put(bcells[b]);
                fb.ncells = int(bcells[b].size());
                fb.items = put(bitems[b]);
                fb.nitems = int(bitems[b].size() / 2);
            }
            p->sc_all = put(sc_all);
            p->n_sc_all = int(sc_all.size() / 2);
            RTX_HIP(hipGetLastError());
        }
        // every bin starts on the generic kernels (one wavefront per cell, a
        // tiled scattering GEMM): no atomics, so results are bit-reproducible,
        // and on large hp meshes they are faster; rtx_autotune moves a bin to
        // the fused per-column kernel where that measures faster
        for (int b = 0; b < RTX_NBINS; ++b) p->fbin[b].generic = true;
        const int rc = rebuild_generic(p);
        if (rc) return rc;
    }
    // the hp column kernel's tables and LDS classes
    {
        const int rc = hpcol_build(p, d);
        if (rc) return rc;
    }
    // sizes of the column G blocks are checked in set_coeffs
    *out = p;
    return RTX_OK;
}

namespace rtx {

// Columns go to the XCD with the least work so far (work: tiles x angles),
// each XCD's queue in column order, then the queues are interleaved so that
// block b runs on XCD b % 8 and a column's items share that XCD's L2.
std::vector<int32_t> scatter_items(const rtx_plan *p, const std::vector<int32_t> &cols) {
    std::vector<std::vector<int32_t>> q(SC_XCD);
    int64_t load[SC_XCD] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int c : cols) {
        const int nsp = p->h_col_nsp[c], na = p->h_col_nang[c];
        const int nst = (nsp + 15) / 16, nrg = (na + SC_RG - 1) / SC_RG;
        int x = 0;
        for (int i = 1; i < SC_XCD; ++i)
            if (load[i] < load[x]) x = i;
        for (int st = 0; st < nst; ++st)
            for (int rg = 0; rg < nrg; ++rg) {
                q[x].push_back(c);
                q[x].push_back(st | (rg << 16));
            }
        load[x] += int64_t(nst) * nrg * na;
    }
    size_t len = 0;
    for (auto &v : q) len = std::max(len, v.size() / 2);
    std::vector<int32_t> out;
    out.reserve(len * SC_XCD * 2);
    for (size_t i = 0; i < len; ++i)
        for (int x = 0; x < SC_XCD; ++x) {
            const bool has = 2 * i < q[x].size();
            out.push_back(has ? q[x][2 * i] : -1);
            out.push_back(has ? q[x][2 * i + 1] : 0);
        }
    return out;
}

int rebuild_generic(rtx_plan *p) {
    std::vector<int32_t> cols = p->h_over_cols;
    for (int b = 0; b < RTX_NBINS; ++b)
        if (p->fbin[b].generic)
            cols.insert(cols.end(), p->h_bin_cols[b].begin(), p->h_bin_cols[b].end());
    std::sort(cols.begin(), cols.end());
    std::vector<int32_t> cells;
    for (int c : cols)
        for (int k = p->h_col_cell_start[c]; k < p->h_col_cell_start[c + 1]; ++k)
            cells.push_back(k);
    const std::vector<int32_t> items = scatter_items(p, cols);
    if (p->gen_alloc) {
        RTX_HIP(hipDeviceSynchronize());   // no launch may still read the old lists
        RTX_HIP(hipFree(p->gen_alloc));
        p->gen_alloc = nullptr;
    }
    p->gen_cols = p->gen_cells = p->sc_gen = nullptr;
    p->n_gen_cols = int(cols.size());
    p->n_gen_cells = int(cells.size());
    p->n_sc_gen = int(items.size() / 2);
    const size_t total = cols.size() + cells.size() + items.size();
    if (total == 0) return RTX_OK;
    RTX_HIP(hipMalloc(&p->gen_alloc, total * sizeof(int32_t)));
    int32_t *base = static_cast<int32_t *>(p->gen_alloc);
    RTX_HIP(hipMemcpy(base, cols.data(), cols.size() * sizeof(int32_t), hipMemcpyHostToDevice));
    RTX_HIP(hipMemcpy(base + cols.size(), cells.data(), cells.size() * sizeof(int32_t),
                      hipMemcpyHostToDevice));
    RTX_HIP(hipMemcpy(base + cols.size() + cells.size(), items.data(),
                      items.size() * sizeof(int32_t), hipMemcpyHostToDevice));
    p->gen_cols = base;
    p->gen_cells = base + cols.size();
    p->sc_gen = base + cols.size() + cells.size();
    return RTX_OK;
}

}  // namespace rtx

extern "C" void rtx_plan_destroy(rtx_plan *p) {
    if (!p) return;
    int cur = 0;
    (void)hipGetDevice(&cur);
    (void)hipSetDevice(p->device);
    if (p->arena) (void)hipFree(p->arena);
    if (p->coef) (void)hipFree(p->coef);
    if (p->bj_off) (void)hipFree(p->bj_off);
    if (p->bj_off16) (void)hipFree(p->bj_off16);
    if (p->bj_off8) (void)hipFree(p->bj_off8);
    if (p->bj_soff) (void)hipFree(p->bj_soff);
    if (p->bj_blk) (void)hipFree(p->bj_blk);
    if (p->bj_apd) (void)hipFree(p->bj_apd);
    if (p->bj_cls) (void)hipFree(p->bj_cls);
    if (p->coarse) (void)hipFree(p->coarse);
    if (p->bins) (void)hipFree(p->bins);
    if (p->gen_alloc) (void)hipFree(p->gen_alloc);
    if (p->hp_alloc) (void)hipFree(p->hp_alloc);
    if (p->hp_es) (void)hipFree(p->hp_es);
    if (p->hp_fork) (void)hipEventDestroy(p->hp_fork);
    if (p->hp_join) (void)hipEventDestroy(p->hp_join);
    if (p->hp_side) (void)hipStreamDestroy(p->hp_side);
    (void)hipSetDevice(cur);
    delete p;
}

// Conversion kernels for reduced-precision copies of the coefficients.
__global__ void k_to_f32(const double *__restrict__ s, float *__restrict__ d, int64_t n) {
    int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    int64_t st = int64_t(gridDim.x) * blockDim.x;
    for (; i < n; i += st) d[i] = float(s[i]);
}

__device__ inline uint16_t f32_to_bf16_rne(float f) {
    uint32_t u = __float_as_uint(f);
    uint32_t lsb = (u >> 16) & 1u;
    u += 0x7fffu + lsb;
    return uint16_t(u >> 16);
}

__global__ void k_to_bf16(const double *__restrict__ s, uint16_t *__restrict__ d, int64_t n) {
    int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    int64_t st = int64_t(gridDim.x) * blockDim.x;
    for (; i < n; i += st) d[i] = f32_to_bf16_rne(float(s[i]));
}

extern "C" int rtx_plan_set_coeffs(rtx_plan *p, int64_t nsp, const double *kappa,
                                   const double *sigma, int64_t nG, const double *G) {
    RTX_CHECK(p && kappa && sigma && G, "rtx_plan_set_coeffs: null argument");
    RTX_CHECK(nsp > 0 && nG > 0, "rtx_plan_set_coeffs: empty coefficient arrays");
    // The caller's tables must cover every column's spatial nodes and G block.
    {
        std::vector<int32_t> nx(p->m.ncol), ny(p->m.ncol), sp(p->m.ncol), na(p->m.ncol);
        std::vector<int64_t> go(p->m.ncol);
        RTX_HIP(hipMemcpy(nx.data(), p->m.col_nx, nx.size() * 4, hipMemcpyDeviceToHost));
        RTX_HIP(hipMemcpy(ny.data(), p->m.col_ny, ny.size() * 4, hipMemcpyDeviceToHost));
        RTX_HIP(hipMemcpy(sp.data(), p->m.col_sp_off, sp.size() * 4, hipMemcpyDeviceToHost));
        RTX_HIP(hipMemcpy(na.data(), p->m.col_nang, na.size() * 4, hipMemcpyDeviceToHost));
        RTX_HIP(hipMemcpy(go.data(), p->m.col_G_off, go.size() * 8, hipMemcpyDeviceToHost));
        for (int c = 0; c < p->m.ncol; ++c) {
            RTX_CHECK(sp[c] >= 0 && int64_t(sp[c]) + nx[c] * ny[c] <= nsp,
                      "column %d: spatial nodes outside kappa/sigma (nsp=%lld)", c,
                      (long long)nsp);
            RTX_CHECK(go[c] >= 0 && go[c] + int64_t(na[c]) * na[c] <= nG,
                      "column %d: G block outside the G pool (nG=%lld)", c, (long long)nG);
        }
    }
    int cur = 0;
    RTX_HIP(hipGetDevice(&cur));
    RTX_CHECK(cur == p->device, "rtx_plan_set_coeffs: plan lives on device %d, current %d",
              p->device, cur);
    auto al = [](size_t x) { return (x + 255) & ~size_t(255); };
    size_t o_k = 0, o_s = al(nsp * 8), o_g = o_s + al(nsp * 8);
    size_t o_kf = o_g + al(nG * 8), o_sf = o_kf + al(nsp * 4), o_gf = o_sf + al(nsp * 4);
    size_t o_gb = o_gf + al(nG * 4);
    size_t total = o_gb + al(nG * 2);
    if (p->coef && p->coef_bytes < total) {
        RTX_HIP(hipFree(p->coef));
        p->coef = nullptr;
    }
    if (!p->coef) {
        RTX_HIP(hipMalloc(&p->coef, total));
        p->coef_bytes = total;
    }
    char *b = static_cast<char *>(p->coef);
    RTX_HIP(hipMemcpy(b + o_k, kappa, nsp * 8, hipMemcpyHostToDevice));
    RTX_HIP(hipMemcpy(b + o_s, sigma, nsp * 8, hipMemcpyHostToDevice));
    RTX_HIP(hipMemcpy(b + o_g, G, nG * 8, hipMemcpyHostToDevice));
    p->m.kappa = (const double *)(b + o_k);
    p->m.sigma = (const double *)(b + o_s);
    p->m.G = (const double *)(b + o_g);
    p->m.kappa_f = (const float *)(b + o_kf);
    p->m.sigma_f = (const float *)(b + o_sf);
    p->m.G_f = (const float *)(b + o_gf);
    p->m.G_bf16 = (const uint16_t *)(b + o_gb);
    int grid = 1024;
    k_to_f32<<<grid, 256>>>(p->m.kappa, (float *)p->m.kappa_f, nsp);
    k_to_f32<<<grid, 256>>>(p->m.sigma, (float *)p->m.sigma_f, nsp);
    k_to_f32<<<grid, 256>>>(p->m.G, (float *)p->m.G_f, nG);
    k_to_bf16<<<grid, 256>>>(p->m.G, (uint16_t *)p->m.G_bf16, nG);
    RTX_HIP(hipGetLastError());
    RTX_HIP(hipDeviceSynchronize());
    p->nsp = nsp;
    p->nG = nG;
    p->has_coeffs = true;
    return RTX_OK;
}
