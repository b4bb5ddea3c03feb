// Internal definitions shared by the rtx translation units.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/rtx.h"

namespace rtx {

void set_error(const char *fmt, ...);

#define RTX_HIP(call)                                                          \
    do {                                                                       \
        hipError_t e_ = (call);                                                \
        if (e_ != hipSuccess) {                                                \
            ::rtx::set_error("%s failed: %s (%s:%d)", #call,                   \
                             hipGetErrorString(e_), __FILE__, __LINE__);       \
            return RTX_ERR_HIP;                                                \
        }                                                                      \
    } while (0)

#define RTX_CHECK(cond, ...)                                                   \
    do {                                                                       \
        if (!(cond)) {                                                         \
            ::rtx::set_error(__VA_ARGS__);                                     \
            return RTX_ERR_ARG;                                                \
        }                                                                      \
    } while (0)

// Device-side view of the plan (passed by value to kernels).
struct DevMesh {
    int32_t ncol, ncell, nlink;
    int64_t ndof;
    const int32_t *col_cell_start, *col_nx, *col_ny, *col_sp_off, *col_nang;
    const int64_t *col_G_off;
    const double *col_geom;
    const int32_t *cell_dof_off, *cell_col, *cell_nth, *cell_quad, *cell_ang_off,
        *cell_bdry;
    const double *cell_geom;
    const int32_t *link_start, *link_src, *link_face, *link_Es_off, *link_Eth_off, *link_target;
    const double *Es_pool, *Eth_pool;
    const double *lgl_x, *lgl_w, *lg_x, *lg_w, *lgl_D;
    // packed per-column tables of the fused kernel
    const int32_t *col_ang_start, *ang_pack, *link_meta, *col_ntrace, *col_nitems;
    const double *ang_trig;
    // coefficients
    const double *kappa, *sigma, *G;
    const float *kappa_f, *sigma_f, *G_f;
    const uint16_t *G_bf16;
};

}  // namespace rtx

struct rtx_plan;
namespace rtx {
size_t fused_lds_bytes(const rtx_plan *p, int H, int prec);
int fused_parts(const rtx_plan *p, int prec);
int launch_uniform(rtx_plan *p, uint32_t ops, int prec, const void *dx, void *dy, hipStream_t st,
                   int accumulate);
bool uniform_shape_compiled(int N, int NTH, int NC);
int uniform_valid_parts(const rtx_plan *p);   // bit H set: H parts valid (H = 1, 2, 4)
int uniform_default_parts(const rtx_plan *p); // parts the uniform kernel launches with
int fused_bin_valid_parts(const rtx_plan *p, int bin, int prec);
// Launch the fused bins (all of them with all_bins, else those not moved to
// the generic kernels; only_bin >= 0 restricts to one bin).
int launch_fused(rtx_plan *p, uint32_t ops, int prec, const void *dx, void *dy, hipStream_t st,
                 int accumulate, int only_bin = -1, bool all_bins = false);
// Scattering work items of a set of columns (pairs: column, tile word).
std::vector<int32_t> scatter_items(const rtx_plan *p, const std::vector<int32_t> &cols);
// Rebuild the generic kernels' column, cell and item lists: the columns no
// fused bin takes, plus the bins marked generic.
int rebuild_generic(rtx_plan *p);
// Limits of the fused per-column kernel: 0 when a column with these shapes
// cannot run on it, else 1 (bytes64: its dynamic LDS per part in FP64).
int fused_column_ok(int ncells, int nsp, int nang, int part_links, size_t bytes64);
size_t fused_bytes(int64_t col_dofs, int64_t nang, int64_t part_dofs, int64_t part_trace,
                   size_t a);

// Column kernel of uneven-order hp meshes (one workgroup per column):
// column-local offsets of each link's angle-contracted trace block and
// the column totals.
constexpr int HP_MAX_CELLS = 64;    // cells per column on this path
constexpr int HP_MAX_LINKS = 192;   // upwind links per column on this path
constexpr int HP_NCLS = 2;          // LDS classes (one launch each)
struct HpTables {
    const int32_t *link_inn = nullptr, *col_inn = nullptr;
    // per cell: column-local offset of its inflow-face sums; per column:
    // their count
    const int32_t *cell_fo = nullptr, *col_ffc = nullptr;
    // E^s staging: per column, its distinct E^s blocks copied contiguously
    // into es_pool from col_es0 (col_nes values); per link, the column-local
    // offset of its block in that copy
    const int32_t *link_esl = nullptr, *col_es0 = nullptr, *col_nes = nullptr;
    const double *es_pool = nullptr;
    // per column a header of HP_HDR ints (first cell, cells, nodes per
    // direction, angles, first DOF, first link, links, first angle, the
    // inn / face-sum / trace / E^s counts, E^s start, first spatial node), so
    // the kernel reads a column's scalars in one step; per cell the mask of
    // faces that carry links; per link face | (neighbour face nodes << 2)
    const int32_t *col_hdr = nullptr, *cell_lf = nullptr, *link_w = nullptr;
};
constexpr int HP_HDR = 16;
struct HpClass {
    int n = 0;                      // columns
    const int32_t *cols = nullptr;  // device list (key order)
    size_t bytes[3] = {0, 0, 0};    // dynamic LDS per workgroup: FP64, FP32, mixed
};
int hpcol_build(rtx_plan *p, const rtx_mesh_desc *d);
int launch_hpcol(rtx_plan *p, uint32_t ops, int prec, const void *dx, void *dy, hipStream_t st,
                 int accumulate);

constexpr int RTX_NBINS = 6;
// Tiled scattering work items: 16 spatial nodes x SC_RG output angles
constexpr int SC_RG = 128;
constexpr int SC_XCD = 8;

// One order bin of the fused kernel: columns that run with the same number
// of parts H, and the bin's shape maxima for every candidate H (index 0, 1,
// 2 for H = 1, 2, 4) so the part count can be re-chosen (autotuned).
struct FusedBin {
    int H = 0;                      // parts in use
    int n = 0;                      // columns
    const int32_t *cols = nullptr;  // device list of column indices
    // largest per-column dynamic LDS of a part, [0]: FP64, [1]: FP32 storage
    size_t bytes[2][3] = {{0, 0, 0}, {0, 0, 0}};
    int part_links[3] = {0, 0, 0};
    int min_cells = 0;
    // the same columns on the generic kernels (cells, scattering items), and
    // whether the apply runs them there (set by the autotuner)
    const int32_t *cells = nullptr, *items = nullptr;
    int ncells = 0, nitems = 0;
    bool generic = false;
};
}  // namespace rtx

struct rtx_plan {
    int device = 0;
    int n_cu = 0;                // compute units of the device
    rtx::DevMesh m{};
    void *arena = nullptr;       // one device allocation for all tables
    size_t arena_bytes = 0;
    void *coef = nullptr;        // kappa, sigma, G (+ reduced-precision copies)
    size_t coef_bytes = 0;
    int64_t nsp = 0, nG = 0;
    int32_t max_nang = 0, max_cell_ndof = 0, max_nsp = 0;
    int32_t max_col_dofs = 0, max_col_cells = 0, max_col_links = 0, max_col_trace = 0;
    int32_t min_col_cells = 0;
    // uniform bin: every column nx = ny = uni_n, every cell nth = uni_nth,
    // uni_ncells cells per column, one shared angular mesh (G block)
    bool uni_ok = false;
    int32_t uni_n = 0, uni_nth = 0, uni_ncells = 0;
    int uni_H = 0;                // parts per column of the uniform kernel (0: default)
    // largest own-DOF count and own-trace count of a column part when the
    // fused kernel splits every column into H = 1, 2, 4 cell ranges
    int32_t max_part_dofs[5] = {0, 0, 0, 0, 0}, max_part_trace[5] = {0, 0, 0, 0, 0};
    int32_t max_part_links[5] = {0, 0, 0, 0, 0};
    int fused_mode = -1;   // -1 auto, 0 generic kernels, 1 fused per-column kernel
    // order bins of the apply (non-uniform meshes): fused-kernel bins by part
    // count, and the columns (and their cells) left to the generic kernels
    // bins b = 3 * cls + h: cls 0 holds columns whose part fits 32 KB of LDS
    // at some H, cls 1 the larger ones; h the part count index (H = 1, 2, 4)
    rtx::FusedBin fbin[rtx::RTX_NBINS];
    const int32_t *gen_cols = nullptr, *gen_cells = nullptr;
    int n_gen_cols = 0, n_gen_cells = 0;
    // work items of the tiled scattering kernel (pairs: column, tile word),
    // over every column and over the generic columns, XCD-interleaved
    const int32_t *sc_all = nullptr, *sc_gen = nullptr;
    int n_sc_all = 0, n_sc_gen = 0;
    void *gen_alloc = nullptr;   // device lists of the generic set
    // column kernel of hp meshes: per-link tables, LDS classes, and the
    // columns it cannot take (run on the generic kernels, lists in hp_alloc)
    bool hp_ok = false;
    bool hp_default = false;     // mode -1 uses the column kernel
    rtx::HpTables hp;
    rtx::HpClass hp_cls[rtx::HP_NCLS];
    std::vector<int32_t> hp_over_cols;
    const int32_t *hp_gen_cells = nullptr, *hp_gen_items = nullptr;
    int hp_n_gen_cells = 0, hp_n_gen_items = 0;
    void *hp_alloc = nullptr, *hp_es = nullptr;
    hipStream_t hp_side = nullptr;
    hipEvent_t hp_fork = nullptr, hp_join = nullptr;
    // host copies: per-column cell start (ncol + 1), spatial nodes, angles;
    // fused-bin columns and the columns no bin takes
    std::vector<int32_t> h_col_cell_start, h_col_nsp, h_col_nang;
    std::vector<int32_t> h_bin_cols[rtx::RTX_NBINS], h_over_cols;
    void *bins = nullptr;
    bool has_coeffs = false;
    // block-Jacobi
    int64_t *bj_off = nullptr;   // ncell + 1 offsets of the column-major inverses
    int64_t bj_entries = 0;
    // packed BF16 copy: columns padded to a multiple of 8 rows (16 bytes)
    int64_t *bj_off16 = nullptr;
    int64_t bj_entries16 = 0;
    // FP8 copy: codes with columns padded to a multiple of 16 rows, then an
    // FP32 scale per column; per block the code offset (bytes) and the
    // scale offset (floats from the start of the scales)
    int64_t *bj_off8 = nullptr, *bj_soff = nullptr;
    int64_t bj_codes8 = 0, bj_bytes8 = 0;
    // FP4 copy: the FP8 layout at half the bytes (two codes per byte, block
    // b's codes at bj_off8[b] / 2), then the same FP32 scales
    int64_t bj_codes4 = 0, bj_bytes4 = 0;
    bool bj_ready = false;
    int bj_method = -1;          // -1 automatic, 0 LDS Gauss-Jordan, 1 register sweeps
    // blocks: (cell, first angle, angle count) triples; a cell of more than
    // BJ_SPLIT DOFs is split into groups of angles unless bj_method == 0
    int32_t *bj_blk = nullptr;
    // per block, for the applies: (n, first DOF, cell angle count, angle
    // count g), one 16-byte load instead of a chain through the mesh tables
    int32_t *bj_apd = nullptr;
    // the register build's size classes (blocks of <= 32, 64, 96, 128
    // DOFs): block indices, concatenated, with the class starts
    int32_t *bj_cls = nullptr;
    int32_t bj_cls_start[5] = {0, 0, 0, 0, 0};
    int32_t bj_nblk = 0, bj_max_block = 0;
    int bj_split = -1;           // split setting the block table was built with
    // coarse grid of the learned correction: per column a list of
    // (coarse cell, area(col & cell) / area(cell), area(col & cell) / area(col))
    void *coarse = nullptr;      // one allocation for the map
    int32_t coarse_n = 0;        // coarse cells
    int32_t coarse_nmap = 0;     // map entries
    const int32_t *cmap_start = nullptr, *cmap_cell = nullptr;
    const double *cmap_fcell = nullptr, *cmap_fcol = nullptr;
};
