/*
 * rtx - MI355X-native DG transport operator and Krylov kernels (C ABI).
 *
 * The reference (jasonltorchinsky/radiative-transfer-amr) is pure Python: its
 * "boundary" for this path is the method surface of rt.Problem
 * (src/rt/class_Problem.py:11-32) which assembles PETSc matrices and calls
 * PETSc KSP (src/rt/solve.py:81-117).  Each entry point below replaces one
 * piece of that path; the Python host package (rt.Problem, rtx.*) binds them
 * with ctypes.  All pointers named d_* are device pointers; everything else
 * is host memory that is copied during the call.  Work is enqueued on the
 * caller's stream (a hipStream_t passed as void*).  Return value: 0 on
 * success, a negative rtx_status on failure, with a message available from
 * rtx_last_error() (thread-local).
 */
#ifndef RTX_H
#define RTX_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    RTX_OK = 0,
    RTX_ERR_ARG = -1,      /* bad argument / size out of range          */
    RTX_ERR_HIP = -2,      /* a HIP runtime call failed                  */
    RTX_ERR_STATE = -3,    /* call out of order (e.g. no coefficients)   */
    RTX_ERR_NUMERIC = -4   /* singular block, non-finite values          */
} rtx_status;

/* Operator parts for rtx_apply (bit mask). */
enum {
    RTX_OP_MASS = 1u,      /* M      : src/rt/mass_matrix.py:140-173                 */
    RTX_OP_SCAT = 2u,      /* -S     : src/rt/scattering_matrix.py:142-237 (returns -S) */
    RTX_OP_ICONV = 4u,     /* -IC    : src/rt/interior_convection_matrix.py:161-243  */
    RTX_OP_BCONV = 8u,     /* +BC    : src/rt/boundary_convection_matrix.py:258-537  */
    RTX_OP_RESTRICT = 16u, /* zero x and y on inflow-boundary DOFs: P_I A P_I^T,
                              src/dg/matrix/get_masks.py:80-110, split_matrix.py:46 */
    RTX_OP_ACCUM = 32u,    /* y += op(x) instead of y = op(x)                      */
    RTX_OP_QUADW = 64u     /* quadrature weights (dx dy dth / 8) w_i w_j w_a: the
                              forcing-vector weighting, src/rt/forcing_vector.py:141-153 */
};
#define RTX_OP_SYSTEM (RTX_OP_MASS | RTX_OP_SCAT | RTX_OP_ICONV | RTX_OP_BCONV)

/* Arithmetic for rtx_apply. */
enum {
    RTX_PREC_F64 = 0,      /* everything in FP64                                    */
    RTX_PREC_F32 = 1,      /* x, y FP32; FP32 arithmetic                            */
    RTX_PREC_MIXED = 2,    /* x, y FP32; convection FP32, scattering BF16 on MFMA
                              with FP32 accumulation                                */
    RTX_PREC_FP8 = 3,      /* block-Jacobi apply only: FP8 (e4m3) inverses with an
                              FP32 scale per column, FP32 vectors and arithmetic    */
    RTX_PREC_FP4 = 4       /* block-Jacobi apply only: FP4 (e2m1) inverses with an
                              FP32 scale per column, FP32 vectors and arithmetic    */
};

/* Largest supported number of nodes per direction in a cell. */
#define RTX_NMAX 10

/*
 * Flattened hp mesh.  Columns are in ascending key order and cells in
 * ascending key order within a column, so the DOF order is the reference's
 * (src/dg/matrix/get_idxs.py:11-80, src/rt/mass_matrix.py:54-74).
 * Inflow face couplings ("links") are those of
 * src/rt/boundary_convection_matrix.py:348-524, one per (cell, inflow face,
 * neighbour cell), sorted by target cell.
 */
typedef struct {
    int32_t ncol, ncell, nlink;
    int64_t ndof;
    /* columns */
    const int32_t *col_cell_start;  /* ncol + 1                                   */
    const int32_t *col_nx, *col_ny; /* LGL orders                                 */
    const int32_t *col_sp_off;      /* first spatial node (kappa/sigma arrays)     */
    const int32_t *col_nang;        /* angular nodes of the column (sum of nth)   */
    const int64_t *col_G_off;       /* offset of the column's G^T in the G pool    */
    const double *col_geom;         /* 4 per column: x0, y0, x1, y1                */
    /* cells */
    const int32_t *cell_dof_off;    /* first DOF of the cell                       */
    const int32_t *cell_col;        /* owning column                               */
    const int32_t *cell_nth;        /* Gauss orders in theta                       */
    const int32_t *cell_quad;       /* quadrant 0..3, -1 when not in one           */
    const int32_t *cell_ang_off;    /* first angular index within the column       */
    const int32_t *cell_bdry;       /* bit F: face F is a domain face whose
                                       directions flow in (masked DOFs)            */
    const double *cell_geom;        /* 2 per cell: th0, th1                        */
    /* inflow links */
    const int32_t *link_start;      /* ncell + 1                                   */
    const int32_t *link_src;        /* source (upwind) cell                        */
    const int32_t *link_face;       /* face F of the target cell                   */
    const int32_t *link_Es_off;     /* E^x or E^y (n1 x n0, row-major) in Es pool  */
    const int32_t *link_Eth_off;    /* E^theta (nth1 x nth0, row-major)            */
    const int32_t *link_target;     /* target cell of each link                    */
    /* Packed per-column metadata for the fused per-column kernel (derived from
       the tables above; checked against them in rtx_plan_create):
       col_ang_start[ncol + 1]: first global angular node of each column
         (angular nodes numbered cell by cell in DOF order);
       ang_pack[nang]: column-local first DOF of the node's cell | nth << 20 |
         local index << 25;
       ang_trig[3 * nang]: cos and sin of the node's physical angle and its
         Gauss weight;
       link_meta[4 * nlink]: source cell's first DOF;
         F | nx1 << 2 | ny1 << 6 | nth1 << 10 | masked faces of the source << 14 |
         column-local target cell << 18;
         column-local offset of the link's trace (nx1 or ny1 times nth1 values);
         column-local offset of the link's face nodes (nx or ny times nth);
       col_ntrace[ncol], col_nitems[ncol]: totals of the last two per column. */
    const int32_t *col_ang_start, *ang_pack;
    const double *ang_trig;
    const int32_t *link_meta, *col_ntrace, *col_nitems;
    int64_t nang;
    int64_t nEs, nEth;
    const double *Es_pool, *Eth_pool;
    /* 1-D rules: for n = 1..RTX_NMAX, index (n * RTX_NMAX + k):
       lgl_x, lgl_w, lg_x, lg_w; and lgl_D[(n * RTX_NMAX + q) * RTX_NMAX + j]
       = l_q'(xi_j) */
    const double *lgl_x, *lgl_w, *lg_x, *lg_w, *lgl_D;
} rtx_mesh_desc;

typedef struct rtx_plan rtx_plan;

const char *rtx_last_error(void);

/* Byte offsets of the rtx_mesh_desc fields in declaration order, followed by
   sizeof(rtx_mesh_desc); writes min(n, count) values and returns count.
   Lets a binding verify its mirror of the struct. */
int rtx_mesh_desc_layout(int64_t *offsets, int n);

/* Copy the flattened mesh to the device (replaces the per-call assembly set-up
   of every src/rt/*_matrix.py).  Uses the current HIP device. */
int rtx_plan_create(const rtx_mesh_desc *desc, rtx_plan **out);
void rtx_plan_destroy(rtx_plan *plan);

/* Nodal coefficients (host arrays).  kappa, sigma: one value per spatial node
   in column order (node (i, j) of column c at col_sp_off[c] + i * ny + j),
   the callables of Problem (src/rt/class_Problem.py:12-23) evaluated there.
   G_pool: per column angular mesh, G^T[A, R] = (dth_K / 8)(dth_K' / 2)
   w_r w_a Phi(th_R, th_A) with R in cell K, A in cell K'
   (src/rt/scattering_matrix.py:160-205). */
int rtx_plan_set_coeffs(rtx_plan *plan, int64_t nsp, const double *kappa,
                        const double *sigma, int64_t nG, const double *G_pool);

/* y = op(x) (or y += op(x) with RTX_OP_ACCUM) for the operator parts in
   op_mask: the matrix-free counterpart of PETSc MatMult with the assembled
   (BC - IC) + (M + S_ret) of src/rt/solve.py:52. */
int rtx_apply(rtx_plan *plan, uint32_t op_mask, int precision, const void *d_x,
              void *d_y, void *stream);

/* Apply path: -1 automatic (the uniform-bin kernel when the whole mesh is
   one compiled bin, else the hp column kernel on the columns it takes and the
   generic kernels on the rest, unless rtx_autotune found the order bins or
   the generic kernels faster), 0 generic per-cell / per-column kernels,
   1 fused per-part kernel, 2 uniform-bin kernel, 3 hp column kernel (one
   workgroup per column; quadrature weights not supported) - rtx_apply fails
   if the mesh does not fit a forced path. */
int rtx_set_path(rtx_plan *plan, int mode);

/* Autotune the apply's launch shapes on this plan: for the uniform-bin kernel
   and every fused order bin (columns grouped by the number of parts H their
   LDS needs), time each valid H in {1, 2, 4} with `reps` launches of the
   interior operator on d_x -> d_y (device vectors of `precision`) and keep
   the fastest; a bin whose columns run faster on the generic kernels moves
   there.  Synchronises the stream.  No counterpart in the reference (its
   assembly has no launch shapes). */
int rtx_autotune(rtx_plan *plan, int precision, const void *d_x, void *d_y, int reps,
                 void *stream);
/* Uniform-bin kernel: parts = 1, 2 or 4 (where valid for the plan's shape)
   forces the line-pass kernel with that many parts per column; 0 restores
   the default: in FP64 the output-layout kernel (one column per workgroup,
   every term of an output added where the scattering MFMA leaves it) when
   all columns are resident at once (four per CU), else the line-pass kernel
   with one part; in FP32 and mixed the line-pass kernel with one part.  No
   counterpart in the reference. */
int rtx_set_uniform_parts(rtx_plan *plan, int parts);
/* Run fused order bin `bin` (0..5) on the generic kernels (generic = 1) or
   on the fused kernel (0); rebuilds the generic lists (synchronises the
   device).  No counterpart in the reference. */
int rtx_set_bin_generic(rtx_plan *plan, int bin, int generic);
/* Launch-shape summary (25 values): uniform bin usable, its parts per column;
   parts, columns and generic flag of the six fused order bins (part count by
   LDS class); columns and cells on the generic kernels; whether the hp
   column kernel is the default path, and the columns it takes; whether the
   uniform bin's FP64 apply runs the output-layout kernel.  Returns the
   count. */
int rtx_plan_info(const rtx_plan *plan, int64_t *out, int n);

/* In-place mask of a full-length vector: keep = 0 zeroes the inflow-boundary
   DOFs (keeps the interior, split_vector.py:39-46), keep = 1 zeroes the
   interior DOFs (keeps the boundary part). */
int rtx_mask(rtx_plan *plan, int keep, int precision, void *d_x, void *stream);

/* Values of one operator part at given positions (the assembled-matrix
   methods of src/rt/*.py).  part: 0 mass (mass_matrix.py), 1 returned
   scattering -S (scattering_matrix.py), 2 interior convection IC
   (interior_convection_matrix.py), 3 boundary convection BC
   (boundary_convection_matrix.py), 4 M - S (preconditioner_matrix.py).
   Entry e sits at (row cell d_kr[e], row local DOF d_lr[e]) x (column cell
   d_kc[e], column local DOF d_lc[e]); for BC, d_link[e] is the upwind link
   or -1 for the outflow own-trace term.  The reference's drop thresholds
   are applied (dropped entries are written as 0). */
int rtx_assemble_entries(rtx_plan *plan, int part, int64_t n, const int32_t *d_kr,
                         const int32_t *d_lr, const int32_t *d_kc, const int32_t *d_lc,
                         const int32_t *d_link, double *d_val, void *stream);

/* ---- per-cell block-Jacobi (PETSc PCBJACOBI of src/rt/solve.py:93-95; the
   reference's discarded dense column inverse, preconditioner_matrix.py:259-262,
   is the closest dense analogue).  Blocks: each cell's diagonal block of the
   interior system (identity on inflow-boundary DOFs when restrict_interior). */

/* Number of doubles of all cell inverses (sum of n_K^2). */
int rtx_bjacobi_size(rtx_plan *plan, int64_t *n_entries);
/* Assemble and invert every cell block (FP64, Gauss-Jordan with partial
   pivoting) into d_inv (column-major per cell); optional FP32 copy into
   d_inv_f32.  With d_inv null the blocks are eliminated in FP32 straight into
   d_inv_f32 (register build only: blocks of up to 128 DOFs, else
   RTX_ERR_ARG), for callers that keep only reduced-precision copies.
   d_scratch (same size as d_inv) is needed only when a block does not fit
   LDS.  Synchronises the stream; RTX_ERR_NUMERIC on a singular block. */
int rtx_bjacobi_factor(rtx_plan *plan, int restrict_interior, double *d_inv, float *d_inv_f32,
                       double *d_scratch, void *stream);
/* Build method: -1 automatic (register-resident sweeps for blocks of up to
   128 DOFs, else the LDS / scratch Gauss-Jordan), 0 LDS / scratch
   Gauss-Jordan with explicit row interchanges, 1 register-resident sweeps
   (fails for larger blocks). */
int rtx_bjacobi_set_method(rtx_plan *plan, int method);
/* Entries (uint16) of the packed BF16 inverse pool: every block's columns
   padded to a multiple of 8 rows so each starts on 16 bytes. */
int rtx_bjacobi_size_bf16(rtx_plan *plan, int64_t *n_entries);
/* Round the FP32 inverses (from rtx_bjacobi_factor) to BF16, nearest even,
   into the packed pool. */
int rtx_bjacobi_pack_bf16(rtx_plan *plan, const float *d_inv32, uint16_t *d_inv16,
                          void *stream);
/* Bytes of the FP8 inverse pool: e4m3 codes with every block's columns
   padded to a multiple of 16 rows (16-byte columns), then one FP32 scale
   per column (at a 16-byte aligned offset). */
int rtx_bjacobi_size_fp8(rtx_plan *plan, int64_t *n_bytes);
/* Quantise the FP32 inverses to FP8 e4m3 (nearest even) with the scale
   max |column| / 448 per column, into the FP8 pool.  No counterpart in the
   reference. */
int rtx_bjacobi_pack_fp8(rtx_plan *plan, const float *d_inv32, uint8_t *d_inv8, void *stream);
/* Bytes of the FP4 inverse pool: the FP8 pool's layout at half the bytes
   (two e2m1 codes per byte, the even row in the low nibble; every column
   8-byte aligned), then one FP32 scale per column (at a 16-byte aligned
   offset). */
int rtx_bjacobi_size_fp4(rtx_plan *plan, int64_t *n_bytes);
/* Quantise the FP32 inverses to FP4 e2m1 (nearest, ties to the even code,
   saturating) with the scale max |column| / 6 per column, into the FP4
   pool.  No counterpart in the reference. */
int rtx_bjacobi_pack_fp4(rtx_plan *plan, const float *d_inv32, uint8_t *d_inv4, void *stream);
/* z = blockdiag(inv) r, precision RTX_PREC_F64 (double inverses, vectors),
   RTX_PREC_F32 (float inverses, vectors), RTX_PREC_MIXED (the packed BF16
   pool of rtx_bjacobi_pack_bf16, float vectors, FP32 accumulation),
   RTX_PREC_FP8 (the pool of rtx_bjacobi_pack_fp8, float vectors) or
   RTX_PREC_FP4 (the pool of rtx_bjacobi_pack_fp4, float vectors); the
   factorisation itself is always FP64. */
int rtx_bjacobi_apply(rtx_plan *plan, int precision, const void *d_inv, const void *d_r,
                      void *d_z, void *stream);
/* rtx_bjacobi_apply(RTX_PREC_FP8 or RTX_PREC_FP4) with d_r in FP16 (IEEE
   binary16) and d_z in FP32: the preconditioner of GMRES on an FP16 basis.
   No counterpart in the reference (PETSc bjacobi, src/rt/solve.py:86-89,
   applies in FP64). */
int rtx_bjacobi_apply_h(rtx_plan *plan, int precision, const void *d_inv, const void *d_r,
                        void *d_z, void *stream);
/* rtx_bjacobi_apply_h(RTX_PREC_FP8, ...). */
int rtx_bjacobi_apply_fp8_h(rtx_plan *plan, const void *d_inv, const void *d_r, void *d_z,
                            void *stream);

/* ---- GMRES building blocks (PETSc KSPGMRES of src/rt/solve.py:81-100). */

/* Workspace (doubles) needed by rtx_arnoldi_step / rtx_mdot. */
int rtx_arnoldi_workspace(int max_vecs, int64_t *n_doubles);
/* Orthogonalise d_w against the nv columns of d_V (length n, stride ld) by
   classical Gram-Schmidt, store w / |w| as column nv of d_V and the
   Hessenberg column (nv + 1 values) in d_hcol.  refine 0: one pass (PETSc's
   default KSPGMRESClassicalGramSchmidtOrthogonalization without refinement,
   which src/rt/solve.py:81-100 relies on); 1: a second pass when |w| dropped
   below 1/sqrt(2) of its value (decided on the device); 2: always two
   passes.  Asynchronous. */
int rtx_arnoldi_step(const double *d_V, int64_t ld, int nv, int64_t n, double *d_w,
                     double *d_hcol, double *d_work, int refine, void *stream);
/* The same step on an FP32 basis d_V / d_w (the inner solves of mixed
   iterative refinement); dot products, norms and d_hcol in FP64.  With
   refine 0 the norm of the projected vector is taken from the dot products
   (|w'|^2 = |w|^2 - sum h_j^2; a relative norm below 1e-6 counts as a
   breakdown: the new column is zero and h_{nv} = 0) and d_w is left
   unprojected.  No counterpart in the reference (PETSc's basis is FP64). */
int rtx_arnoldi_step_f32(const float *d_V, int64_t ld, int nv, int64_t n, float *d_w,
                         double *d_hcol, double *d_work, int refine, void *stream);
/* Arnoldi step j of a GMRES cycle of m steps (nv = j + 1 basis vectors,
   FP64 basis for basis_precision 0, FP32 for 1) fused with its Givens
   update: the same results as rtx_arnoldi_step (or _f32) followed by
   rtx_gmres_givens(j, m, d_hcol, d_state).  basis_precision 2: an FP16
   basis (IEEE binary16) with d_w in FP32, one pass only (refine 0), the
   steps of rtx_arnoldi_step_f32 with the new column rounded to FP16 after
   the normalisation.  No counterpart in the reference (PETSc keeps the
   least-squares problem on the host and its basis in FP64). */
int rtx_gmres_arnoldi(int basis_precision, const void *d_V, int64_t ld, int64_t n, void *d_w,
                      double *d_hcol, double *d_work, int refine, int j, int m,
                      double *d_state, void *stream);
/* rtx_gmres_arnoldi without the workspace argument: the same kernels and
   results, with the reduction partials, norms and re-orthogonalisation flag
   in a static device array of the library (2.1 MB per device, held until
   the process ends).  No counterpart in the reference: it is for callers
   that cannot allocate device memory (code under HIP graph capture, serving
   loops).  Limit: m < 255 (RTX_ERR_ARG beyond).  Allocates nothing; calls
   that overlap on one device share the array and race. */
int rtx_gmres_arnoldi_static(int basis_precision, const void *d_V, int64_t ld, int64_t n,
                             void *d_w, double *d_hcol, int refine, int j, int m,
                             double *d_state, void *stream);
/* HIP graphs of a stream's work (the GMRES iterations of a chunk are
   captured once and replayed).  No counterpart in the reference.
   capture_begin: start capturing the calls this thread enqueues on `stream`
   (a created stream, not the legacy default one); capture_end: stop and
   instantiate into *graph_exec; launch: replay on `stream`; destroy: free. */
int rtx_graph_capture_begin(void *stream);
int rtx_graph_capture_end(void *stream, void **graph_exec);
int rtx_graph_launch(void *graph_exec, void *stream);
int rtx_graph_destroy(void *graph_exec);
/* Least-squares state of one restart cycle of m steps kept on the device
   (Givens QR of the Hessenberg matrix; PETSc keeps it on the host, here it
   stays on the GPU so whole chunks of iterations can run as one HIP graph).
   Layout (doubles): cs[m], sn[m], g[m+1], R[m*m] (column j at R + j*m),
   H[(m+1)*m] (raw Hessenberg, column j at H + j*(m+1)), est[m]. */
int rtx_gmres_state_size(int m, int64_t *n_doubles);
/* g = (beta, 0, ..., 0). */
int rtx_gmres_begin(int m, double beta, double *d_state, void *stream);
/* Fold Hessenberg column j (d_hcol[0..j+1], as written by rtx_arnoldi_step)
   into the factorisation; est[j] = residual norm after j + 1 steps. */
int rtx_gmres_givens(int j, int m, const double *d_hcol, double *d_state, void *stream);
/* d_y[0..k-1] = minimiser over the first k directions (back substitution). */
int rtx_gmres_solve(int k, int m, const double *d_state, double *d_y, void *stream);
/* d_out[j] = V_j . w, j < nv. */
int rtx_mdot(const double *d_V, int64_t ld, int nv, int64_t n, const double *d_w,
             double *d_out, double *d_work, void *stream);
/* d_x = (accumulate ? d_x : 0) + sum_j y_j V_j, y on the device. */
int rtx_combine(const double *d_V, int64_t ld, int nv, const double *d_y, int64_t n,
                double *d_x, int accumulate, void *stream);
/* d_x (+)= sum_j y_j V_j for an FP32 basis; d_x FP64 (x_precision 0) or
   FP32 (1); accumulated in FP64.  No counterpart in the reference. */
int rtx_combine_f32(const float *d_V, int64_t ld, int nv, const double *d_y, int64_t n,
                    void *d_x, int x_precision, int accumulate, void *stream);
/* The same for an FP16 basis (IEEE binary16).  No counterpart in the
   reference. */
int rtx_combine_f16(const void *d_V, int64_t ld, int nv, const double *d_y, int64_t n,
                    void *d_x, int x_precision, int accumulate, void *stream);
/* y = a x + b y (precision RTX_PREC_F64 or RTX_PREC_F32). */
int rtx_axpby(int precision, int64_t n, double a, const void *d_x, double b, void *d_y,
              void *stream);
/* y = a x + b y, x FP32 and y FP64 (direction 0), x FP64 and y FP32 (1), or
   x FP64 and y FP16 (2). */
int rtx_axpby_mixed(int direction, int64_t n, double a, const void *d_x, double b, void *d_y,
                    void *stream);
/* direction 0: FP64 -> FP32, 1: FP32 -> FP64. */
int rtx_convert(int direction, int64_t n, const void *d_src, void *d_dst, void *stream);

/* ---- analytic L2 error (src/amr/Error_Indicator/error/error_analytic.py:22-149:
   interpolation of every cell to n + 3 nodes per direction, weighted squared
   error against the analytic solution at those nodes). */
#define RTX_ERR_TAB_STRIDE (2 * RTX_NMAX * (RTX_NMAX + 3) + 2 * (RTX_NMAX + 3))
/* d_tab: (RTX_NMAX + 1) * RTX_ERR_TAB_STRIDE doubles; for n = 1..RTX_NMAX
   at n * RTX_ERR_TAB_STRIDE: the LGL interpolation matrix n -> n + 3
   (P[i (n + 3) + p] = l_i(xi_p), error_analytic.py:57-80), at
   + RTX_NMAX (RTX_NMAX + 3) the Gauss-Legendre one (:100-111), at
   + 2 RTX_NMAX (RTX_NMAX + 3) the n + 3 LGL weights, then the n + 3 LG
   weights.  d_u_hr: the analytic solution at each cell's (nx + 3)(ny + 3)
   (nth + 3) nodes, (p, q, r) row-major, cell k from d_hr_off[k].
   d_cell_err[k] = (dx dy dth / 8) sum w_p w_q w_r (u - I uh)^2 (:136-137). */
int rtx_error_l2(rtx_plan *plan, const double *d_uh, const double *d_u_hr,
                 const int64_t *d_hr_off, const double *d_tab, double *d_cell_err,
                 void *stream);

/* The same error against the nodal solution of a p-refined mesh with the
   same column and cell keys (error_high_resolution.py:103-203): cell k is
   interpolated to d_hr_dims[3k + 0..2] = (mx, my, mth) nodes, each at most
   RTX_NMAX; its high-resolution values start at d_u_hr + d_hr_off[k].
   d_tab: (RTX_NMAX + 1)(RTX_NMAX + 4) * RTX_ERR_TAB_STRIDE doubles; the
   block of (n, m) at (n (RTX_NMAX + 4) + m) * RTX_ERR_TAB_STRIDE has the
   layout of rtx_error_l2's block of n with m nodes in place of n + 3
   (:122-131, 143-155).  Also d_hr_norm2[k] = (dx dy dth / 8) sum w w w
   u_hr^2, the cell's share of the normalisation (:66-97). */
int rtx_error_l2_nodal(rtx_plan *plan, const double *d_uh, const double *d_u_hr,
                       const int64_t *d_hr_off, const int32_t *d_hr_dims, const double *d_tab,
                       double *d_cell_err, double *d_hr_norm2, void *stream);

/* ---- error indicators without an analytic solution
   (src/amr/Error_Indicator/error/error_angular_jump.py,
   error_spatial_jump.py) and the projection integrals hp steering reads
   (src/dg/projection/cell_intg_xy.py, col_intg_th.py). */

/* d_intg_xy[col_ang_start + cell_ang_off + a] = (dx dy / 4) sum_ij w_i w_j
   u[i, j, a] for every cell (cell_intg_xy.py:13-36);
   d_intg_th[col_sp_off + i ny + j] = sum over the column's cells (dth / 2)
   sum_a w_a u[i, j, a] (col_intg_th.py:13-37); d_norm2 (may be NULL):
   (dx dy / 4)(dth / 2) sum w w w u^2 per cell. */
int rtx_indicator_moments(rtx_plan *plan, const double *d_u, double *d_intg_xy,
                          double *d_intg_th, double *d_norm2, void *stream);

/* Angular-jump error per cell (error_angular_jump.py:26-81):
   d_cell_err[k] = (1/4) sum_ij w_i w_j (1/2)((bot_k - top_lo)^2 +
   (top_k - bot_up)^2), the traces at theta pulled back to -1 (bot) and +1
   (top); d_nbr[2k], d_nbr[2k + 1]: global cell indices of the lower and
   upper angular neighbours (Cell.nhbr_keys); d_lend[(n 2 + s) RTX_NMAX + a]
   = l_a(2 s - 1) of the n-point Gauss-Legendre basis.  The caller takes the
   square roots and column sums (:69-79). */
int rtx_error_angular_jump(rtx_plan *plan, const double *d_u, const int32_t *d_nbr,
                           const double *d_lend, double *d_cell_err, void *stream);

/* Spatial-jump error per column (error_spatial_jump.py:30-227).  The pairs
   (own cell, neighbour cell) of column c across each face are
   d_pair[RTX_JUMP_PAIR_INTS p ..] for p in d_col_pair_start[c] ..
   d_col_pair_start[c + 1] - 1: own cell, neighbour cell, face F, nB, nH,
   and d_pool offsets of the row-major projections psi_0 (n_0 x nB), psi_1
   (n_1 x nB), xsi_0 (nth_0 x nH), xsi_1 (nth_1 x nH) (:84-207);
   d_scale[p] = d_A dth_G / 4 (:212, 222).  d_col_err[c] = sum_p d_scale[p]
   sum wB wH (U_0 - U_1)^2 / ((2 dx + 2 dy) 2 pi) (:227). */
#define RTX_JUMP_PAIR_INTS 9
int rtx_error_spatial_jump(rtx_plan *plan, const double *d_u, const int32_t *d_col_pair_start,
                           const int32_t *d_pair, const double *d_scale, const double *d_pool,
                           double *d_col_err, void *stream);

/* ---- measurement.  No counterpart in the reference: d_dst[i] = d_src[i]
   for i < n, one double per lane per iteration with ordinary vector loads
   and stores (a copy of known size for calibrating HBM traffic counters). */
int rtx_copy_f64(const double *d_src, double *d_dst, int64_t n, void *stream);

/* ---- learned coarse scalar-flux correction.  No counterpart in the
   reference: its preconditioners are PETSc bjacobi / kaczmarz
   (src/rt/solve.py:86-95).  A coarse grid of n_coarse cells covers the
   domain; every column maps to the coarse cells it overlaps. */

/* Coarse map (host arrays, copied): entries col_start[c] .. col_start[c+1]-1
   of column c list coarse cells cell[e] with fcell[e] = area(col & cell) /
   area(cell) and fcol[e] = area(col & cell) / area(col). */
int rtx_coarse_setup(rtx_plan *plan, int32_t n_coarse, int32_t nmap, const int32_t *col_start,
                     const int32_t *cell, const double *fcell, const double *fcol);
/* mode 0: d_out[q * n_coarse + g] = sum_c fcol * sum_{beta in c} v_beta t_q(theta_beta),
   t = (1, cos, sin) - angular moments of a weak-form residual (3 n_coarse floats);
   mode 1: d_out[g] = sum_c fcell * quadrature-weighted column mean of v (n_coarse floats).
   v is FP64 (precision RTX_PREC_F64) or FP32. */
int rtx_coarse_restrict(rtx_plan *plan, int mode, int precision, const void *d_v, float *d_out,
                        void *stream);
/* z_beta (=, or += when accumulate) = s * sum_e fcol[e] phi[cell[e]] on the interior DOFs of
   each column and 0 on inflow-boundary DOFs; s = *d_scale (1 when null). */
int rtx_coarse_prolong(rtx_plan *plan, int precision, const float *d_phi, const float *d_scale,
                       int accumulate, void *d_z, void *stream);
/* *d_scale = |m|_2 (1 when m = 0), d_x = m / *d_scale (one workgroup). */
int rtx_scale_features(int64_t n, const float *d_m, float *d_x, float *d_scale, void *stream);

/* MLP kernels (FP32 training on MFMA, BF16 inference). */
/* C[m, n] = sum_k A[m, k] B[k, n] with A[m, k] at d_A[m * sam + k * sak] and
   B[k, n] at d_B[k * sbk + n * sbn]; epilogue 0: C = AB; 1: C = AB + bias[n]
   and, if d_H, H = relu(C) (same ldc); 2: C = AB * (Z[m, n] > 0), Z at ldz.
   d_work (rtx_mlp_gemm_workspace floats) enables split-K for tall K. */
int rtx_mlp_gemm_workspace(int M, int N, int K, int64_t *n_floats);
int rtx_mlp_gemm(int M, int N, int K, const float *d_A, int64_t sam, int64_t sak,
                 const float *d_B, int64_t sbk, int64_t sbn, float *d_C, int64_t ldc,
                 int epilogue, const float *d_bias, float *d_H, const float *d_Z, int64_t ldz,
                 float *d_work, int64_t work_floats, void *stream);
/* d_out[n] = sum_m A[m, n] (A is M x N row-major). */
int rtx_mlp_colsum(int M, int N, const float *d_A, float *d_out, void *stream);
/* y[n] = act(sum_k W[n, k] x[k] + bias[n]), W BF16 with row stride ldw, act = relu if relu. */
int rtx_mlp_gemv_bf16(int N, int K, const uint16_t *d_W, int64_t ldw, const float *d_x,
                      const float *d_bias, int relu, float *d_y, void *stream);
int rtx_mlp_to_bf16(int64_t n, const float *d_src, uint16_t *d_dst, void *stream);
/* AdamW step with bias correction (step >= 1). */
int rtx_adam_step(int64_t n, float *d_p, const float *d_g, float *d_m, float *d_v, float lr,
                  float beta1, float beta2, float eps, float weight_decay, int step, void *stream);
/* d_dout = scale (out - y); *d_loss += sum (out - y)^2 (d_loss may be null). */
int rtx_mse_grad(int64_t n, float scale, const float *d_out, const float *d_y, float *d_dout,
                 double *d_loss, void *stream);
/* d_alpha[0] = max(0, r.q / q.q), d_alpha[1] = r.q, d_alpha[2] = q.q, d_alpha[3] = r.r
   (4 doubles);
   d_work holds rtx_line_search_workspace doubles. */
int rtx_line_search_workspace(int64_t *n_doubles);
int rtx_line_search(int precision, int64_t n, const void *d_r, const void *d_q, double *d_alpha,
                    double *d_work, void *stream);
/* Same four outputs for two coarse moment vectors (n floats each): the step
   minimising |R r - alpha R q|, which is 1 for an exact Galerkin coarse
   solve.  One workgroup. */
int rtx_coarse_line_search(int64_t n, const float *d_mr, const float *d_mq, double *d_alpha,
                           void *stream);
/* y += sign * d_alpha[0] * x (the scalar stays on the device). */
int rtx_axpy_dev(int precision, int64_t n, const double *d_alpha, double sign, const void *d_x,
                 void *d_y, void *stream);

#ifdef __cplusplus
}
#endif
#endif /* RTX_H */
