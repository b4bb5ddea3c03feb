"""ctypes binding of librtx.so (the C ABI declared in include/rtx.h).

The library is built in-tree by ``csrc/build.py``.  Loading fails loudly when
it is missing: there is no CPU fallback for any device operation.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "librtx.so")

OP_MASS, OP_SCAT, OP_ICONV, OP_BCONV, OP_RESTRICT, OP_ACCUM, OP_QUADW = 1, 2, 4, 8, 16, 32, 64
OP_SYSTEM = OP_MASS | OP_SCAT | OP_ICONV | OP_BCONV
PREC_F64, PREC_F32, PREC_MIXED, PREC_FP8, PREC_FP4 = 0, 1, 2, 3, 4
NMAX = 10

_i32p = C.POINTER(C.c_int32)
_i64p = C.POINTER(C.c_int64)
_f64p = C.POINTER(C.c_double)


class MeshDesc(C.Structure):
    _fields_ = [
        ("ncol", C.c_int32), ("ncell", C.c_int32), ("nlink", C.c_int32),
        ("ndof", C.c_int64),
        ("col_cell_start", _i32p), ("col_nx", _i32p), ("col_ny", _i32p),
        ("col_sp_off", _i32p), ("col_nang", _i32p), ("col_G_off", _i64p),
        ("col_geom", _f64p),
        ("cell_dof_off", _i32p), ("cell_col", _i32p), ("cell_nth", _i32p),
        ("cell_quad", _i32p), ("cell_ang_off", _i32p), ("cell_bdry", _i32p),
        ("cell_geom", _f64p),
        ("link_start", _i32p), ("link_src", _i32p), ("link_face", _i32p),
        ("link_Es_off", _i32p), ("link_Eth_off", _i32p), ("link_target", _i32p),
        ("col_ang_start", _i32p), ("ang_pack", _i32p), ("ang_trig", _f64p),
        ("link_meta", _i32p), ("col_ntrace", _i32p), ("col_nitems", _i32p),
        ("nang", C.c_int64),
        ("nEs", C.c_int64), ("nEth", C.c_int64),
        ("Es_pool", _f64p), ("Eth_pool", _f64p),
        ("lgl_x", _f64p), ("lgl_w", _f64p), ("lg_x", _f64p), ("lg_w", _f64p),
        ("lgl_D", _f64p),
    ]


class RtxError(RuntimeError):
    pass


_lib = None

# (name, restype, argtypes) of every exported entry point.
SIGNATURES = [
    ("rtx_last_error", C.c_char_p, []),
    ("rtx_mesh_desc_layout", C.c_int, [C.POINTER(C.c_int64), C.c_int]),
    ("rtx_plan_create", C.c_int, [C.POINTER(MeshDesc), C.POINTER(C.c_void_p)]),
    ("rtx_plan_destroy", None, [C.c_void_p]),
    ("rtx_plan_set_coeffs", C.c_int, [C.c_void_p, C.c_int64, _f64p, _f64p, C.c_int64, _f64p]),
    ("rtx_apply", C.c_int, [C.c_void_p, C.c_uint32, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]),
    ("rtx_mask", C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_void_p]),
    ("rtx_set_path", C.c_int, [C.c_void_p, C.c_int]),
    ("rtx_autotune", C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_int,
                               C.c_void_p]),
    ("rtx_plan_info", C.c_int, [C.c_void_p, C.POINTER(C.c_int64), C.c_int]),
    ("rtx_set_uniform_parts", C.c_int, [C.c_void_p, C.c_int]),
    ("rtx_set_bin_generic", C.c_int, [C.c_void_p, C.c_int, C.c_int]),
    ("rtx_assemble_entries", C.c_int, [C.c_void_p, C.c_int, C.c_int64] + [C.c_void_p] * 7),
    ("rtx_bjacobi_size", C.c_int, [C.c_void_p, C.POINTER(C.c_int64)]),
    ("rtx_bjacobi_factor", C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p,
                                     C.c_void_p]),
    ("rtx_bjacobi_set_method", C.c_int, [C.c_void_p, C.c_int]),
    ("rtx_bjacobi_size_bf16", C.c_int, [C.c_void_p, C.POINTER(C.c_int64)]),
    ("rtx_bjacobi_size_fp8", C.c_int, [C.c_void_p, C.POINTER(C.c_int64)]),
    ("rtx_bjacobi_pack_fp8", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    ("rtx_bjacobi_size_fp4", C.c_int, [C.c_void_p, C.POINTER(C.c_int64)]),
    ("rtx_bjacobi_pack_fp4", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    ("rtx_bjacobi_pack_bf16", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    ("rtx_bjacobi_apply", C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p,
                                    C.c_void_p]),
    ("rtx_bjacobi_apply_fp8_h", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                          C.c_void_p]),
    ("rtx_bjacobi_apply_h", C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p,
                                      C.c_void_p]),
    ("rtx_arnoldi_workspace", C.c_int, [C.c_int, C.POINTER(C.c_int64)]),
    ("rtx_arnoldi_step", C.c_int, [C.c_void_p, C.c_int64, C.c_int, C.c_int64, C.c_void_p,
                                   C.c_void_p, C.c_void_p, C.c_int, C.c_void_p]),
    ("rtx_mdot", C.c_int, [C.c_void_p, C.c_int64, C.c_int, C.c_int64, C.c_void_p, C.c_void_p,
                           C.c_void_p, C.c_void_p]),
    ("rtx_arnoldi_step_f32", C.c_int, [C.c_void_p, C.c_int64, C.c_int, C.c_int64, C.c_void_p,
                                       C.c_void_p, C.c_void_p, C.c_int, C.c_void_p]),
    ("rtx_combine", C.c_int, [C.c_void_p, C.c_int64, C.c_int, C.c_void_p, C.c_int64,
                              C.c_void_p, C.c_int, C.c_void_p]),
    ("rtx_combine_f32", C.c_int, [C.c_void_p, C.c_int64, C.c_int, C.c_void_p, C.c_int64,
                                  C.c_void_p, C.c_int, C.c_int, C.c_void_p]),
    ("rtx_combine_f16", C.c_int, [C.c_void_p, C.c_int64, C.c_int, C.c_void_p, C.c_int64,
                                  C.c_void_p, C.c_int, C.c_int, C.c_void_p]),
    ("rtx_axpby_mixed", C.c_int, [C.c_int, C.c_int64, C.c_double, C.c_void_p, C.c_double,
                                  C.c_void_p, C.c_void_p]),
    ("rtx_axpby", C.c_int, [C.c_int, C.c_int64, C.c_double, C.c_void_p, C.c_double,
                            C.c_void_p, C.c_void_p]),
    ("rtx_convert", C.c_int, [C.c_int, C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p]),
    ("rtx_error_l2", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                               C.c_void_p, C.c_void_p]),
    ("rtx_error_l2_nodal", C.c_int, [C.c_void_p] * 9),
    ("rtx_copy_f64", C.c_int, [C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p]),
    ("rtx_indicator_moments", C.c_int, [C.c_void_p] * 6),
    ("rtx_error_angular_jump", C.c_int, [C.c_void_p] * 6),
    ("rtx_error_spatial_jump", C.c_int, [C.c_void_p] * 8),
    ("rtx_gmres_state_size", C.c_int, [C.c_int, C.POINTER(C.c_int64)]),
    ("rtx_gmres_begin", C.c_int, [C.c_int, C.c_double, C.c_void_p, C.c_void_p]),
    ("rtx_gmres_givens", C.c_int, [C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]),
    ("rtx_gmres_arnoldi", C.c_int, [C.c_int, C.c_void_p, C.c_int64, C.c_int64, C.c_void_p,
                                    C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int,
                                    C.c_void_p, C.c_void_p]),
    ("rtx_gmres_arnoldi_static", C.c_int, [C.c_int, C.c_void_p, C.c_int64, C.c_int64,
                                           C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int,
                                           C.c_void_p, C.c_void_p]),
    ("rtx_gmres_solve", C.c_int, [C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]),
    ("rtx_graph_capture_begin", C.c_int, [C.c_void_p]),
    ("rtx_graph_capture_end", C.c_int, [C.c_void_p, C.POINTER(C.c_void_p)]),
    ("rtx_graph_launch", C.c_int, [C.c_void_p, C.c_void_p]),
    ("rtx_graph_destroy", C.c_int, [C.c_void_p]),
    # learned coarse correction
    ("rtx_coarse_setup", C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_void_p, C.c_void_p,
                                   C.c_void_p, C.c_void_p]),
    ("rtx_coarse_restrict", C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_void_p,
                                      C.c_void_p]),
    ("rtx_coarse_prolong", C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_int,
                                     C.c_void_p, C.c_void_p]),
    ("rtx_scale_features", C.c_int, [C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    ("rtx_mlp_gemm_workspace", C.c_int, [C.c_int, C.c_int, C.c_int, C.POINTER(C.c_int64)]),
    ("rtx_mlp_gemm", C.c_int, [C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_int64, C.c_int64,
                               C.c_void_p, C.c_int64, C.c_int64, C.c_void_p, C.c_int64, C.c_int,
                               C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p,
                               C.c_int64, C.c_void_p]),
    ("rtx_mlp_colsum", C.c_int, [C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]),
    ("rtx_mlp_gemv_bf16", C.c_int, [C.c_int, C.c_int, C.c_void_p, C.c_int64, C.c_void_p,
                                    C.c_void_p, C.c_int, C.c_void_p, C.c_void_p]),
    ("rtx_mlp_to_bf16", C.c_int, [C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p]),
    ("rtx_adam_step", C.c_int, [C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                C.c_float, C.c_float, C.c_float, C.c_float, C.c_float, C.c_int,
                                C.c_void_p]),
    ("rtx_mse_grad", C.c_int, [C.c_int64, C.c_float, C.c_void_p, C.c_void_p, C.c_void_p,
                               C.c_void_p, C.c_void_p]),
    ("rtx_line_search_workspace", C.c_int, [C.POINTER(C.c_int64)]),
    ("rtx_coarse_line_search", C.c_int, [C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p,
                                         C.c_void_p]),
    ("rtx_line_search", C.c_int, [C.c_int, C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p,
                                  C.c_void_p, C.c_void_p]),
    ("rtx_axpy_dev", C.c_int, [C.c_int, C.c_int64, C.c_void_p, C.c_double, C.c_void_p,
                               C.c_void_p, C.c_void_p]),
]


def load():
    """Load librtx.so once; raise if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RtxError("librtx.so not found at {} - build it with "
                       "radiative-transfer-amr_amd/csrc/build.py".format(LIB_PATH))
    lib = C.CDLL(LIB_PATH)
    for name, res, args in SIGNATURES:
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(status: int, what: str = "") -> None:
    if status != 0:
        msg = load().rtx_last_error().decode(errors="replace")
        raise RtxError("{} failed ({}): {}".format(what or "rtx call", status, msg))


def ptr(arr, ctype):
    """ctypes pointer to a contiguous numpy array (kept alive by the caller)."""
    return arr.ctypes.data_as(C.POINTER(ctype))
