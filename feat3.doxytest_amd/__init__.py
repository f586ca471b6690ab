"""Python binding of the MI355X LAFEM + Solver hot path.

Loads the in-tree libraries built by build.py:
  lib/libfeat3amd.so       C-ABI of include/feat3amd.h (HIP kernels, gfx950)
  lib/libfeat3amd_host.so  C++ host layer, include/feat3amd_host.h
and declares their signatures for ctypes.  Device buffers are torch tensors
(plumbing only); every computation runs in the HIP library.  There is no
fallback: if the libraries are missing, loading fails.

The directory name contains a dot, so import it by path:
    spec = importlib.util.spec_from_file_location("feat3amd", ".../feat3.doxytest_amd/__init__.py")
"""
import ctypes as C
import os

import numpy as np

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.join(PKG_DIR, "lib")
KERNEL_LIB = os.path.join(LIB_DIR, "libfeat3amd.so")
HOST_LIB = os.path.join(LIB_DIR, "libfeat3amd_host.so")

FEAT_OK = 0
STATUS = {0: "undefined", 1: "progress", 2: "success", 3: "aborted", 4: "diverged",
          5: "max_iter", 6: "stagnated"}
COARSE = {"none": 0, "smoother": 1, "jacobi": 2, "direct": 3}

_vp = C.c_void_p
_u64 = C.c_uint64
_d = C.c_double
_f = C.c_float
_i = C.c_int


class FeatError(RuntimeError):
    pass


class CsrF64(C.Structure):
    _fields_ = [("rows", _u64), ("cols", _u64), ("nnz", _u64), ("val", _vp), ("col", _vp),
                ("rowptr", _vp)]


class CsrF32(C.Structure):
    _fields_ = [("rows", _u64), ("cols", _u64), ("nnz", _u64), ("val", _vp), ("col", _vp),
                ("rowptr", _vp)]


class Matrix(C.Structure):
    """feat_matrix_f64_t / feat_matrix_f32_t (same layout)."""
    _fields_ = [("format", _i), ("index", _i), ("rows", _u64), ("cols", _u64), ("nnz", _u64),
                ("val", _vp), ("col", _vp), ("ptr", _vp), ("rowlen", _vp), ("value_kind", _i),
                ("scale", _f), ("tune", _i), ("slice_words", C.c_uint32)]


FMT_CSR, FMT_SELL, FMT_DLINEAR, FMT_BANDED = 0, 1, 2, 3
IDX_U64, IDX_U32, IDX_D16, IDX_DICT, IDX_DIAG = 0, 1, 2, 3, 4
SELL_NO_DICT, SELL_NO_DIAG, SELL_NO_BANDED = 1, 2, 4
VAL_NATIVE, VAL_F16, VAL_CODE8 = 0, 1, 2
Q1_HOST_ASSEMBLY = 1


class ReduceWs(C.Structure):
    _fields_ = [("partials", _vp), ("counter", _vp)]


class MgPcgCfg(C.Structure):
    _fields_ = [("nu_pre", _i), ("nu_post", _i), ("omega_jac", _d), ("omega_rich", _d),
                ("coarse", _i), ("coarse_omega", _d), ("tol_rel", _d), ("max_iter", _u64),
                ("force_generic", _i), ("csr_only", _i), ("precision", _i), ("no_graphs", _i),
                ("no_autotune", _i), ("no_coarse_fusion", _i)]


class LevelHost(C.Structure):
    _fields_ = [("rows", _u64), ("val", _vp), ("col", _vp), ("rowptr", _vp), ("filt", _vp),
                ("nfilt", _u64), ("pcols", _u64), ("pval", _vp), ("pcol", _vp), ("prowptr", _vp)]


class SolveResult(C.Structure):
    _fields_ = [("status", _i), ("iters", _u64), ("def_init", _d), ("def_final", _d)]


# (name, restype, argtypes) of every entry point of include/feat3amd.h
KERNEL_API = [
    ("feat_last_error", C.c_char_p, []),
    ("feat_handle_create", _i, [C.POINTER(_vp), _i, _vp]),
    ("feat_handle_destroy", _i, [_vp]),
    ("feat_handle_stream", _i, [_vp, C.POINTER(_vp)]),
    ("feat_sync", _i, [_vp]),
    ("feat_stream_create", _i, [C.POINTER(_vp)]),
    ("feat_stream_destroy", _i, [_vp]),
    ("feat_handle_wait", _i, [_vp, _vp]),
    ("feat_graph_begin", _i, [_vp]),
    ("feat_graph_end", _i, [_vp, C.POINTER(_vp)]),
    ("feat_graph_launch", _i, [_vp, _vp]),
    ("feat_graph_destroy", _i, [_vp]),
    ("feat_copy_scalar_to_host", _i, [_vp, _vp, _vp]),
    ("feat_malloc", _i, [C.POINTER(_vp), C.c_size_t]),
    ("feat_free", _i, [_vp]),
    ("feat_host_malloc", _i, [C.POINTER(_vp), C.c_size_t]),
    ("feat_host_free", _i, [_vp]),
    ("feat_copy", _i, [_vp, _vp, _vp, C.c_size_t]),
    ("feat_copy_device", _i, [_vp, _vp, _vp, C.c_size_t]),
    ("feat_stream_read_f64", _i, [_vp, _vp, _u64, _vp]),
    ("feat_set_memory_f64", _i, [_vp, _vp, _d, _u64]),
    ("feat_set_memory_f32", _i, [_vp, _vp, _f, _u64]),
    ("feat_convert_f64_f32", _i, [_vp, _vp, _vp, _u64]),
    ("feat_convert_f32_f64", _i, [_vp, _vp, _vp, _u64]),
    ("feat_convert_f64_f16", _i, [_vp, _vp, _vp, _u64]),
    ("feat_convert_f16_f64", _i, [_vp, _vp, _vp, _u64]),
    ("feat_convert_u64_u32", _i, [_vp, _vp, _vp, _u64]),
    ("feat_csr_apply_f64_u64", _i, [_vp, _vp, _d, _vp, _d, _vp, _vp, _vp, _vp, _u64, _u64, _u64, _i]),
    ("feat_csr_apply_transposed_noalloc_f64_u64", _i, [_vp, _vp, _d, _vp, _d, _vp, _vp, _vp, _vp, _u64, _u64, _u64]),
    ("feat_csr_apply_f64_u32", _i, [_vp, _vp, _d, _vp, _d, _vp, _vp, _vp, _vp, _u64, _u64, _u64, _i]),
    ("feat_csr_apply_f32_u32", _i, [_vp, _vp, _f, _vp, _f, _vp, _vp, _vp, _vp, _u64, _u64, _u64, _i]),
    ("feat_csr_apply_f32_u64", _i, [_vp, _vp, _f, _vp, _f, _vp, _vp, _vp, _vp, _u64, _u64, _u64, _i]),
    ("feat_bcsr_apply_f64_u64", _i, [_vp, _i, _i, _vp, _d, _vp, _d, _vp, _vp, _vp, _vp, _u64, _u64, _u64]),
    ("feat_axpy_f64", _i, [_vp, _vp, _d, _vp, _vp, _u64]),
    ("feat_axpy_f32", _i, [_vp, _vp, _f, _vp, _vp, _u64]),
    ("feat_dot_f64", _i, [_vp, _vp, _vp, _vp, _u64]),
    ("feat_dot_f64_host", _i, [_vp, C.POINTER(_d), _vp, _vp, _u64]),
    ("feat_triple_dot_f64_host", _i, [_vp, C.POINTER(_d), _vp, _vp, _vp, _u64]),
    ("feat_norm2_f64", _i, [_vp, _vp, _vp, _u64]),
    ("feat_norm2_f64_host", _i, [_vp, C.POINTER(_d), _vp, _u64]),
    ("feat_scale_f64", _i, [_vp, _vp, _vp, _d, _u64]),
    ("feat_component_product_f64", _i, [_vp, _vp, _vp, _vp, _u64]),
    ("feat_component_invert_f64", _i, [_vp, _vp, _vp, _d, _u64]),
    ("feat_component_invert_f32", _i, [_vp, _vp, _vp, _f, _u64]),
    ("feat_csr_diag_index_u64", _i, [_vp, _vp, _vp, _vp, _u64]),
    ("feat_csr_extract_diag_f64_u64", _i, [_vp, _vp, _vp, _vp, _vp, _u64]),
    ("feat_unit_filter_rhs_f64_u64", _i, [_vp, _vp, _vp, _vp, _u64]),
    ("feat_unit_filter_def_f64_u64", _i, [_vp, _vp, _vp, _u64]),
    ("feat_unit_filter_mask_u64", _i, [_vp, _vp, _u64, _vp, _u64]),
    ("feat_unit_filter_blocked_rhs_f64_u64", _i, [_vp, _i, _vp, _vp, _vp, _u64, _i]),
    ("feat_unit_filter_blocked_def_f64_u64", _i, [_vp, _i, _vp, _vp, _vp, _u64, _i]),
    ("feat_mean_filter_f64", _i, [_vp, _vp, _vp, _vp, _d, _u64]),
    ("feat_matrix_from_csr_f64", _i, [C.POINTER(CsrF64), C.POINTER(Matrix)]),
    ("feat_matrix_dlinear_f64", _i, [_i, C.c_uint32, _i, C.POINTER(Matrix)]),
    ("feat_matrix_dlinear_f32", _i, [_i, C.c_uint32, _i, C.POINTER(Matrix)]),
    ("feat_sell_create_f64", _i, [_vp, C.POINTER(CsrF64), C.POINTER(Matrix)]),
    ("feat_sell_create_f32", _i, [_vp, C.POINTER(CsrF64), C.POINTER(Matrix)]),
    ("feat_sell_create_ex_f64", _i, [_vp, C.POINTER(CsrF64), _i, C.POINTER(Matrix)]),
    ("feat_sell_create_ex_f32", _i, [_vp, C.POINTER(CsrF64), _i, C.POINTER(Matrix)]),
    ("feat_sell_create_f16", _i, [_vp, C.POINTER(CsrF64), C.POINTER(Matrix)]),
    ("feat_sell_create_ex_f16", _i, [_vp, C.POINTER(CsrF64), _i, C.POINTER(Matrix)]),
    ("feat_sell_destroy_f64", _i, [C.POINTER(Matrix)]),
    ("feat_sell_destroy_f32", _i, [C.POINTER(Matrix)]),
    ("feat_matrix_apply_f64", _i, [_vp, C.POINTER(Matrix), _vp, _d, _vp, _d, _vp]),
    ("feat_q1_device_sizes", _i, [_i, _i, _vp]),
    ("feat_mg_tail_f64", _i, [_vp, _vp, _vp, _vp, _vp]),
    ("feat_mg_tail_f32", _i, [_vp, _vp, _vp, _vp, _vp]),
    ("feat_sor_init_symbolic", _i, [_vp, C.POINTER(CsrF64), _i, C.POINTER(_vp)]),
    ("feat_sor_init_numeric", _i, [_vp, _vp, C.POINTER(CsrF64)]),
    ("feat_sor_colors", _i, [_vp, _i, C.POINTER(_i), _vp, _vp]),
    ("feat_sor_apply", _i, [_vp, _vp, _vp, _vp, _d]),
    ("feat_ssor_forward_apply", _i, [_vp, _vp, _vp, _vp, _d]),
    ("feat_ssor_backward_apply", _i, [_vp, _vp, _vp, _d]),
    ("feat_sor_done_symbolic", _i, [_vp]),
    ("feat_ilu_init_symbolic", _i, [_vp, C.POINTER(CsrF64), C.POINTER(_vp)]),
    ("feat_ilu_init_numeric", _i, [_vp, _vp, C.POINTER(CsrF64)]),
    ("feat_ilu_init_symbolic_p", _i, [_vp, C.POINTER(CsrF64), _i, C.POINTER(_vp)]),
    ("feat_ilu_apply", _i, [_vp, _vp, _vp, _vp]),
    ("feat_ilu_info", _i, [_vp, C.POINTER(_u64), C.POINTER(_u64), C.POINTER(_i), C.POINTER(_i)]),
    ("feat_ilu_factors", _i, [_vp, _vp, _vp, _vp, _vp]),
    ("feat_ilu_done_symbolic", _i, [_vp]),
    ("feat_q1_assemble_device", _i, [_vp, _i, _i] + [_vp] * 10),
    ("feat_stokes_device_sizes", _i, [_i, _vp]),
    ("feat_stokes_assemble_device", _i, [_vp, _i, _vp]),
    ("feat_defect_f64", _i, [_vp, C.POINTER(Matrix), _vp, _vp, _vp, _vp, _vp, C.POINTER(ReduceWs)]),
    ("feat_apply_masked_f64", _i, [_vp, C.POINTER(Matrix), _vp, _vp, _vp, _vp, _vp, _vp, C.POINTER(ReduceWs)]),
    ("feat_prol_add_f64", _i, [_vp, C.POINTER(Matrix), _vp, _vp, _vp]),
    ("feat_rj_smooth_f64", _i, [_vp, C.POINTER(Matrix), _vp, _vp, _i, _d, _vp, _vp, _vp, _vp, _vp, _vp,
                                _vp, _vp, C.POINTER(ReduceWs)]),
    ("feat_defect_first_f64", _i, [_vp, C.POINTER(Matrix), _vp, _vp, _vp, _vp, _vp, _vp, _d, _vp,
                                   C.POINTER(ReduceWs)]),
    ("feat_apply_masked_first_f64", _i, [_vp, C.POINTER(Matrix), _vp, _vp, _vp, _vp, _vp, _d, _vp,
                                         C.POINTER(ReduceWs)]),
    ("feat_rj_smooth_from1_f64", _i, [_vp, C.POINTER(Matrix), _vp, _vp, _i, _d, _vp, _vp, _vp, _vp, _vp, _vp,
                                      _vp, _vp, C.POINTER(ReduceWs)]),
    ("feat_pcg_update_first_f64", _i, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _u64, _vp, _vp, _d, _vp,
                                       C.POINTER(ReduceWs)]),
    ("feat_dense_gemv_f64", _i, [_vp, _i, _vp, _vp, _vp, _vp]),
    ("feat_pcg_update_f64", _i, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _u64, C.POINTER(ReduceWs)]),
    ("feat_pcg_direction_f64", _i, [_vp, _vp, _vp, _vp, _vp, _u64]),
    ("feat_read_scalar_f64", _i, [_vp, _vp, C.POINTER(_d)]),
    ("feat_apply_masked_f32", _i, [_vp, C.POINTER(Matrix), _vp, _vp, _vp, _vp, C.POINTER(ReduceWs)]),
    ("feat_defect_f32", _i, [_vp, C.POINTER(Matrix), _vp, _vp, _vp, _vp, _vp, C.POINTER(ReduceWs)]),
    ("feat_prol_add_f32", _i, [_vp, C.POINTER(Matrix), _vp, _vp, _vp]),
    ("feat_rj_smooth_f32", _i, [_vp, C.POINTER(Matrix), _vp, _vp, _i, _f, _vp, _vp, _vp, _vp, _vp, _vp,
                                C.POINTER(ReduceWs)]),
    ("feat_dense_gemv_f32", _i, [_vp, _i, _vp, _vp, _vp, _vp]),
    ("feat_convert_dot_f32_f64", _i, [_vp, _vp, _vp, _vp, _u64, _vp, C.POINTER(ReduceWs)]),
    ("feat_dense_gemv_mixed", _i, [_vp, _i, _vp, _vp, _vp, _vp]),
    ("feat_tuple_apply_f64", _i, [_vp, _vp, _vp, _vp, _vp, _d, _vp]),
    ("feat_stokes_tail_f64", _i, [_vp, _vp, _vp, _vp]),
    ("feat_tuple_apply_post_f64", _i, [_vp, _vp, _vp, _vp, _vp, _d, _vp, _vp, _vp, _d, _i, _i, _vp, _d, _vp, _i,
                                       _vp]),
    ("feat_stokes_transfer_post_f64", _i, [_vp, _vp, _vp, _vp, _u64, _vp, _vp, _vp, _vp, _d, _i, _vp, _vp]),
    ("feat_mean_filter_passes_f64", _i, [_vp, _vp, _vp, _vp, _d, _u64, _i, _i]),
    ("feat_tuple_norm2_f64", _i, [_vp, _vp, _u64, _u64, _vp, _vp]),
    ("feat_axpy_gated_f64", _i, [_vp, _vp, _d, _vp, _u64, _vp]),
    ("feat_csrsb2_apply_f64_u64", _i, [_vp, _vp, _vp, _vp, _vp, _vp, _u64]),
    ("feat_vanka_numeric_f64", _i, [_vp, _vp, _vp, _vp, _d, _i, _vp, _vp, _vp]),
    ("feat_batch_dot_f64", _i, [_vp, _vp, _vp, _vp, _vp]),
    ("feat_batch_pcg_update_f64", _i, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    ("feat_batch_pcg_direction_f64", _i, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _i]),
    ("feat_batch_copy_scalars_f64", _i, [_vp, C.c_uint32, _vp, _vp, _vp]),
    ("feat_batch_status_f64", _i, [_vp, C.c_uint32, _vp, _vp, _vp, _i]),
    ("feat_batch_gather_f64", _i, [_vp, C.c_uint32, _vp, _vp, _vp, _vp, _vp]),
    ("feat_batch_tail_f64", _i, [_vp, _vp, _vp, _vp]),
    ("feat_batch_dot_partials_f64", _i, [_vp, _vp, _vp, _vp, _vp]),
    ("feat_batch_pcg_update_partials_f64", _i, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    ("feat_batch_status_partials_f64", _i, [_vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    ("feat_batch_pcg_direction_partials_f64", _i, [_vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    ("feat_batch_pre_tiled_f64", _i, [_vp, _vp]),
    ("feat_batch_post_tiled_f64", _i, [_vp, _vp]),
    ("feat_copy_int_to_host", _i, [_vp, _vp, _vp]),
    ("feat_autotune_matrix_f64", _i, [_vp, _vp, _vp, _vp, _i, _vp]),
    ("feat_mlp_to_bf16", _i, [_vp, _vp, _vp]),
    ("feat_mlp_infer", _i, [_vp, _vp, _vp, _vp, _u64, _vp]),
    ("feat_mlp_grad", _i, [_vp, _vp, _vp, _vp, _vp, _u64, _vp, _u64, _vp]),
    ("feat_mlp_adam", _i, [_vp, _vp, _vp, _vp, _vp, _vp, _i, _f, _f, _f, _f]),
    ("feat_mlp_features", _i, [_vp, _u64, _vp, _vp, _vp, _vp, _vp, _vp, _u64, _vp, _vp]),
    ("feat_mlp_smooth_f64", _i, [_vp, _u64, _vp, _vp, _vp, _vp, _vp, _d, _d, _vp]),
    ("feat_mlp_features_full", _i, [_vp, _u64, _vp, _vp, _vp, _vp, _vp, _vp]),
    ("feat_mlp_energy", _i, [_vp, _u64, _i, _d, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    ("feat_mlp_step_f64", _i, [_vp, _u64, _vp, _vp, _vp, _vp, _vp, _vp]),
    ("feat_mlp_direction_f64", _i, [_vp, _u64, _vp, _vp, _vp, _vp, _vp, _d, _d, _vp, _vp, C.POINTER(ReduceWs)]),
]

HOST_API = [
    ("feat_host_last_error", C.c_char_p, []),
    ("feat_q1_sizes", _i, [_i, _i, _vp]),
    ("feat_csr_write_file", _i, [_i, C.c_char_p, _u64, _u64, _vp, _vp, _vp, _i]),
    ("feat_csr_read_file", _i, [_i, C.c_char_p, C.POINTER(_vp), C.POINTER(_u64), C.POINTER(_u64),
                                C.POINTER(_u64)]),
    ("feat_io_fetch_csr", _i, [_vp, _vp, _vp, _vp]),
    ("feat_dv_write_file", _i, [_i, C.c_char_p, _vp, _u64]),
    ("feat_dv_read_file", _i, [_i, C.c_char_p, C.POINTER(_vp), C.POINTER(_u64)]),
    ("feat_io_fetch_dv", _i, [_vp, _vp]),
    ("feat_io_free", _i, [_vp]),
    ("feat_mgpcg_write_matrix", _i, [_vp, _i, _i, C.c_char_p, _i]),
    ("feat_krylov_create", _i, [C.POINTER(_vp), _i, _u64, _vp, _vp, _vp, _vp, _u64, _i, _d, _d, _d, _u64]),
    ("feat_krylov_solve", _i, [_vp, _vp, _vp, C.POINTER(SolveResult), _vp, _u64]),
    ("feat_krylov_precond_apply", _i, [_vp, _vp, _vp]),
    ("feat_krylov_destroy", _i, [_vp]),
    ("feat_q1_assemble", _i, [_i, _i, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    ("feat_q1_rhs_sine", _i, [_i, _i, _vp]),
    ("feat_rhs_random", _i, [_u64, _u64, _d, _d, _vp]),
    ("feat_stokes_sizes", _i, [_i, _vp]),
    ("feat_stokes_assemble", _i, [_i, _vp]),
    ("feat_stokes_create", _i, [C.POINTER(_vp), _i, _i, _vp]),
    ("feat_stokes_solve", _i, [_vp, _vp, _vp, C.POINTER(SolveResult), _vp, _u64]),
    ("feat_stokes_init_numeric", _i, [_vp, C.POINTER(_d)]),
    ("feat_stokes_sizes_of", _i, [_vp, _i, _vp]),
    ("feat_stokes_vanka_copy", _i, [_vp, _i, _vp, _vp, _vp, _vp]),
    ("feat_stokes_apply", _i, [_vp, _i, _vp, _vp, _vp, _d]),
    ("feat_stokes_vanka_apply", _i, [_vp, _i, _vp, _vp, _i]),
    ("feat_stokes_smooth", _i, [_vp, _i, _vp, _vp]),
    ("feat_stokes_filter", _i, [_vp, _i, _vp, _i]),
    ("feat_stokes_transfer", _i, [_vp, _i, _i, _vp, _vp]),
    ("feat_stokes_destroy", _i, [_vp]),
    ("feat_batchsolver_create", _i, [C.POINTER(_vp), _vp]),
    ("feat_batchsolver_layout", _i, [_vp, _vp, _vp]),
    ("feat_batchsolver_solve", _i, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    ("feat_batchsolver_tuning", _i, [_vp, _vp, _vp, _i, C.POINTER(_i)]),
    ("feat_batchsolver_tail_levels", _i, [_vp, C.POINTER(_i), C.POINTER(_i)]),
    ("feat_batchsolver_tiled_levels", _i, [_vp, _vp, _i, C.POINTER(_i)]),
    ("feat_batchsolver_destroy", _i, [_vp]),
    ("feat_batch_levels", _i, [_u64, _i, _i, _i, _vp]),
    ("feat_lognormal_cells", _i, [_u64, _u64, _d, _vp]),
    ("feat_q1_diffusion_assemble", _i, [_i, _i, _vp, _vp, _vp, _vp]),
    ("feat_q1_diffusion3d_assemble", _i, [_i, _i, _vp, _d, _vp, _vp, _vp]),
    ("feat_learned_create", _i, [C.POINTER(_vp), _vp]),
    ("feat_learned_generate", _i, [_vp, _i, C.POINTER(_d)]),
    ("feat_learned_train", _i, [_vp, _i, _u64, _vp]),
    ("feat_learned_solve", _i, [_vp, _i, _i, _vp, _vp, _vp, _u64]),
    ("feat_learned_size", _i, [_vp, C.POINTER(_u64), C.POINTER(_u64)]),
    ("feat_learned_energy", _i, [_vp, C.POINTER(_d)]),
    ("feat_learned_fit", _i, [_vp, _i, _i, _i, _u64, C.POINTER(_i), C.POINTER(_u64), C.POINTER(_u64), _vp]),
    ("feat_learned_params", _i, [_vp, _vp, _vp]),
    ("feat_learned_param_ptrs", _i, [_vp, C.POINTER(_vp), C.POINTER(_vp)]),
    ("feat_learned_destroy", _i, [_vp]),
    ("feat_mgpcg_create", _i, [C.POINTER(_vp), C.POINTER(LevelHost), _i, C.POINTER(MgPcgCfg)]),
    ("feat_mgpcg_create_q1", _i, [C.POINTER(_vp), _i, _i, _i, C.POINTER(MgPcgCfg)]),
    ("feat_mgpcg_create_q1_ex", _i, [C.POINTER(_vp), _i, _i, _i, C.POINTER(MgPcgCfg), _i]),
    ("feat_mgpcg_solve", _i, [_vp, _vp, _vp, C.POINTER(SolveResult), _vp, _u64]),
    ("feat_mgpcg_matrix", _i, [_vp, _i, C.POINTER(_u64), C.POINTER(_u64), C.POINTER(_vp), C.POINTER(_vp),
                               C.POINTER(_vp)]),
    ("feat_mgpcg_coarse_fusion_level", _i, [_vp, C.POINTER(_i)]),
    ("feat_mgpcg_coarse_fusion_grid", _i, [_vp, C.POINTER(_i)]),
    ("feat_mgpcg_set_values", _i, [_vp, _i, _vp]),
    ("feat_mgpcg_destroy", _i, [_vp]),
    ("feat_runtime_init", _i, [_i, _vp]),
    ("feat_runtime_sell_flags", _i, [_i]),
    ("feat_runtime_grid_forms", _i, [_i]),
]


def _declare(lib, api):
    for name, res, args in api:
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args


_LIBS = {}


def libs():
    """(kernel_lib, host_lib); raises if the HIP library is not built."""
    if not _LIBS:
        if not os.path.exists(KERNEL_LIB) or not os.path.exists(HOST_LIB):
            raise FeatError(f"HIP extension not built: run python {PKG_DIR}/build.py")
        k = C.CDLL(KERNEL_LIB, mode=C.RTLD_GLOBAL)
        _declare(k, KERNEL_API)
        h = C.CDLL(HOST_LIB)
        _declare(h, HOST_API)
        _LIBS["k"], _LIBS["h"] = k, h
    return _LIBS["k"], _LIBS["h"]


def check(rc, what=""):
    if rc != FEAT_OK:
        k, h = libs()
        msg = (k.feat_last_error() or b"").decode() or (h.feat_host_last_error() or b"").decode()
        raise FeatError(f"{what} failed ({rc}): {msg}")


def ptr(t):
    """Device pointer of a torch tensor (or None)."""
    return None if t is None else C.c_void_p(t.data_ptr())


class Handle:
    """feat_handle_t on the current torch stream of `device`."""

    def __init__(self, device=0, stream=None):
        k, _ = libs()
        self.k = k
        self.h = C.c_void_p()
        check(k.feat_handle_create(C.byref(self.h), device, stream), "feat_handle_create")

    def sync(self):
        check(self.k.feat_sync(self.h), "feat_sync")

    def stream(self):
        s = C.c_void_p()
        check(self.k.feat_handle_stream(self.h, C.byref(s)), "feat_handle_stream")
        return s.value

    def close(self):
        if self.h:
            self.k.feat_handle_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def csr_desc(val, col, rowptr, rows, cols):
    cls = CsrF64 if val.dtype.itemsize == 8 else CsrF32
    return cls(rows, cols, val.numel(), val.data_ptr(), col.data_ptr(), rowptr.data_ptr())


def reduce_ws(torch_mod, device):
    """A reduction workspace: (keep-alive tensors, ReduceWs)."""
    part = torch_mod.zeros(2 * 16384, dtype=torch_mod.float64, device=device)
    cnt = torch_mod.zeros(16, dtype=torch_mod.int32, device=device)
    return (part, cnt), ReduceWs(part.data_ptr(), cnt.data_ptr())


# ---------------------------------------------------------------- host layer
class StokesArrays(C.Structure):
    _fields_ = [(n, _vp) for n in ("a_val", "a_col", "a_ptr", "b_val", "b_col", "b_ptr", "d_val", "d_col",
                                   "d_ptr", "filt", "filt_val", "prim", "dual", "pv_val", "pv_col", "pv_ptr",
                                   "pp_val", "pp_col", "pp_ptr")]


def stokes_problem(level):
    """Q2/P1dc lid-driven cavity level (lexicographic numbering), host numpy arrays:
    nv, np, A/B/D as (val, col, rowptr) with 2x2 / 2x1 / 1x2 row-major blocks, the
    velocity filter (nodes, 2 values each), the mean filter vectors and the
    prolongations from level-1 (CSR)."""
    _, h = libs()
    s = np.zeros(10, np.uint64)
    check(h.feat_stokes_sizes(level, s.ctypes.data), "feat_stokes_sizes")
    nv, np_, na, nb, nd, nf, npv, npp, nvc, npc = (int(v) for v in s)
    out = dict(nv=nv, np=np_,
               A=(np.empty(4 * na), np.empty(na, np.uint64), np.empty(nv + 1, np.uint64)),
               B=(np.empty(2 * nb), np.empty(nb, np.uint64), np.empty(nv + 1, np.uint64)),
               D=(np.empty(2 * nd), np.empty(nd, np.uint64), np.empty(np_ + 1, np.uint64)),
               filt=np.empty(nf, np.uint64), filt_val=np.empty(2 * nf), prim=np.empty(np_), dual=np.empty(np_))
    if level > 0:
        out["Pv"] = (np.empty(npv), np.empty(npv, np.uint64), np.empty(nv + 1, np.uint64), nvc)
        out["Pp"] = (np.empty(npp), np.empty(npp, np.uint64), np.empty(np_ + 1, np.uint64), npc)
    a = StokesArrays()
    for key, names in (("A", ("a_val", "a_col", "a_ptr")), ("B", ("b_val", "b_col", "b_ptr")),
                       ("D", ("d_val", "d_col", "d_ptr")), ("Pv", ("pv_val", "pv_col", "pv_ptr")),
                       ("Pp", ("pp_val", "pp_col", "pp_ptr"))):
        if key in out:
            for nm, arr in zip(names, out[key][:3]):
                setattr(a, nm, arr.ctypes.data)
    for nm in ("filt", "filt_val", "prim", "dual"):
        setattr(a, nm, out[nm].ctypes.data)
    check(h.feat_stokes_assemble(level, C.byref(a)), "feat_stokes_assemble")
    return out


class StokesLevelDev(C.Structure):
    """feat_stokes_level_dev_t: device output arrays of feat_stokes_assemble_device."""
    _fields_ = [(n, _vp) for n in ("a_val", "a_col", "a_ptr", "b_val", "b_col", "b_ptr", "d_val", "d_col",
                                   "d_ptr", "filt", "filt_val", "prim", "dual", "pv_val", "pv_col", "pv_ptr",
                                   "pp_val", "pp_col", "pp_ptr", "rv_val", "rv_col", "rv_ptr", "rp_val",
                                   "rp_col", "rp_ptr")]


def stokes_problem_device(torch_mod, handle, level):
    """The level of stokes_problem() assembled on the device (torch tensors, index arrays
    as int64), plus the restrictions Rv = Pv^T and Rp = Pp^T on levels > 0."""
    k, _ = libs()
    s = np.zeros(10, np.uint64)
    check(k.feat_stokes_device_sizes(level, s.ctypes.data), "feat_stokes_device_sizes")
    nv, np_, na, nb, nd, nf, npv, npp, nvc, npc = (int(v) for v in s)
    dev = torch_mod.device("cuda", torch_mod.cuda.current_device())
    f64 = lambda m: torch_mod.empty(m, dtype=torch_mod.float64, device=dev)
    i64 = lambda m: torch_mod.empty(m, dtype=torch_mod.int64, device=dev)
    out = dict(nv=nv, np=np_, A=(f64(4 * na), i64(na), i64(nv + 1)), B=(f64(2 * nb), i64(nb), i64(nv + 1)),
               D=(f64(2 * nd), i64(nd), i64(np_ + 1)), filt=i64(nf), filt_val=f64(2 * nf), prim=f64(np_),
               dual=f64(np_))
    if level > 0:
        out["Pv"] = (f64(npv), i64(npv), i64(nv + 1), nvc)
        out["Pp"] = (f64(npp), i64(npp), i64(np_ + 1), npc)
        out["Rv"] = (f64(npv), i64(npv), i64(nvc + 1), nv)
        out["Rp"] = (f64(npp), i64(npp), i64(npc + 1), np_)
    a = StokesLevelDev()
    for key, pre in (("A", "a"), ("B", "b"), ("D", "d"), ("Pv", "pv"), ("Pp", "pp"), ("Rv", "rv"), ("Rp", "rp")):
        if key in out:
            for nm, t in zip(("val", "col", "ptr"), out[key][:3]):
                setattr(a, f"{pre}_{nm}", t.data_ptr())
    for nm in ("filt", "filt_val", "prim", "dual"):
        setattr(a, nm, out[nm].data_ptr())
    check(k.feat_stokes_assemble_device(handle.h, level, C.byref(a)), "feat_stokes_assemble_device")
    return out


def q1_problem(dim, level):
    """Structured Q1 Poisson level (lexicographic numbering), host numpy arrays."""
    _, h = libs()
    s = np.zeros(5, np.uint64)
    check(h.feat_q1_sizes(dim, level, s.ctypes.data), "feat_q1_sizes")
    n, nnz, nb, pnnz, pcols = (int(v) for v in s)
    val = np.empty(nnz)
    col = np.empty(nnz, np.uint64)
    rp = np.empty(n + 1, np.uint64)
    bnd = np.empty(nb, np.uint64)
    pval = np.empty(max(pnnz, 1))
    pcol = np.empty(max(pnnz, 1), np.uint64)
    prp = np.empty(n + 1, np.uint64)
    check(h.feat_q1_assemble(dim, level, val.ctypes.data, col.ctypes.data, rp.ctypes.data,
                             bnd.ctypes.data, pval.ctypes.data, pcol.ctypes.data, prp.ctypes.data),
          "feat_q1_assemble")
    out = dict(n=n, val=val, col=col, rowptr=rp, filt=bnd)
    if level > 0:
        out.update(pval=pval[:pnnz], pcol=pcol[:pnnz], prowptr=prp, pcols=pcols)
    return out


def q1_problem_device(torch_mod, handle, dim, level, transfer=True):
    """The level of q1_problem() assembled on the device (torch tensors on the
    handle's device, bit-identical to the host arrays); with transfer and level > 0
    also the restriction R = P^T (rval, rcol, rrowptr)."""
    k, _ = libs()
    s = np.zeros(5, np.uint64)
    check(k.feat_q1_device_sizes(dim, level, s.ctypes.data), "feat_q1_device_sizes")
    n, nnz, nb, pnnz, nc = (int(v) for v in s)
    dev = torch_mod.device("cuda", torch_mod.cuda.current_device())
    u64 = torch_mod.int64   # the uint64 indices, as int64 tensors (all below 2^63)
    e = lambda m, dt: torch_mod.empty(m, dtype=dt, device=dev)
    out = dict(n=n, val=e(nnz, torch_mod.float64), col=e(nnz, u64), rowptr=e(n + 1, u64), filt=e(nb, u64))
    tr = [None] * 6
    if transfer and level > 0:
        out.update(pval=e(pnnz, torch_mod.float64), pcol=e(pnnz, u64), prowptr=e(n + 1, u64), pcols=nc,
                   rval=e(pnnz, torch_mod.float64), rcol=e(pnnz, u64), rrowptr=e(nc + 1, u64))
        tr = [out[nm].data_ptr() for nm in ("pval", "pcol", "prowptr", "rval", "rcol", "rrowptr")]
    check(k.feat_q1_assemble_device(handle.h, dim, level, out["val"].data_ptr(), out["col"].data_ptr(),
                                    out["rowptr"].data_ptr(), out["filt"].data_ptr(), *tr),
          "feat_q1_assemble_device")
    return out


FILE_MODE = {"exp": 0, "dv": 1, "mtx": 2, "csr": 4, "binary": 13}   # LAFEM FileMode values


def _fmode(mode):
    return FILE_MODE[mode] if isinstance(mode, str) else int(mode)


def csr_write_file(mode, path, rows, cols, val, col, rowptr, symmetric=False):
    """SparseMatrixCSR::write_out of host arrays (mode "csr", "binary" or "mtx")."""
    _, h = libs()
    val = np.ascontiguousarray(val, np.float64)
    col = np.ascontiguousarray(col, np.uint64)
    rowptr = np.ascontiguousarray(rowptr, np.uint64)
    check(h.feat_csr_write_file(_fmode(mode), str(path).encode(), rows, cols, val.ctypes.data, col.ctypes.data,
                                rowptr.ctypes.data, int(symmetric)), "feat_csr_write_file")


def csr_read_file(mode, path):
    """SparseMatrixCSR::read_from: dict(rows, cols, val, col, rowptr) of host arrays."""
    _, h = libs()
    f = C.c_void_p()
    rows, cols, nnz = _u64(), _u64(), _u64()
    check(h.feat_csr_read_file(_fmode(mode), str(path).encode(), C.byref(f), C.byref(rows), C.byref(cols),
                               C.byref(nnz)), "feat_csr_read_file")
    try:
        val, col = np.empty(nnz.value), np.empty(nnz.value, np.uint64)
        rp = np.empty(rows.value + 1, np.uint64)
        check(h.feat_io_fetch_csr(f, val.ctypes.data, col.ctypes.data, rp.ctypes.data), "feat_io_fetch_csr")
    finally:
        h.feat_io_free(f)
    return dict(rows=rows.value, cols=cols.value, val=val, col=col, rowptr=rp)


def dv_write_file(mode, path, v):
    """DenseVector::write_out of a host array (mode "dv", "binary", "mtx" or "exp")."""
    _, h = libs()
    v = np.ascontiguousarray(v, np.float64)
    check(h.feat_dv_write_file(_fmode(mode), str(path).encode(), v.ctypes.data, v.size), "feat_dv_write_file")


def dv_read_file(mode, path):
    """DenseVector::read_from: host array."""
    _, h = libs()
    f = C.c_void_p()
    n = _u64()
    check(h.feat_dv_read_file(_fmode(mode), str(path).encode(), C.byref(f), C.byref(n)), "feat_dv_read_file")
    try:
        v = np.empty(n.value)
        check(h.feat_io_fetch_dv(f, v.ctypes.data), "feat_io_fetch_dv")
    finally:
        h.feat_io_free(f)
    return v


SOR_MULTICOLOR = 1


class SorPlan:
    """SOR / SSOR plan of a device CSR matrix (feat_sor_*): level-set order (the generic
    sequential sweeps exactly) or, with multicolor=True, the CUDA backend's colour order."""

    def __init__(self, handle, desc, multicolor=False):
        self.k, _ = libs()
        self.h = handle
        self.desc = desc
        self.s = C.c_void_p()
        check(self.k.feat_sor_init_symbolic(handle.h, C.byref(desc), SOR_MULTICOLOR if multicolor else 0,
                                            C.byref(self.s)), "feat_sor_init_symbolic")
        nc = C.c_int()
        check(self.k.feat_sor_colors(self.s, 0, C.byref(nc), None, None), "feat_sor_colors")
        self.ncolors = nc.value

    def colors(self, backward=False):
        """(rank of each row's colour, rows per colour in processing order)."""
        nc = C.c_int()
        check(self.k.feat_sor_colors(self.s, int(backward), C.byref(nc), None, None), "feat_sor_colors")
        rank = np.empty(self.desc.rows, np.uint32)
        rows = np.empty(max(nc.value, 1), np.uint64)
        check(self.k.feat_sor_colors(self.s, int(backward), None, rank.ctypes.data, rows.ctypes.data),
              "feat_sor_colors")
        return rank, rows[:nc.value]

    def init_numeric(self):
        check(self.k.feat_sor_init_numeric(self.h.h, self.s, C.byref(self.desc)), "feat_sor_init_numeric")

    def sor(self, y, x, omega):
        check(self.k.feat_sor_apply(self.h.h, self.s, ptr(y), ptr(x), omega), "feat_sor_apply")

    def ssor(self, y, x, omega):
        """forward and backward sweeps (without the omega (2 - omega) scaling)"""
        check(self.k.feat_ssor_forward_apply(self.h.h, self.s, ptr(y), ptr(x), omega), "feat_ssor_forward_apply")
        check(self.k.feat_ssor_backward_apply(self.h.h, self.s, ptr(y), omega), "feat_ssor_backward_apply")

    def close(self):
        if self.s:
            self.k.feat_sor_done_symbolic(self.s)
            self.s = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class IluPlan:
    """ILU(p) of a device CSR matrix (feat_ilu_*), factorised at construction."""

    def __init__(self, handle, desc, p=0):
        self.k, _ = libs()
        self.h = handle
        self.desc = desc
        self.s = C.c_void_p()
        check(self.k.feat_ilu_init_symbolic_p(handle.h, C.byref(desc), int(p), C.byref(self.s)),
              "feat_ilu_init_symbolic_p")
        self.init_numeric()
        nl, nu, ll, lu = _u64(), _u64(), _i(), _i()
        check(self.k.feat_ilu_info(self.s, C.byref(nl), C.byref(nu), C.byref(ll), C.byref(lu)), "feat_ilu_info")
        self.nnz_l, self.nnz_u, self.levels_l, self.levels_u = nl.value, nu.value, ll.value, lu.value

    def init_numeric(self):
        check(self.k.feat_ilu_init_numeric(self.h.h, self.s, C.byref(self.desc)), "feat_ilu_init_numeric")

    def factors(self):
        l, u, d = np.empty(max(self.nnz_l, 1)), np.empty(max(self.nnz_u, 1)), np.empty(self.desc.rows)
        check(self.k.feat_ilu_factors(self.h.h, self.s, l.ctypes.data, u.ctypes.data, d.ctypes.data), "feat_ilu_factors")
        return l[:self.nnz_l], u[:self.nnz_u], d

    def apply(self, y, x):
        check(self.k.feat_ilu_apply(self.h.h, self.s, ptr(y), ptr(x)), "feat_ilu_apply")

    def close(self):
        if self.s:
            self.k.feat_ilu_done_symbolic(self.s)
            self.s = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def q1_rhs_sine(dim, level, n):
    _, h = libs()
    b = np.empty(n)
    check(h.feat_q1_rhs_sine(dim, level, b.ctypes.data), "feat_q1_rhs_sine")
    return b


def rhs_random(seed, n, a=-1.0, b=1.0):
    _, h = libs()
    v = np.empty(n)
    check(h.feat_rhs_random(seed, n, a, b, v.ctypes.data), "feat_rhs_random")
    return v


def mgpcg_cfg(nu_pre=4, nu_post=4, omega_jac=0.5, omega_rich=1.0, coarse="direct",
              coarse_omega=1.0, tol_rel=1e-8, max_iter=100, force_generic=False, csr_only=False,
              precision="f64", graphs=True, autotune=True, coarse_fusion=True):
    return MgPcgCfg(nu_pre, nu_post, omega_jac, omega_rich, COARSE[coarse], coarse_omega, tol_rel,
                    max_iter, int(force_generic), int(csr_only), {"f64": 0, "f32": 1, "f16": 2}[precision],
                    int(not graphs), int(not autotune), int(not coarse_fusion))


def csr_matrix(val, col, rowptr, rows, cols):
    """CSR descriptor (Matrix) of device tensors."""
    k, _ = libs()
    c = csr_desc(val, col, rowptr, rows, cols)
    m = Matrix()
    if val.dtype.itemsize == 8 and col.dtype.itemsize == 8:
        check(k.feat_matrix_from_csr_f64(C.byref(c), C.byref(m)), "feat_matrix_from_csr_f64")
    else:
        idx = IDX_U64 if col.dtype.itemsize == 8 else IDX_U32
        m = Matrix(FMT_CSR, idx, rows, cols, val.numel(), val.data_ptr(), col.data_ptr(),
                   rowptr.data_ptr(), None, 0, 1.0)
    return m


class SellPlan:
    """SELL-64 plan of a device CSR (f64 values with u64 indices); fp='f32' rounds values.
    dict=False keeps full values and column offsets where the pair dictionary applies;
    diag=False keeps the dictionary in entry order where the diagonal layout applies;
    banded=False keeps the plain plan where the banded plan (grid stencil rows) applies."""

    def __init__(self, handle, val, col, rowptr, rows, cols, fp="f64", dict=True, diag=True, banded=True):
        self.k, _ = libs()
        self.fp = fp
        c = csr_desc(val, col, rowptr, rows, cols)
        self.m = Matrix()
        flags = (0 if dict else SELL_NO_DICT) | (0 if diag else SELL_NO_DIAG) | (0 if banded else SELL_NO_BANDED)
        fn = {"f64": self.k.feat_sell_create_ex_f64, "f32": self.k.feat_sell_create_ex_f32,
              "f16": self.k.feat_sell_create_ex_f16}[fp]
        check(fn(handle.h, C.byref(c), flags, C.byref(self.m)), fn.__name__)

    def close(self):
        if self.m.format in (FMT_SELL, FMT_BANDED):
            fn = self.k.feat_sell_destroy_f64 if self.fp == "f64" else self.k.feat_sell_destroy_f32
            fn(C.byref(self.m))

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def runtime_init(device=0, stream=None):
    """Binds the host layer to a HIP stream (default: torch's current stream)."""
    _, h = libs()
    if stream is None:
        import torch
        stream = torch.cuda.current_stream(device).cuda_stream
    check(h.feat_runtime_init(device, stream), "feat_runtime_init")


def runtime_grid_forms(on):
    """Multigrid solvers initialised from now on use the grid forms where the levels qualify
    (d-linear transfer descriptors, fused coarse levels' stencils; default on)."""
    _, h = libs()
    check(h.feat_runtime_grid_forms(int(bool(on))), "feat_runtime_grid_forms")


def runtime_sell_flags(flags):
    """Flags of the SELL-64 plans the host layer builds from now on (SELL_NO_DICT: plain plans)."""
    _, h = libs()
    check(h.feat_runtime_sell_flags(int(flags)), "feat_runtime_sell_flags")


class StokesCfg(C.Structure):
    _fields_ = [("steps", _i), ("omega_rich", _d), ("vanka_omega", _d), ("coarse_skip_singular", _i),
                ("omega_outer", _d), ("tol_rel", _d), ("max_iter", _u64), ("no_graphs", _i),
                ("no_coarse_fusion", _i), ("tail_max_level_dofs", _u64)]


class Stokes:
    """Richardson(MG-V(Richardson(Schwarz(AmaVanka)))) on the Q2/P1dc lid-driven cavity,
    levels lf..lc, of the C++ host layer (AmaVanka factorised on the device)."""

    KIND = {"rhs": 0, "sol": 1, "def": 2, "cor": 3}

    def __init__(self, lf, lc, steps=5, omega_rich=0.7, vanka_omega=1.0, coarse_skip_singular=True,
                 omega_outer=1.0, tol_rel=1e-8, max_iter=100, graphs=True, coarse_fusion=True,
                 tail_max_level_dofs=0):
        _, self.hl = libs()
        runtime_init()
        self.s = C.c_void_p()
        self.cfg = StokesCfg(steps, omega_rich, vanka_omega, int(coarse_skip_singular), omega_outer, tol_rel,
                             max_iter, int(not graphs), int(not coarse_fusion), tail_max_level_dofs)
        check(self.hl.feat_stokes_create(C.byref(self.s), lf, lc, C.byref(self.cfg)), "feat_stokes_create")

    def sizes(self, lvl=0):
        o = np.zeros(8, np.uint64)
        check(self.hl.feat_stokes_sizes_of(self.s, lvl, o.ctypes.data), "feat_stokes_sizes_of")
        return dict(nv=int(o[0]), np=int(o[1]), macros=int(o[2]), stride=int(o[3]),
                    nnz=[int(v) for v in o[4:8]])

    def n(self, lvl=0):
        z = self.sizes(lvl)
        return 2 * z["nv"] + z["np"]

    def vanka(self, lvl=0):
        z = self.sizes(lvl)
        nv, np_ = z["nv"], z["np"]
        shapes = [(nv, 2, 2), (nv, 2, 1), (np_, 1, 2), (np_, 1, 1)]
        blocks = [dict(rowptr=np.zeros(r + 1, np.uint64), col=np.zeros(max(k, 1), np.uint64),
                       val=np.zeros(max(k, 1) * bh * bw)) for (r, bh, bw), k in zip(shapes, z["nnz"])]
        P = (C.c_void_p * 4)(*[b["rowptr"].ctypes.data for b in blocks])
        Q = (C.c_void_p * 4)(*[b["col"].ctypes.data for b in blocks])
        V = (C.c_void_p * 4)(*[b["val"].ctypes.data for b in blocks])
        mask = np.zeros(max(z["macros"], 1), np.int32)
        check(self.hl.feat_stokes_vanka_copy(self.s, lvl, P, Q, V, mask.ctypes.data), "feat_stokes_vanka_copy")
        return dict(blocks=blocks, mask=mask[:z["macros"]], **z)

    def apply(self, lvl, r, x, y=None, alpha=1.0):
        check(self.hl.feat_stokes_apply(self.s, lvl, ptr(r), ptr(x), ptr(y) if y is not None else None, alpha),
              "feat_stokes_apply")

    def vanka_apply(self, lvl, c, d, passes=1):
        check(self.hl.feat_stokes_vanka_apply(self.s, lvl, ptr(c), ptr(d), passes), "feat_stokes_vanka_apply")

    def smooth(self, lvl, sol, rhs):
        check(self.hl.feat_stokes_smooth(self.s, lvl, ptr(sol), ptr(rhs)), "feat_stokes_smooth")

    def filter(self, lvl, v, kind):
        check(self.hl.feat_stokes_filter(self.s, lvl, ptr(v), self.KIND[kind]), "feat_stokes_filter")

    def transfer(self, lvl, prol, r, x):
        check(self.hl.feat_stokes_transfer(self.s, lvl, int(prol), ptr(r), ptr(x)), "feat_stokes_transfer")

    def init_numeric(self):
        t = C.c_double()
        check(self.hl.feat_stokes_init_numeric(self.s, C.byref(t)), "feat_stokes_init_numeric")
        return t.value

    def solve(self, x, b, hist_cap=1000):
        res = SolveResult()
        hist = np.full(hist_cap, np.nan)
        check(self.hl.feat_stokes_solve(self.s, ptr(x), ptr(b), C.byref(res), hist.ctypes.data, hist_cap),
              "feat_stokes_solve")
        it = int(res.iters)
        return dict(status=STATUS[res.status], iters=it, def_init=res.def_init, def_final=res.def_final,
                    hist=hist[:it + 1])

    def close(self):
        if self.s:
            self.hl.feat_stokes_destroy(self.s)
            self.s = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class BatchCfg(C.Structure):
    _fields_ = [("seed", _u64), ("ninst", _i), ("lmin", _i), ("lmax", _i), ("sigma", _d), ("tol_rel", _d),
                ("max_iter", _u64), ("autotune", _i), ("graphs", _i), ("tail_levels", _i),
                ("no_tiled", _i), ("tiled_prol_arrays", _i), ("tail_ell", _i), ("no_tiled_restriction", _i)]


def batch_levels(seed, ninst, lmin, lmax):
    _, h = libs()
    out = np.zeros(ninst, np.int32)
    check(h.feat_batch_levels(seed, ninst, lmin, lmax, out.ctypes.data), "feat_batch_levels")
    return out


def lognormal_cells(seed, n, sigma):
    _, h = libs()
    out = np.zeros(n)
    check(h.feat_lognormal_cells(seed, n, sigma, out.ctypes.data), "feat_lognormal_cells")
    return out


def q1_diffusion(level, level_c, coef):
    """Q1 diffusion matrix values on `level` (pattern of q1_problem(2, level))."""
    _, h = libs()
    q = q1_problem(2, level)
    val = np.zeros(q["val"].size)
    coef = np.ascontiguousarray(coef, dtype=np.float64)
    check(h.feat_q1_diffusion_assemble(level, level_c, coef.ctypes.data, val.ctypes.data, None, None),
          "feat_q1_diffusion_assemble")
    return val


def q1_diffusion3d(level, level_c, coef, kz):
    """3D Q1 diffusion matrix values of diag(1, 1, kz) on `level` (pattern of q1_problem(3, level))."""
    _, h = libs()
    q = q1_problem(3, level)
    val = np.zeros(q["val"].size)
    coef = np.ascontiguousarray(coef, dtype=np.float64)
    check(h.feat_q1_diffusion3d_assemble(level, level_c, coef.ctypes.data, kz, val.ctypes.data, None, None),
          "feat_q1_diffusion3d_assemble")
    return val


def diffusion3d_levels(lf, lc, seed, sigma, kz):
    """Host levels (finest first, as for MgPcg(levels=...)) of the 3D Q1 anisotropic diffusion
    hierarchy: a log-normal(sigma) coefficient per finest cell (seed), tensor diag(1, 1, kz)."""
    coef = lognormal_cells(seed, 1 << (3 * lf), sigma)
    levels = []
    for k in range(lf, lc - 1, -1):
        q = dict(q1_problem(3, k))
        q["val"] = q1_diffusion3d(k, lf, coef, kz)
        levels.append(q)
    return levels


def batch_rhs(bs, seed=28054777172512):
    """Right-hand sides of the standard batch in the PCG layout (host): instance i
    uniform [-1, 1] from seed + 2000 + i, boundary rows zero."""
    b = np.zeros(int(bs.off[-1]))
    for i, L in enumerate(bs.levels):
        lo, hi = int(bs.off[i]), int(bs.off[i + 1])
        bi = rhs_random(seed + 2000 + i, hi - lo)
        m = 2 ** int(L) + 1
        idx = np.arange(hi - lo)
        bi[(idx % m == 0) | (idx % m == m - 1) | (idx // m == 0) | (idx // m == m - 1)] = 0.0
        b[lo:hi] = bi
    return b


def batch_instance_levels(L, coef):
    """Host levels (finest first, as for MgPcg(levels=...)) of one batch instance."""
    levels = []
    for k in range(L, 0, -1):
        q = dict(q1_problem(2, k))
        q["val"] = q1_diffusion(k, L, coef)
        levels.append(q)
    return levels


class BatchSolver:
    """The batch of independent Q1 diffusion MG-PCG solves (C++ host layer)."""

    def __init__(self, seed=28054777172512, ninst=64, lmin=3, lmax=8, sigma=1.0, tol_rel=1e-8, max_iter=100,
                 autotune=True, graphs=True, tail_levels=5, tiled=True, tiled_prol_arrays=False, tail_ell=False,
                 tiled_restriction=True):
        _, self.hl = libs()
        runtime_init()
        self.s = C.c_void_p()
        self.cfg = BatchCfg(seed, ninst, lmin, lmax, sigma, tol_rel, max_iter, int(autotune), int(graphs),
                            int(tail_levels), int(not tiled), int(tiled_prol_arrays), int(tail_ell),
                            int(not tiled_restriction))
        check(self.hl.feat_batchsolver_create(C.byref(self.s), C.byref(self.cfg)), "feat_batchsolver_create")
        self.levels = np.zeros(ninst, np.int32)
        self.off = np.zeros(ninst + 1, np.uint64)
        check(self.hl.feat_batchsolver_layout(self.s, self.levels.ctypes.data, self.off.ctypes.data),
              "feat_batchsolver_layout")
        self.ninst, self.max_iter = ninst, max_iter

    @property
    def n(self):
        return int(self.off[-1])

    def solve(self, x, b):
        ni = self.ninst
        st = np.zeros(ni, np.int32)
        it = np.zeros(ni, np.uint64)
        d0, d1 = np.zeros(ni), np.zeros(ni)
        hist = np.zeros(ni * (self.max_iter + 1))
        check(self.hl.feat_batchsolver_solve(self.s, ptr(x), ptr(b), st.ctypes.data, it.ctypes.data, d0.ctypes.data,
                                             d1.ctypes.data, hist.ctypes.data), "feat_batchsolver_solve")
        return dict(status=[STATUS[int(v)] for v in st], iters=it.astype(int), def_init=d0, def_final=d1,
                    hist=hist.reshape(ni, self.max_iter + 1))

    @property
    def tail_levels(self):
        d, g = C.c_int(), C.c_int()
        check(self.hl.feat_batchsolver_tail_levels(self.s, C.byref(d), C.byref(g)), "feat_batchsolver_tail_levels")
        return d.value

    @property
    def tail_grid(self):
        d, g = C.c_int(), C.c_int()
        check(self.hl.feat_batchsolver_tail_levels(self.s, C.byref(d), C.byref(g)), "feat_batchsolver_tail_levels")
        return bool(g.value)

    @property
    def tiled_levels(self):
        v = np.zeros(16, np.int32)
        cnt = C.c_int()
        check(self.hl.feat_batchsolver_tiled_levels(self.s, v.ctypes.data, 16, C.byref(cnt)),
              "feat_batchsolver_tiled_levels")
        return [int(x) for x in v[:cnt.value]]

    def tuning(self):
        t = np.zeros(64, np.int32)
        u = np.zeros(64)
        cnt = C.c_int()
        check(self.hl.feat_batchsolver_tuning(self.s, t.ctypes.data, u.ctypes.data, 64, C.byref(cnt)),
              "feat_batchsolver_tuning")
        return t[:cnt.value], u[:cnt.value]

    def close(self):
        if self.s:
            self.hl.feat_batchsolver_destroy(self.s)
            self.s = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class LearnedCfg(C.Structure):
    _fields_ = [("level_fine", _i), ("level_coarse", _i), ("sigma", _d), ("kz", _d), ("seed", _u64), ("nu", _i),
                ("omega_jac", _d), ("omega_learned", _d), ("tol_rel", _d), ("max_iter", _u64), ("n_init", _i),
                ("n_snap", _i), ("batch", _u64), ("lr", _f), ("fcg_m", _i), ("learned_steps", _i)]


class LearnedResult(C.Structure):
    _fields_ = [("status", _i), ("iters", _u64), ("fell_back", _i), ("fallback_iter", _u64), ("def_init", _d),
                ("def_final", _d), ("seconds", _d)]


MLP_NPARAM, MLP_NPARAM_PAD = 8385, 8448


class LearnedMG:
    """Flexible PCG(MG-V) on the 3D Q1 anisotropic variable-coefficient problem with
    a Jacobi or a learned (BF16 MLP) finest-level smoother (C++ host layer)."""

    def __init__(self, level_fine=6, level_coarse=1, sigma=1.5, kz=10.0, seed=28054777172512, nu=4, omega_jac=0.5,
                 omega_learned=1.0, tol_rel=1e-8, max_iter=200, n_init=4, n_snap=2, batch=8192, lr=1e-3, fcg_m=4,
                 learned_steps=0):
        _, self.hl = libs()
        runtime_init()
        self.s = C.c_void_p()
        self.cfg = LearnedCfg(level_fine, level_coarse, sigma, kz, seed, nu, omega_jac, omega_learned, tol_rel, max_iter,
                              n_init, n_snap, batch, lr, fcg_m, learned_steps)
        check(self.hl.feat_learned_create(C.byref(self.s), C.byref(self.cfg)), "feat_learned_create")
        n, m = _u64(), _u64()
        check(self.hl.feat_learned_size(self.s, C.byref(n), C.byref(m)), "feat_learned_size")
        self.n, self.m = n.value, m.value

    def generate(self, on_policy=False):
        rho = C.c_double()
        check(self.hl.feat_learned_generate(self.s, int(on_policy), C.byref(rho)), "feat_learned_generate")
        return rho.value

    def energy(self):
        e = C.c_double()
        check(self.hl.feat_learned_energy(self.s, C.byref(e)), "feat_learned_energy")
        return e.value

    def fit(self, rounds=4, steps=300, on_policy=False, seed=1):
        """Training with model selection on a second right-hand side (see feat_learned_fit).
        on_policy refreshes the snapshots with the learned smoother's own cycles before rounds
        2.. (off by default: networks trained on them failed their validation solves)."""
        br, ji, bi = C.c_int(), _u64(), _u64()
        it = np.zeros(rounds, np.uint64)
        check(self.hl.feat_learned_fit(self.s, rounds, steps, int(on_policy), seed, C.byref(br), C.byref(ji), C.byref(bi),
                                       it.ctypes.data), "feat_learned_fit")
        return dict(best_round=br.value, jacobi_iters=ji.value, best_iters=bi.value, round_iters=[int(v) for v in it])

    def train(self, steps, seed=1):
        loss = np.zeros(max(steps, 1), np.float32)
        check(self.hl.feat_learned_train(self.s, steps, seed, loss.ctypes.data), "feat_learned_train")
        return loss[:steps]

    def solve(self, learned, allow_fallback=True, x=None, hist_cap=1000):
        res = LearnedResult()
        hist = np.full(hist_cap, np.nan)
        check(self.hl.feat_learned_solve(self.s, int(learned), int(allow_fallback), ptr(x) if x is not None else None,
                                         C.byref(res), hist.ctypes.data, hist_cap), "feat_learned_solve")
        it = int(res.iters)
        return dict(status=STATUS[res.status], iters=it, fell_back=bool(res.fell_back),
                    fallback_iter=int(res.fallback_iter), def_init=res.def_init, def_final=res.def_final,
                    seconds=res.seconds, hist=hist[:it + 1])

    def params(self, new=None):
        cur = np.zeros(MLP_NPARAM, np.float32)
        nw = None if new is None else np.ascontiguousarray(new, dtype=np.float32)
        check(self.hl.feat_learned_params(self.s, cur.ctypes.data, nw.ctypes.data if nw is not None else None),
              "feat_learned_params")
        return cur

    def param_ptrs(self):
        p, w = _vp(), _vp()
        check(self.hl.feat_learned_param_ptrs(self.s, C.byref(p), C.byref(w)), "feat_learned_param_ptrs")
        return p, w

    def close(self):
        if self.s:
            self.hl.feat_learned_destroy(self.s)
            self.s = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


PRECOND = {"none": 0, "jacobi": 1, "sor": 2, "ssor": 3, "ilu": 4, "sor_mc": 5, "ssor_mc": 6}


class KrylovSolver:
    """PCG ("pcg") or Richardson ("richardson") around a device preconditioner
    ("none", "jacobi", "sor", "ssor", "ilu") of the C++ host layer."""

    def __init__(self, solver, A, filt, precond, omega_pre=1.0, omega_rich=1.0, tol_rel=1e-8, max_iter=1000):
        _, self.hl = libs()
        runtime_init()
        self.s = C.c_void_p()
        self.n = int(A.rows)
        val = np.ascontiguousarray(A.val, np.float64)
        col = np.ascontiguousarray(A.col, np.uint64)
        rp = np.ascontiguousarray(A.rowptr, np.uint64)
        f = np.ascontiguousarray(filt, np.uint64)
        check(self.hl.feat_krylov_create(C.byref(self.s), {"pcg": 0, "richardson": 1}[solver], self.n,
                                         val.ctypes.data, col.ctypes.data, rp.ctypes.data,
                                         f.ctypes.data if f.size else None, f.size, PRECOND[precond],
                                         omega_pre, omega_rich, tol_rel, max_iter), "feat_krylov_create")

    def solve(self, x, b, hist_cap=2000):
        res = SolveResult()
        hist = np.full(hist_cap, np.nan)
        check(self.hl.feat_krylov_solve(self.s, ptr(x), ptr(b), C.byref(res), hist.ctypes.data, hist_cap),
              "feat_krylov_solve")
        it = int(res.iters)
        return dict(status=STATUS[res.status], iters=it, def_init=res.def_init, def_final=res.def_final,
                    hist=hist[:it + 1])

    def precond(self, cor, d):
        check(self.hl.feat_krylov_precond_apply(self.s, ptr(cor), ptr(d)), "feat_krylov_precond_apply")

    def close(self):
        if self.s:
            self.hl.feat_krylov_destroy(self.s)
            self.s = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class MgPcg:
    """PCG(MultiGrid-V(Richardson(Jacobi))) of the C++ host layer."""

    def __init__(self, cfg, levels=None, q1=None, host_assembly=False):
        _, self.hl = libs()
        runtime_init()
        self.s = C.c_void_p()
        self.cfg = cfg
        if q1 is not None:
            dim, lf, lc = q1
            check(self.hl.feat_mgpcg_create_q1_ex(C.byref(self.s), dim, lf, lc, C.byref(cfg),
                                                  Q1_HOST_ASSEMBLY if host_assembly else 0),
                  "feat_mgpcg_create_q1_ex")
        else:
            arr = (LevelHost * len(levels))()
            keep = []
            for i, lv in enumerate(levels):
                a = {k: np.ascontiguousarray(v) for k, v in lv.items() if isinstance(v, np.ndarray)}
                keep.append(a)
                hasp = "pval" in a and i + 1 < len(levels)
                arr[i] = LevelHost(lv["n"], a["val"].ctypes.data, a["col"].ctypes.data,
                                   a["rowptr"].ctypes.data,
                                   a["filt"].ctypes.data if a["filt"].size else None, a["filt"].size,
                                   lv.get("pcols", 0) if hasp else 0,
                                   a["pval"].ctypes.data if hasp else None,
                                   a["pcol"].ctypes.data if hasp else None,
                                   a["prowptr"].ctypes.data if hasp else None)
            check(self.hl.feat_mgpcg_create(C.byref(self.s), arr, len(levels), C.byref(cfg)),
                  "feat_mgpcg_create")

    def solve(self, x, b, hist_cap=1000):
        """x (initial guess, overwritten) and b: torch float64 device tensors."""
        res = SolveResult()
        hist = np.full(hist_cap, np.nan)
        check(self.hl.feat_mgpcg_solve(self.s, ptr(x), ptr(b), C.byref(res), hist.ctypes.data,
                                       hist_cap), "feat_mgpcg_solve")
        it = int(res.iters)
        return dict(status=STATUS[res.status], iters=it, def_init=res.def_init,
                    def_final=res.def_final, hist=hist[:it + 1])

    def write_matrix(self, lvl, mode, path, symmetric=False):
        """SparseMatrixCSR::write_out of level lvl's device-resident matrix."""
        check(self.hl.feat_mgpcg_write_matrix(self.s, lvl, _fmode(mode), str(path).encode(), int(symmetric)),
              "feat_mgpcg_write_matrix")

    def matrix(self, lvl=0):
        rows, nnz = _u64(), _u64()
        v, c, r = _vp(), _vp(), _vp()
        check(self.hl.feat_mgpcg_matrix(self.s, lvl, C.byref(rows), C.byref(nnz), C.byref(v),
                                        C.byref(c), C.byref(r)), "feat_mgpcg_matrix")
        return CsrF64(rows.value, rows.value, nnz.value, v.value, c.value, r.value)

    def set_values(self, lvl, val):
        """New values (host array, same pattern) for level lvl's matrix, then the
        numeric re-initialisation of the hierarchy and the solver tree."""
        v = np.ascontiguousarray(val, np.float64)
        check(self.hl.feat_mgpcg_set_values(self.s, lvl, v.ctypes.data), "feat_mgpcg_set_values")

    def coarse_fusion_grid(self):
        lv = C.c_int()
        check(self.hl.feat_mgpcg_coarse_fusion_grid(self.s, C.byref(lv)), "feat_mgpcg_coarse_fusion_grid")
        return lv.value

    def coarse_fusion_level(self):
        """First level of the coarse levels fused into one workgroup (0: none)."""
        lv = C.c_int(0)
        check(self.hl.feat_mgpcg_coarse_fusion_level(self.s, C.byref(lv)), "feat_mgpcg_coarse_fusion_level")
        return lv.value

    def close(self):
        if self.s:
            self.hl.feat_mgpcg_destroy(self.s)
            self.s = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
