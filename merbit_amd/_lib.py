"""ctypes binding of libmerbit_amd.so (include/merbit_amd.h, include/merbit_gen.h).

The library is the product: there is no Python or CPU fallback for any
device entry point.  If the shared library is missing this module raises on
first use, naming the build command.
"""
from __future__ import annotations

import ctypes
import os

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG, "libmerbit_amd.so")

# Every symbol include/merbit_amd.h and include/merbit_gen.h declare.
ABI_SYMBOLS = (
    "mbx_plan_create", "mbx_plan_create_rows", "mbx_plan_destroy", "mbx_plan_clone",
    "mbx_plan_info",
    "mbx_plan_partition", "mbx_plan_set_stream", "mbx_batch_set_stream",
    "mbx_plan_set_residency", "mbx_batch_set_residency", "mbx_plan_compress_cols",
    "mbx_plan_sort_cols", "mbx_batch_sort_cols", "mbx_colblock_sort_cols",
    "mbx_plan_set_accumulation",
    "mbx_plan_set_events", "mbx_spmv_f64", "mbx_spmv_f32", "mbx_spmv_bf16",
    "mbx_spmm_workspace_bytes", "mbx_spmm_f32",
    "mbx_batch_create", "mbx_batch_destroy", "mbx_batch_info", "mbx_batch_spmv",
    "mbx_batch_graph_create", "mbx_spmv_graph_create", "mbx_graph_launch", "mbx_graph_destroy",
    "mbx_autotune_spmv", "mbx_autotune_spmv_launch", "mbx_autotune_spmv_form",
    "mbx_autotune_spmm", "mbx_autotune_batch",
    "mbx_adjoint_create", "mbx_adjoint_destroy", "mbx_adjoint_transpose", "mbx_adjoint_dx",
    "mbx_adjoint_set_blocks", "mbx_adjoint_auto",
    "mbx_sddmm", "mbx_ingest_ref", "mbx_plan_stats",
    "mbx_pagerank_workspace_bytes", "mbx_pagerank_f64", "mbx_pagerank_graph_create",
    "mbx_colblock_create", "mbx_colblock_destroy", "mbx_colblock_info", "mbx_colblock_values",
    "mbx_colblock_spmv", "mbx_colblock_sddmm", "mbx_colblock_values_to_csr", "mbx_colblock_set_stream",
    "mbx_colblock_set_residency", "mbx_colblock_adjoint_prepare", "mbx_colblock_dx",
    "mbx_csr_validate", "mbx_spmv_f64_noalloc", "mbx_plan_split_hot",
    "mbx_launch_rule", "mbx_plan_create_auto", "mbx_status_str",
)
GEN_SYMBOLS = (
    "mbg_fill_uniform_f64", "mbg_fill_uniform_f32", "mbg_fill_normal_f64",
    "mbg_fill_normal_f32", "mbg_f32_to_bf16", "mbg_laplacian5_nnz", "mbg_laplacian5",
    "mbg_rmat", "mbg_banded_nnz", "mbg_banded", "mbg_poisson_rows",
)
HOST_SYMBOLS = ("mbh_config_lookup", "mbh_mtx_read", "mbh_ref_ingest")

MBX_OK = 0
MBX_COL_MAJOR = 0
MBX_ROW_MAJOR = 1
MBX_VALUE_F64 = 0
MBX_VALUE_F32 = 1
MBX_VALUE_BF16 = 2
MBX_STREAM_AUTO = 0
MBX_STREAM_CACHED = 1
MBX_STREAM_NONTEMPORAL = 2


class BatchItem(ctypes.Structure):
    """mbx_batch_item of include/merbit_amd.h."""
    _fields_ = [("row_ptr", ctypes.c_void_p), ("col", ctypes.c_void_p),
                ("val", ctypes.c_void_p), ("x", ctypes.c_void_p), ("y", ctypes.c_void_p),
                ("nrow", ctypes.c_int32), ("ncol", ctypes.c_int32), ("nnz", ctypes.c_int32)]

class PatternStats(ctypes.Structure):
    """mbx_pattern_stats of include/merbit_amd.h."""
    _fields_ = [("nrow", ctypes.c_int32), ("ncol", ctypes.c_int32), ("nnz", ctypes.c_int32),
                ("mean", ctypes.c_double), ("var", ctypes.c_double), ("skew", ctypes.c_double),
                ("kurt", ctypes.c_double), ("empty_rows", ctypes.c_int32),
                ("max_row", ctypes.c_int32), ("ntiles", ctypes.c_int32),
                ("tile_items", ctypes.c_int32), ("single_row_tiles", ctypes.c_int32),
                ("split_tiles", ctypes.c_int32), ("max_rows_per_tile", ctypes.c_int32),
                ("long_rows", ctypes.c_int32)]


class LaunchChoice(ctypes.Structure):
    """mbx_launch_choice of include/merbit_amd.h."""
    _fields_ = [("rule", ctypes.c_int32), ("items_per_thread", ctypes.c_int32),
                ("stream_mode", ctypes.c_int32), ("wg_per_cu", ctypes.c_int32),
                ("accumulation", ctypes.c_int32), ("compress_cols", ctypes.c_int32),
                ("hot_columns", ctypes.c_int32), ("column_blocks", ctypes.c_int32),
                ("row_tiles", ctypes.c_int32), ("sort_cols", ctypes.c_int32),
                ("cus", ctypes.c_int32), ("ntiles8", ctypes.c_int32),
                ("max_tile_span", ctypes.c_int32), ("max_row", ctypes.c_int32),
                ("x_bytes", ctypes.c_double), ("hot_share", ctypes.c_double)]


RULE_NAMES = {1: "few_tiles", 2: "local", 3: "x_fits", 4: "hot_columns", 5: "spread_columns"}


class CsrReport(ctypes.Structure):
    """mbx_csr_report of include/merbit_amd.h."""
    _fields_ = [("row_ptr_errors", ctypes.c_int64), ("col_errors", ctypes.c_int64),
                ("unsorted_pairs", ctypes.c_int64), ("nnz_mismatch", ctypes.c_int32),
                ("first_bad_col", ctypes.c_int32)]


_lib = None


class MerbitError(RuntimeError):
    """A negative MBX_* status returned by the C ABI."""

    def __init__(self, status: int, what: str):
        self.status = status
        super().__init__(f"{what}: {status_str(status)} (status {status})")


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is not built; run `python -m merbit_amd.build` "
            "(the HIP library is required: there is no CPU fallback)")
    L = ctypes.CDLL(LIB_PATH)
    vp, i32, i64, u64 = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, ctypes.c_uint64
    f64, f32 = ctypes.c_double, ctypes.c_float
    L.mbx_plan_create.argtypes = [vp, i32, i32, i32, i32, i32, vp, ctypes.POINTER(vp)]
    L.mbx_plan_create_rows.argtypes = [vp, i32, i32, i32, i32, i32, vp, ctypes.POINTER(vp)]
    L.mbx_plan_destroy.argtypes = [vp]
    L.mbx_plan_clone.argtypes = [vp, ctypes.POINTER(vp)]
    L.mbx_plan_set_stream.argtypes = [vp, i32]
    L.mbx_plan_set_residency.argtypes = [vp, i32]
    L.mbx_batch_set_residency.argtypes = [vp, i32]
    L.mbx_batch_set_stream.argtypes = [vp, i32]
    L.mbx_plan_destroy.restype = None
    L.mbx_plan_info.argtypes = [vp, vp, vp, vp]
    L.mbx_plan_partition.argtypes = [vp, vp, vp, i32]
    L.mbx_plan_set_events.argtypes = [vp, vp, vp]
    L.mbx_spmv_f64.argtypes = [vp, vp, vp, vp, f64, vp, f64, vp, vp]
    L.mbx_spmv_f32.argtypes = [vp, vp, vp, vp, f32, vp, f32, vp, vp]
    L.mbx_spmv_bf16.argtypes = [vp, vp, vp, vp, f32, vp, f32, vp, vp]
    L.mbx_spmm_workspace_bytes.argtypes = [vp, i32, vp, i64, i32]
    L.mbx_spmm_workspace_bytes.restype = ctypes.c_size_t
    L.mbx_spmm_f32.argtypes = [vp, vp, vp, vp, i32, f32, vp, i64, i32, f32, vp, i64, i32,
                               vp, ctypes.c_size_t, vp]
    L.mbx_batch_create.argtypes = [vp, i32, i32, i32, vp, ctypes.POINTER(vp)]
    L.mbx_batch_destroy.argtypes = [vp]
    L.mbx_batch_destroy.restype = None
    L.mbx_batch_info.argtypes = [vp, vp, vp, vp]
    L.mbx_batch_spmv.argtypes = [vp, f64, f64, vp]
    L.mbx_batch_graph_create.argtypes = [vp, f64, f64, i32, ctypes.POINTER(vp)]
    L.mbx_spmv_graph_create.argtypes = [vp, i32, vp, vp, vp, f64, vp, f64, vp, i32,
                                        ctypes.POINTER(vp)]
    L.mbx_graph_launch.argtypes = [vp, vp]
    L.mbx_graph_destroy.argtypes = [vp]
    L.mbx_graph_destroy.restype = None
    L.mbx_autotune_spmv.argtypes = [vp, i32, i32, i32, vp, vp, i32, vp, vp, i32, vp,
                                    ctypes.POINTER(i32), ctypes.POINTER(f32)]
    L.mbx_autotune_spmv_launch.argtypes = [vp, i32, i32, i32, vp, vp, i32, vp, vp, i32, vp,
                                           ctypes.POINTER(i32), ctypes.POINTER(i32),
                                           ctypes.POINTER(i32), ctypes.POINTER(f32)]
    L.mbx_autotune_spmv_form.argtypes = [vp, i32, i32, i32, vp, vp, i32, vp, vp, i32, vp,
                                         ctypes.POINTER(i32), ctypes.POINTER(i32),
                                         ctypes.POINTER(i32), ctypes.POINTER(i32),
                                         ctypes.POINTER(f32)]
    L.mbx_autotune_spmm.argtypes = [vp, i32, i32, i32, vp, vp, i32, vp, i64, i32, vp, i64, i32,
                                    vp, ctypes.c_size_t, i32, vp, ctypes.POINTER(i32),
                                    ctypes.POINTER(f32)]
    L.mbx_autotune_batch.argtypes = [vp, i32, i32, i32, vp, ctypes.POINTER(i32),
                                     ctypes.POINTER(f32)]
    L.mbx_plan_compress_cols.argtypes = [vp, vp, vp]
    L.mbx_plan_sort_cols.argtypes = [vp, vp, vp]
    L.mbx_batch_sort_cols.argtypes = [vp, i32, vp]
    L.mbx_colblock_sort_cols.argtypes = [vp, i32, vp]
    L.mbx_plan_set_accumulation.argtypes = [vp, i32]
    L.mbx_colblock_create.argtypes = [vp, vp, i32, i32, i32, i32, i32, vp, ctypes.POINTER(vp)]
    L.mbx_colblock_destroy.argtypes = [vp]
    L.mbx_colblock_destroy.restype = None
    L.mbx_colblock_info.argtypes = [vp, ctypes.POINTER(i32), ctypes.POINTER(ctypes.c_int64), vp]
    L.mbx_colblock_values.argtypes = [vp, i32, vp, vp, vp]
    L.mbx_colblock_spmv.argtypes = [vp, i32, vp, f64, vp, f64, vp, vp]
    L.mbx_colblock_sddmm.argtypes = [vp, i32, vp, vp, vp, vp]
    L.mbx_colblock_values_to_csr.argtypes = [vp, i32, vp, vp, vp]
    L.mbx_colblock_set_stream.argtypes = [vp, i32]
    L.mbx_colblock_set_residency.argtypes = [vp, i32]
    L.mbx_colblock_adjoint_prepare.argtypes = [vp, vp, vp, i32, i32, vp, ctypes.POINTER(i32)]
    L.mbx_colblock_dx.argtypes = [vp, i32, vp, f64, vp, f64, vp, vp]
    L.mbx_adjoint_create.argtypes = [vp, vp, i32, i32, i32, i32, vp, ctypes.POINTER(vp)]
    L.mbx_adjoint_destroy.argtypes = [vp]
    L.mbx_adjoint_destroy.restype = None
    L.mbx_adjoint_transpose.argtypes = [vp, vp, vp, vp]
    L.mbx_adjoint_dx.argtypes = [vp, i32, vp, f64, vp, f64, vp, vp]
    L.mbx_adjoint_set_blocks.argtypes = [vp, i32, vp]
    L.mbx_sddmm.argtypes = [vp, i32, vp, vp, vp, vp, vp, vp]
    L.mbx_ingest_ref.argtypes = [i32, i32, vp, vp, i32, vp, vp, vp, vp]
    L.mbx_plan_stats.argtypes = [vp, vp, ctypes.POINTER(PatternStats)]
    L.mbx_pagerank_workspace_bytes.argtypes = [i32]
    L.mbx_pagerank_workspace_bytes.restype = ctypes.c_size_t
    L.mbx_pagerank_f64.argtypes = [vp, vp, vp, vp, f64, i32, vp, vp, vp, ctypes.c_size_t, vp]
    L.mbx_pagerank_graph_create.argtypes = [vp, vp, vp, vp, f64, i32, vp, vp, vp,
                                            ctypes.c_size_t, ctypes.POINTER(vp)]
    L.mbx_csr_validate.argtypes = [vp, vp, i32, i32, i32, vp, ctypes.POINTER(CsrReport)]
    L.mbx_launch_rule.argtypes = [vp, vp, i32, i32, i32, i32, vp, ctypes.POINTER(LaunchChoice)]
    L.mbx_adjoint_auto.argtypes = [vp, i32, vp, ctypes.POINTER(LaunchChoice)]
    L.mbx_plan_create_auto.argtypes = [vp, vp, i32, i32, i32, i32, vp, ctypes.POINTER(vp),
                                       ctypes.POINTER(LaunchChoice)]
    L.mbx_plan_split_hot.argtypes = [vp, vp, i32, vp, ctypes.POINTER(f64)]
    L.mbx_spmv_f64_noalloc.argtypes = [vp, vp, vp, i32, i32, i32, f64, vp, f64, vp, vp]
    L.mbx_status_str.argtypes = [ctypes.c_int]
    L.mbx_status_str.restype = ctypes.c_char_p
    L.mbg_fill_uniform_f64.argtypes = [vp, i64, f64, f64, u64]
    L.mbg_fill_uniform_f32.argtypes = [vp, i64, f32, f32, u64]
    L.mbg_fill_normal_f64.argtypes = [vp, i64, u64]
    L.mbg_fill_normal_f32.argtypes = [vp, i64, u64]
    L.mbg_f32_to_bf16.argtypes = [vp, vp, i64]
    L.mbg_laplacian5_nnz.argtypes = [i32]
    L.mbg_laplacian5_nnz.restype = i64
    L.mbg_laplacian5.argtypes = [i32, vp, vp, vp]
    L.mbg_rmat.argtypes = [i32, i64, f64, f64, f64, u64, vp, vp]
    L.mbg_banded_nnz.argtypes = [i32, i32]
    L.mbg_banded_nnz.restype = i64
    L.mbg_banded.argtypes = [i32, i32, vp, vp]
    L.mbg_poisson_rows.argtypes = [i32, i32, f64, u64, vp, vp]
    L.mbh_config_lookup.argtypes = [ctypes.c_char_p, i32, ctypes.c_char_p, i32]
    L.mbh_mtx_read.argtypes = [ctypes.c_char_p, ctypes.POINTER(i32), ctypes.POINTER(i64),
                               vp, vp, vp]
    L.mbh_ref_ingest.argtypes = [i32, i64, vp, vp, vp, vp, vp]
    _lib = L
    return L


def status_str(status: int) -> str:
    return lib().mbx_status_str(int(status)).decode()


def check(status: int, what: str) -> None:
    if status != MBX_OK:
        raise MerbitError(status, what)
