"""ctypes wrapper of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by ``__graft_entry__.smoke()``
and by the ``cpu_baseline`` leg of ``bench.py``; never by the product package
``merbit_amd``.  Each wrapper names the reference file:line it restates (paths
relative to the reference repository root); the C source documents how each
function is pinned.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_u32p = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
_u64p = np.ctypeslib.ndpointer(dtype=np.uint64, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")


def build(force: bool = False) -> str:
    """Compile liboracle.so with the Makefile beside this file."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-C", _HERE] + (["-B"] if force else []),
                       check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB_PATH):
        build()
    L = ctypes.CDLL(_LIB_PATH)
    i32, i64, f64 = ctypes.c_int32, ctypes.c_int64, ctypes.c_double
    L.orc_merge_search.argtypes = [_i32p, i32, i32, i64,
                                   ctypes.POINTER(i32), ctypes.POINTER(i32)]
    L.orc_merge_search.restype = None
    L.orc_ref_lane_count.argtypes = [i32, i32, i32]
    L.orc_ref_lane_count.restype = i32
    L.orc_ref_plan.argtypes = [_i32p, i32, i32, i32, _u32p, _u32p, _u64p]
    L.orc_ref_plan.restype = i32
    L.orc_spmv_csr_f64.argtypes = [i32, _i32p, _i32p, _f64p, f64, _f64p, f64, _f64p]
    L.orc_spmv_csr_f64.restype = None
    L.orc_spmv_accurate.argtypes = [i32, _i32p, _i32p, ctypes.c_void_p, i32,
                                    ctypes.c_void_p, i32, _f64p, ctypes.c_void_p]
    L.orc_spmv_accurate.restype = None
    L.orc_merbit_emulate.argtypes = [i32, i32, _i32p, _i32p, _f64p, f64, _f64p,
                                     i32, _f64p]
    L.orc_merbit_emulate.restype = i32
    L.orc_ref_ingest.argtypes = [i32, i32, _i32p, _i32p, _i32p, _i32p, _f64p]
    L.orc_ref_ingest.restype = None
    L.orc_csr_transpose.argtypes = [i32, i32, i32, _i32p, _i32p, _i32p, _i32p, _i32p]
    L.orc_csr_transpose.restype = None
    L.orc_spmv_csr_omp_time.argtypes = [i32, _i32p, _i32p, _f64p, _f64p, _f64p,
                                        i32, i32, i32, ctypes.POINTER(i32)]
    L.orc_spmv_csr_omp_time.restype = f64
    _lib = L
    return L


def _c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


def merge_search(row_ptr, nnz: int, diag: int):
    """include/utils.cuh:5-29 -> (coord_x, coord_y)."""
    rp = _c(row_ptr, np.int32)
    cx, cy = ctypes.c_int32(), ctypes.c_int32()
    lib().orc_merge_search(rp, len(rp) - 1, int(nnz), int(diag),
                           ctypes.byref(cx), ctypes.byref(cy))
    return cx.value, cy.value


def ref_plan(row_ptr, nnz: int, sigma: int):
    """include/tile.cuh:28-122 -> (tile_x, tile_y, lane_desc)."""
    rp = _c(row_ptr, np.int32)
    nrow = len(rp) - 1
    lanes = lib().orc_ref_lane_count(nrow, int(nnz), int(sigma))
    if lanes < 0:
        raise ValueError("sigma out of range")
    tiles = (lanes + 31) // 32
    tx = np.zeros(tiles + 1, np.uint32)
    ty = np.zeros(tiles + 1, np.uint32)
    desc = np.zeros(max(lanes, 1), np.uint64)
    lib().orc_ref_plan(rp, nrow, int(nnz), int(sigma), tx, ty, desc)
    return tx, ty, desc[:lanes]


def spmv_csr(row_ptr, col, val, x, alpha=1.0, beta=0.0, y=None):
    """Row-sequential FP64 CSR y = alpha*A*x + beta*y."""
    rp = _c(row_ptr, np.int32)
    nrow = len(rp) - 1
    out = np.zeros(nrow) if y is None else _c(y, np.float64).copy()
    lib().orc_spmv_csr_f64(nrow, rp, _c(col, np.int32), _c(val, np.float64),
                           float(alpha), _c(x, np.float64), float(beta), out)
    return out


_VTYPE = {np.dtype(np.float64): 0, np.dtype(np.float32): 1, np.dtype(np.uint16): 2}
_XTYPE = {np.dtype(np.float64): 0, np.dtype(np.float32): 1}


def spmv_accurate(row_ptr, col, val, x):
    """Dot2-accurate y = A*x and |A||x| (values: f64, f32 or bf16 bits as u16)."""
    rp = _c(row_ptr, np.int32)
    nrow = len(rp) - 1
    v = np.ascontiguousarray(val)
    xv = np.ascontiguousarray(x)
    if v.dtype not in _VTYPE or xv.dtype not in _XTYPE:
        raise TypeError("unsupported dtype")
    y = np.zeros(nrow)
    absr = np.zeros(nrow)
    lib().orc_spmv_accurate(nrow, rp, _c(col, np.int32), v.ctypes.data,
                            _VTYPE[v.dtype], xv.ctypes.data, _XTYPE[xv.dtype],
                            y, absr.ctypes.data)
    return y, absr


def merbit_emulate(row_ptr, col, val, x, sigma: int = 14, alpha: float = 1.0):
    """One merbit_kernel launch, reference summation grouping (tile.cuh:476-561)."""
    rp = _c(row_ptr, np.int32)
    nrow = len(rp) - 1
    y = np.zeros(nrow)
    rc = lib().orc_merbit_emulate(nrow, int(rp[-1]), rp, _c(col, np.int32),
                                  _c(val, np.float64), float(alpha),
                                  _c(x, np.float64), int(sigma), y)
    if rc != 0:
        raise ValueError("sigma out of range")
    return y


def ref_ingest(nrow: int, row_ind, col_ind):
    """src/class.cu:22-34,143-177 + tile.cuh:571-574: CSR of A^T, 1/outdeg values."""
    ri, ci = _c(row_ind, np.int32), _c(col_ind, np.int32)
    nnz = len(ri)
    rp = np.zeros(nrow + 1, np.int32)
    col = np.zeros(max(nnz, 1), np.int32)
    val = np.zeros(max(nnz, 1))
    lib().orc_ref_ingest(nrow, nnz, ri, ci, rp, col, val)
    return rp, col[:nnz], val[:nnz]


def csr_transpose(row_ptr, col, ncol: int):
    rp = _c(row_ptr, np.int32)
    nrow = len(rp) - 1
    nnz = int(rp[-1])
    cp = np.zeros(ncol + 1, np.int32)
    ri = np.zeros(max(nnz, 1), np.int32)
    perm = np.zeros(max(nnz, 1), np.int32)
    lib().orc_csr_transpose(nrow, ncol, nnz, rp, _c(col, np.int32), cp, ri, perm)
    return cp, ri[:nnz], perm[:nnz]


def pagerank(row_ptr, col, val, damp: float, iters: int, x0):
    """Damped power iterations on the CSR of P^T (values 1/out-degree), the
    use of PARAMETER::damp (include/class.cuh:119-146) that the reference's
    loop leaves out: y = damp*A*x (Dot2-accurate rows), x' = y + (sum(x) -
    sum(y))/n with correctly rounded sums (math.fsum).  Pinned in the tests
    by the fixed point from a dense linear solve and by mass conservation."""
    import math
    x = np.ascontiguousarray(x0, dtype=np.float64).copy()
    n = len(x)
    for _ in range(int(iters)):
        y = damp * spmv_accurate(row_ptr, col, val, x)[0]
        x = y + (math.fsum(x) - math.fsum(y)) / n
    return x


def spmv_csr_omp_time(row_ptr, col, val, x, nthreads: int, warmup=3, reps=20):
    """Median seconds of the OpenMP FP64 CSR baseline and threads used."""
    rp = _c(row_ptr, np.int32)
    nrow = len(rp) - 1
    y = np.zeros(nrow)
    used = ctypes.c_int32()
    t = lib().orc_spmv_csr_omp_time(nrow, rp, _c(col, np.int32),
                                    _c(val, np.float64), _c(x, np.float64), y,
                                    int(nthreads), int(warmup), int(reps),
                                    ctypes.byref(used))
    return t, used.value, y
