"""Measured comparators for the benchmark (not used by the product path).

``RocsparseSpmv`` wraps rocSPARSE's CSR and COO SpMV (libmerbit_cmp.so, built from
rocsparse_cmp.cpp) on device tensors; the benchmark times it beside the
merge-path kernel on the same arrays, as the reference times its vendor
library rows (src/spmv.cu:6-64, 124-157)."""
from __future__ import annotations

import ctypes
import os

import torch

from .._lib import MBX_VALUE_F32, MBX_VALUE_F64, MerbitError

_PATH = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                     "libmerbit_cmp.so")
_ALGS = {"adaptive": 0, "lrb": 1, "rowsplit": 2, "nnzsplit": 3, "coo": 4, "coo_atomic": 5}
_L = None


def _lib():
    global _L
    if _L is None:
        if not os.path.exists(_PATH):
            raise ImportError(f"{_PATH} is not built; run `python -m merbit_amd.build`")
        L = ctypes.CDLL(_PATH)
        vp, i32 = ctypes.c_void_p, ctypes.c_int32
        L.mbc_rocsparse_create.argtypes = [i32, i32, i32, i32, i32, vp, vp, vp, vp, vp, vp,
                                           ctypes.POINTER(vp)]
        L.mbc_rocsparse_create.restype = ctypes.c_int
        L.mbc_rocsparse_spmv.argtypes = [vp]
        L.mbc_rocsparse_spmv.restype = ctypes.c_int
        L.mbc_rocsparse_buffer_bytes.argtypes = [vp]
        L.mbc_rocsparse_buffer_bytes.restype = ctypes.c_int64
        L.mbc_rocsparse_destroy.argtypes = [vp]
        L.mbc_rocsparse_destroy.restype = None
        _L = L
    return _L


class RocsparseSpmv:
    """y = A x by rocSPARSE SpMV (FP64 or FP32), preprocessed once.  ``alg``:
    CSR adaptive / lrb / rowsplit / nnzsplit, or COO coo / coo_atomic (COO
    row indices converted from row_ptr at construction)."""

    def __init__(self, row_ptr: torch.Tensor, col: torch.Tensor, val: torch.Tensor,
                 x: torch.Tensor, y: torch.Tensor, alg: str = "adaptive"):
        kind = {torch.float64: MBX_VALUE_F64, torch.float32: MBX_VALUE_F32}[val.dtype]
        self._keep = (row_ptr, col, val, x, y)
        self._h = ctypes.c_void_p()
        stream = torch.cuda.current_stream().cuda_stream
        st = _lib().mbc_rocsparse_create(kind, _ALGS[alg], row_ptr.numel() - 1, x.numel(),
                                         col.numel(), row_ptr.data_ptr(), col.data_ptr(),
                                         val.data_ptr(), x.data_ptr(), y.data_ptr(), stream,
                                         ctypes.byref(self._h))
        if st != 0:
            raise MerbitError(st, f"rocSPARSE {alg} create")
        self.alg = alg

    @property
    def buffer_bytes(self) -> int:
        return int(_lib().mbc_rocsparse_buffer_bytes(self._h))

    def __call__(self) -> None:
        st = _lib().mbc_rocsparse_spmv(self._h)
        if st != 0:
            raise MerbitError(st, f"rocSPARSE {self.alg} spmv")

    def close(self) -> None:
        if self._h:
            _lib().mbc_rocsparse_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_PROBE_PATH = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                           "libmerbit_probe.so")
_PL = None


def probe(plan, which: str, row_ptr, col, val, x, y, stream=None) -> None:
    """One bandwidth probe launch at the plan's tile shape (probes.hip):
    which = "stream" (row_ptr, col and values as the tile kernel stages them;
    y: nrow doubles) or "gather" (columns and x gathers; y: ntiles doubles)."""
    global _PL
    if _PL is None:
        if not os.path.exists(_PROBE_PATH):
            raise ImportError(f"{_PROBE_PATH} is not built; run `python -m merbit_amd.build`")
        L = ctypes.CDLL(_PROBE_PATH)
        vp = ctypes.c_void_p
        L.mbp_probe.argtypes = [vp, ctypes.c_int32, vp, vp, vp, vp, vp, vp]
        L.mbp_probe.restype = ctypes.c_int
        _PL = L
    s = torch.cuda.current_stream() if stream is None else stream
    st = _PL.mbp_probe(plan.handle, {"stream": 0, "gather": 1}[which], row_ptr.data_ptr(),
                       col.data_ptr(), val.data_ptr() if val is not None else None,
                       x.data_ptr() if x is not None else None, y.data_ptr(),
                       int(s.cuda_stream))
    if st != 0:
        raise MerbitError(st, "mbp_probe")
