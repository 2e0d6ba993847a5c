"""Batches of matrices, HIP-graph replay and the timing autotuner (Python
front end of the batch / graph / autotune entry points of merbit_amd.h)."""
from __future__ import annotations

import ctypes

import torch

from ._lib import (MBX_VALUE_BF16, MBX_VALUE_F32, MBX_VALUE_F64, BatchItem, check, lib)
from .spmv import _dptr, _stream_handle

_KINDS = {torch.float64: MBX_VALUE_F64, torch.float32: MBX_VALUE_F32,
          torch.uint16: MBX_VALUE_BF16, torch.bfloat16: MBX_VALUE_BF16}


def value_kind(val: torch.Tensor) -> int:
    if val.dtype not in _KINDS:
        raise TypeError(f"unsupported value dtype {val.dtype}")
    return _KINDS[val.dtype]


def _vec_dtype(kind: int):
    return torch.float64 if kind == MBX_VALUE_F64 else torch.float32


class Graph:
    """An executable HIP graph of repeated products (hipGraphExec_t)."""

    def __init__(self, handle: ctypes.c_void_p, keep=()):
        self._h = handle
        self._keep = keep  # operands the graph points at stay alive with it

    def launch(self, stream=None) -> None:
        check(lib().mbx_graph_launch(self._h, _stream_handle(stream)), "mbx_graph_launch")

    def close(self) -> None:
        if getattr(self, "_h", None) is not None:
            lib().mbx_graph_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _items(mats):
    """mats: sequence of (row_ptr, col, val, x, y, ncol) device tensors."""
    kind = None
    arr = (BatchItem * len(mats))()
    for i, (rp, col, val, x, y, ncol) in enumerate(mats):
        k = value_kind(val)
        if kind is None:
            kind = k
        elif k != kind:
            raise TypeError("all matrices of a batch share one value kind")
        v = val.view(torch.uint16) if val.dtype == torch.bfloat16 else val
        vdt = _vec_dtype(k)
        nrow, nnz = rp.numel() - 1, col.numel()
        if val.numel() != nnz or x.numel() != ncol or y.numel() != nrow:
            raise ValueError(f"matrix {i}: operand sizes do not match")
        arr[i] = BatchItem(_dptr(rp, torch.int32, "row_ptr"), _dptr(col, torch.int32, "col"),
                           _dptr(v, v.dtype, "val"), _dptr(x, vdt, "x"), _dptr(y, vdt, "y"),
                           nrow, int(ncol), nnz)
    return arr, kind


class MerbitBatch:
    """Many CSR matrices computed in one persistent launch."""

    def __init__(self, mats, items_per_thread: int = 8, stream=None):
        self._arr, self.kind = _items(mats)
        self._mats = list(mats)
        h = ctypes.c_void_p()
        check(lib().mbx_batch_create(self._arr, len(self._arr), self.kind, items_per_thread,
                                     _stream_handle(stream), ctypes.byref(h)),
              "mbx_batch_create")
        self._h = h
        tt, ti, gr = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
        check(lib().mbx_batch_info(h, ctypes.byref(tt), ctypes.byref(ti), ctypes.byref(gr)),
              "mbx_batch_info")
        self.total_tiles, self.tile_items, self.grid = tt.value, ti.value, gr.value
        self.items_per_thread = items_per_thread

    def set_stream(self, mode: str) -> None:
        """Cache policy of the batch's stream loads (mbx_batch_set_stream):
        "auto", "cached" or "nontemporal" (MerbitPlan.set_stream)."""
        from .spmv import STREAM_MODES
        check(lib().mbx_batch_set_stream(self._h, STREAM_MODES[mode]), "mbx_batch_set_stream")

    def set_residency(self, wg_per_cu: int) -> None:
        """Tiles resident per CU for the batch's launches
        (mbx_batch_set_residency); 0 = as many as fit."""
        check(lib().mbx_batch_set_residency(self._h, int(wg_per_cu)), "mbx_batch_set_residency")

    def sort_cols(self, enable: bool = True, stream=None) -> bool:
        """Tile-sorted gathers for every member (mbx_batch_sort_cols); False
        when the batch's shape or a member rules them out.  Results are
        bit-identical either way."""
        st = lib().mbx_batch_sort_cols(self._h, int(bool(enable)), _stream_handle(stream))
        if st == -3:  # MBX_EUNSUPPORTED
            return False
        check(st, "mbx_batch_sort_cols")
        return bool(enable)

    def spmv(self, alpha: float = 1.0, beta: float = 0.0, stream=None) -> None:
        check(lib().mbx_batch_spmv(self._h, alpha, beta, _stream_handle(stream)),
              "mbx_batch_spmv")

    def graph(self, alpha: float = 1.0, beta: float = 0.0, repeats: int = 1) -> Graph:
        g = ctypes.c_void_p()
        check(lib().mbx_batch_graph_create(self._h, alpha, beta, repeats, ctypes.byref(g)),
              "mbx_batch_graph_create")
        return Graph(g, keep=(self,))

    def close(self) -> None:
        if getattr(self, "_h", None) is not None:
            lib().mbx_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def spmv_graph(plan, row_ptr, col, val, x, y, alpha=1.0, beta=0.0, repeats=1) -> Graph:
    """Capture `repeats` products y = alpha*A*x + beta*y on `plan`."""
    kind = value_kind(val)
    v = val.view(torch.uint16) if val.dtype == torch.bfloat16 else val
    vdt = _vec_dtype(kind)
    g = ctypes.c_void_p()
    check(lib().mbx_spmv_graph_create(plan.handle, kind, _dptr(row_ptr, torch.int32, "row_ptr"),
                                      _dptr(col, torch.int32, "col"), _dptr(v, v.dtype, "val"),
                                      alpha, _dptr(x, vdt, "x"), beta, _dptr(y, vdt, "y"),
                                      repeats, ctypes.byref(g)), "mbx_spmv_graph_create")
    return Graph(g, keep=(plan, row_ptr, col, val, x, y))


def autotune_spmv(row_ptr, col, val, x, y, ncol=None, reps: int = 10, stream=None):
    """(best items_per_thread, microseconds per product) for this matrix."""
    kind = value_kind(val)
    v = val.view(torch.uint16) if val.dtype == torch.bfloat16 else val
    vdt = _vec_dtype(kind)
    ncol = x.numel() if ncol is None else ncol
    ipt, us = ctypes.c_int32(), ctypes.c_float()
    check(lib().mbx_autotune_spmv(_dptr(row_ptr, torch.int32, "row_ptr"), row_ptr.numel() - 1,
                                  ncol, col.numel(), _dptr(col, torch.int32, "col"),
                                  _dptr(v, v.dtype, "val"), kind, _dptr(x, vdt, "x"),
                                  _dptr(y, vdt, "y"), reps, _stream_handle(stream),
                                  ctypes.byref(ipt), ctypes.byref(us)), "mbx_autotune_spmv")
    return ipt.value, us.value


def autotune_spmv_launch(row_ptr, col, val, x, y, ncol=None, reps: int = 10, stream=None):
    """Best launch shape for this matrix (mbx_autotune_spmv_launch):
    (items_per_thread, stream mode name, tiles per CU, microseconds)."""
    kind = value_kind(val)
    v = val.view(torch.uint16) if val.dtype == torch.bfloat16 else val
    vdt = _vec_dtype(kind)
    ncol = x.numel() if ncol is None else ncol
    ipt, mode, wg, us = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32(), ctypes.c_float()
    check(lib().mbx_autotune_spmv_launch(
        _dptr(row_ptr, torch.int32, "row_ptr"), row_ptr.numel() - 1, ncol, col.numel(),
        _dptr(col, torch.int32, "col"), _dptr(v, v.dtype, "val"), kind, _dptr(x, vdt, "x"),
        _dptr(y, vdt, "y"), reps, _stream_handle(stream), ctypes.byref(ipt), ctypes.byref(mode),
        ctypes.byref(wg), ctypes.byref(us)), "mbx_autotune_spmv_launch")
    from .spmv import STREAM_MODES
    name = {v: k for k, v in STREAM_MODES.items()}[mode.value]
    return ipt.value, name, wg.value, us.value


def autotune_spmv_form(row_ptr, col, val, x, y, ncol=None, reps: int = 10, stream=None):
    """As autotune_spmv_launch with tile-sorted gathers as a further
    candidate (mbx_autotune_spmv_form): (items_per_thread, stream mode
    name, tiles per CU, sorted, microseconds)."""
    kind = value_kind(val)
    v = val.view(torch.uint16) if val.dtype == torch.bfloat16 else val
    vdt = _vec_dtype(kind)
    ncol = x.numel() if ncol is None else ncol
    ipt, mode, wg, srt = (ctypes.c_int32() for _ in range(4))
    us = ctypes.c_float()
    check(lib().mbx_autotune_spmv_form(
        _dptr(row_ptr, torch.int32, "row_ptr"), row_ptr.numel() - 1, ncol, col.numel(),
        _dptr(col, torch.int32, "col"), _dptr(v, v.dtype, "val"), kind, _dptr(x, vdt, "x"),
        _dptr(y, vdt, "y"), reps, _stream_handle(stream), ctypes.byref(ipt), ctypes.byref(mode),
        ctypes.byref(wg), ctypes.byref(srt), ctypes.byref(us)), "mbx_autotune_spmv_form")
    from .spmv import STREAM_MODES
    name = {v: k for k, v in STREAM_MODES.items()}[mode.value]
    return ipt.value, name, wg.value, bool(srt.value), us.value


def autotune_spmm(row_ptr, col, val, X, Y, reps: int = 5, stream=None):
    """(best items_per_thread, microseconds per product) for this SpMM."""
    from .spmv import MerbitPlan
    xl, ldx = MerbitPlan._layout(X, "X")
    yl, ldy = MerbitPlan._layout(Y, "Y")
    k = int(X.shape[1])
    nrow, ncol, nnz = row_ptr.numel() - 1, int(X.shape[0]), col.numel()
    probe = MerbitPlan(row_ptr, ncol, nnz, items_per_thread=2)
    need = lib().mbx_spmm_workspace_bytes(probe.handle, k, X.data_ptr(), ldx, xl)
    probe.close()
    ws = torch.empty(max(int(need), 1), dtype=torch.uint8, device=X.device)
    ipt, us = ctypes.c_int32(), ctypes.c_float()
    check(lib().mbx_autotune_spmm(_dptr(row_ptr, torch.int32, "row_ptr"), nrow, ncol, nnz,
                                  _dptr(col, torch.int32, "col"),
                                  _dptr(val, torch.float32, "val"), k, X.data_ptr(), ldx, xl,
                                  Y.data_ptr(), ldy, yl, ws.data_ptr(), ws.numel(), reps,
                                  _stream_handle(stream), ctypes.byref(ipt), ctypes.byref(us)),
          "mbx_autotune_spmm")
    return ipt.value, us.value


def autotune_batch(mats, reps: int = 5, stream=None):
    arr, kind = _items(mats)
    ipt, us = ctypes.c_int32(), ctypes.c_float()
    check(lib().mbx_autotune_batch(arr, len(arr), kind, reps, _stream_handle(stream),
                                   ctypes.byref(ipt), ctypes.byref(us)), "mbx_autotune_batch")
    return ipt.value, us.value
