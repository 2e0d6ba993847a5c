"""Python mirror of the merge-path SpMV host API over the C ABI.

``MerbitPlan`` plays the reference's ``TILE<BLOCK,SIGMA>`` plan object
(include/tile.cuh:6-25,106-122) and ``MerbitPlan.spmv`` its
``merbit_kernel`` launch (tile.cuh:476-561, host loop 596-603).  Tensors are
torch device tensors; the call runs on torch's current stream.  There is no
CPU path: every call goes through libmerbit_amd.so.
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from ._lib import check, lib

# Stream cache policies of mbx_plan_set_stream / mbx_batch_set_stream.
STREAM_MODES = {"auto": 0, "cached": 1, "nontemporal": 2}

# Accumulation modes of mbx_plan_set_accumulation (FP32 / bfloat16 values).
ACCUMULATION_MODES = {"compensated": 0, "f64": 1, "f32prod": 2}
ACCUMULATION_NAMES = {v: k for k, v in ACCUMULATION_MODES.items()}

# Items per thread with an instantiated tile kernel (the reference's SIGMA;
# the tiles of 32 hold 8192 merge items and run as 512 threads of 16).
ITEMS_PER_THREAD = (2, 4, 6, 8, 10, 12, 14, 16, 32)

_VALUE_KINDS = {torch.float64: "f64", torch.float32: "f32", torch.uint16: "bf16",
                torch.bfloat16: "bf16"}


def _stream_handle(stream) -> int:
    if stream is None:
        stream = torch.cuda.current_stream()
    return int(stream.cuda_stream)


def _dptr(t: torch.Tensor, dtype, name: str) -> int:
    if not t.is_cuda:
        raise ValueError(f"{name} must be a device tensor")
    if t.dtype != dtype:
        raise TypeError(f"{name}: expected {dtype}, got {t.dtype}")
    if not t.is_contiguous():
        raise ValueError(f"{name} must be contiguous")
    return t.data_ptr()


def validate_csr(row_ptr: torch.Tensor, col: torch.Tensor, ncol: int, stream=None) -> dict:
    """Structural check of a device CSR pattern (mbx_csr_validate): counts of
    row_ptr errors, out-of-range columns (with the first such position),
    descending column pairs within rows and whether row_ptr[nrow] != nnz,
    plus ``ok`` (no error but possibly unsorted columns)."""
    from ._lib import CsrReport
    nrow = int(row_ptr.numel()) - 1
    nnz = int(col.numel())
    rep = CsrReport()
    st = lib().mbx_csr_validate(_dptr(row_ptr, torch.int32, "row_ptr"),
                                _dptr(col, torch.int32, "col") if nnz else None, nrow,
                                int(ncol), nnz, _stream_handle(stream), ctypes.byref(rep))
    if st not in (0, -1):
        check(st, "mbx_csr_validate")
    out = {name: getattr(rep, name) for name, _ in CsrReport._fields_}
    out["ok"] = st == 0
    return out


_KIND_OF_DTYPE = {torch.float64: 0, torch.float32: 1, torch.bfloat16: 2, torch.uint16: 2}


def _choice_dict(ch) -> dict:
    from ._lib import RULE_NAMES, LaunchChoice
    out = {name: getattr(ch, name) for name, _ in LaunchChoice._fields_}
    out["rule"] = RULE_NAMES.get(out["rule"], out["rule"])
    return out


def launch_rule(row_ptr: torch.Tensor, col: torch.Tensor, ncol: int,
                value_dtype=torch.float64, stream=None) -> dict:
    """The launch the pattern's statistics choose (mbx_launch_rule): items
    per thread, tiles per CU, accumulation, 16-bit columns, hot columns and
    a column-blocked recommendation, with the features that decided."""
    from ._lib import LaunchChoice
    nrow, nnz = int(row_ptr.numel()) - 1, int(col.numel())
    ch = LaunchChoice()
    check(lib().mbx_launch_rule(_dptr(row_ptr, torch.int32, "row_ptr"),
                                _dptr(col, torch.int32, "col") if nnz else None, nrow, int(ncol),
                                nnz, _KIND_OF_DTYPE[value_dtype], _stream_handle(stream),
                                ctypes.byref(ch)), "mbx_launch_rule")
    return _choice_dict(ch)


class MerbitPlan:
    """Merge-path partition of one CSR pattern (``TILE`` of the reference).

    items_per_thread None (the default) with the column array ``col``: the
    launch rule (mbx_plan_create_auto) picks the plan's shape and options
    from the pattern's statistics for values of ``value_dtype``; the choice
    is kept in ``self.choice``.  Without ``col`` the default is 8 items per
    thread.  row_tiles=True with an explicit items_per_thread: row-aligned
    tiles (mbx_plan_create_rows; no row crosses a tile boundary)."""

    def __init__(self, row_ptr: torch.Tensor, ncol: int, nnz: int | None = None,
                 items_per_thread: int | None = None, wg_size: int = 256, stream=None,
                 col: torch.Tensor | None = None, value_dtype=torch.float64,
                 row_tiles: bool = False):
        self.nrow = int(row_ptr.numel()) - 1
        self.ncol = int(ncol)
        self.nnz = int(row_ptr[-1].item()) if nnz is None else int(nnz)
        self.choice = None
        h = ctypes.c_void_p()
        if items_per_thread is None and col is not None and wg_size == 256:
            from ._lib import LaunchChoice
            ch = LaunchChoice()
            check(lib().mbx_plan_create_auto(
                _dptr(row_ptr, torch.int32, "row_ptr"),
                _dptr(col, torch.int32, "col") if self.nnz else None, self.nrow, self.ncol,
                self.nnz, _KIND_OF_DTYPE[value_dtype], _stream_handle(stream), ctypes.byref(h),
                ctypes.byref(ch)), "mbx_plan_create_auto")
            self.choice = _choice_dict(ch)
        else:
            create = lib().mbx_plan_create_rows if row_tiles else lib().mbx_plan_create
            check(create(_dptr(row_ptr, torch.int32, "row_ptr"), self.nrow, self.ncol,
                         self.nnz, items_per_thread or 8, wg_size, _stream_handle(stream),
                         ctypes.byref(h)),
                  "mbx_plan_create_rows" if row_tiles else "mbx_plan_create")
        self._h = h
        nt, ti, wg = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
        check(lib().mbx_plan_info(h, ctypes.byref(nt), ctypes.byref(ti), ctypes.byref(wg)),
              "mbx_plan_info")
        self.ntiles, self.tile_items, self.wg_size = nt.value, ti.value, wg.value
        self.items_per_thread = self.tile_items // self.wg_size

    @property
    def handle(self) -> ctypes.c_void_p:
        if self._h is None:
            raise RuntimeError("plan is closed")
        return self._h

    def partition(self):
        """Tile start coordinates (rows consumed, nnz consumed), ntiles+1 each."""
        n = self.ntiles + 1
        rows = np.empty(n, np.int32)
        nnz = np.empty(n, np.int32)
        check(lib().mbx_plan_partition(self.handle, rows.ctypes.data, nnz.ctypes.data, n),
              "mbx_plan_partition")
        return rows, nnz

    def stats(self, row_ptr: torch.Tensor) -> dict:
        """Pattern statistics (mbx_plan_stats): row-length moments, empty rows,
        tile statistics of this plan's shape."""
        from ._lib import PatternStats
        st = PatternStats()
        check(lib().mbx_plan_stats(self.handle, _dptr(row_ptr, torch.int32, "row_ptr"),
                                   ctypes.byref(st)), "mbx_plan_stats")
        return {name: getattr(st, name) for name, _ in PatternStats._fields_}

    def set_stream(self, mode: str) -> None:
        """Cache policy of the row_ptr/col/value stream (mbx_plan_set_stream):
        "auto" (non-temporal past 240 MiB per product), "cached" (plain
        loads: the matrix can stay in the MALL between products) or
        "nontemporal"."""
        check(lib().mbx_plan_set_stream(self.handle, STREAM_MODES[mode]), "mbx_plan_set_stream")

    def set_residency(self, wg_per_cu: int) -> None:
        """Tiles resident per CU for this plan's launches
        (mbx_plan_set_residency); 0 = as many as fit."""
        check(lib().mbx_plan_set_residency(self.handle, int(wg_per_cu)),
              "mbx_plan_set_residency")

    def set_accumulation(self, mode: str) -> None:
        """Sums of the FP32 / bfloat16 value forms: "compensated" (TwoSum
        compensated FP32, the default), "f64" (exact FP64 products, FP64
        sums) or "f32prod" (products rounded once to FP32, FP64 sums: half
        the LDS per tile)."""
        check(lib().mbx_plan_set_accumulation(self.handle, ACCUMULATION_MODES[mode]),
              "mbx_plan_set_accumulation")

    def compress_cols(self, col: torch.Tensor | None, stream=None) -> bool:
        """Keep the columns as 16-bit offsets from each tile's base
        (mbx_plan_compress_cols); False when a tile spans too many columns
        or the plan's items per thread is not 8 or 12.  None drops them."""
        ptr = None if col is None else _dptr(col, torch.int32, "col")
        st = lib().mbx_plan_compress_cols(self.handle, ptr, _stream_handle(stream))
        if st == -3:  # MBX_EUNSUPPORTED
            return False
        check(st, "mbx_plan_compress_cols")
        return col is not None

    def sort_cols(self, col: torch.Tensor | None, stream=None) -> bool:
        """Keep every tile's columns sorted in gather order
        (mbx_plan_sort_cols), so the x gathers share lines; False when the
        plan's shape, ncol or another column option rules it out.  None
        drops the copy.  Results are bit-identical either way."""
        ptr = None if col is None else _dptr(col, torch.int32, "col")
        st = lib().mbx_plan_sort_cols(self.handle, ptr, _stream_handle(stream))
        if st == -3:  # MBX_EUNSUPPORTED
            return False
        check(st, "mbx_plan_sort_cols")
        return col is not None

    def split_hot(self, col: torch.Tensor | None, max_hot: int, stream=None) -> float:
        """Keep the max_hot most frequent columns' x values in LDS during
        products (mbx_plan_split_hot); returns the share of the nonzeros in
        those columns.  None or 0 drops the split."""
        cov = ctypes.c_double()
        ptr = None if col is None else _dptr(col, torch.int32, "col")
        check(lib().mbx_plan_split_hot(self.handle, ptr, int(max_hot), _stream_handle(stream),
                                       ctypes.byref(cov)), "mbx_plan_split_hot")
        return cov.value

    def set_events(self, begin, end) -> None:
        """Attach a torch.cuda.Event pair (timing enabled, recorded once so
        that the underlying events exist) to the next eager call's tile
        kernel: their elapsed time is that kernel's own execution."""
        check(lib().mbx_plan_set_events(self.handle, begin.cuda_event, end.cuda_event),
              "mbx_plan_set_events")

    def spmv(self, row_ptr: torch.Tensor, col: torch.Tensor, val: torch.Tensor,
             x: torch.Tensor, y: torch.Tensor | None = None, alpha: float = 1.0,
             beta: float = 0.0, stream=None) -> torch.Tensor:
        """y = alpha*A*x + beta*y; FP64, FP32 or bfloat16 values by val.dtype."""
        kind = _VALUE_KINDS.get(val.dtype)
        if kind is None:
            raise TypeError(f"unsupported value dtype {val.dtype}")
        if row_ptr.numel() != self.nrow + 1 or col.numel() != self.nnz or \
                val.numel() != self.nnz:
            raise ValueError("CSR arrays do not match the plan")
        xdt = torch.float64 if kind == "f64" else torch.float32
        if x.numel() != self.ncol:
            raise ValueError("x has the wrong length")
        if y is None:
            y = torch.empty(self.nrow, dtype=xdt, device=x.device)
            beta = 0.0
        if y.numel() != self.nrow:
            raise ValueError("y has the wrong length")
        args = (self.handle, _dptr(row_ptr, torch.int32, "row_ptr"),
                _dptr(col, torch.int32, "col"))
        s = _stream_handle(stream)
        if kind == "f64":
            st = lib().mbx_spmv_f64(*args, _dptr(val, torch.float64, "val"), alpha,
                                    _dptr(x, xdt, "x"), beta, _dptr(y, xdt, "y"), s)
        elif kind == "f32":
            st = lib().mbx_spmv_f32(*args, _dptr(val, torch.float32, "val"), alpha,
                                    _dptr(x, xdt, "x"), beta, _dptr(y, xdt, "y"), s)
        else:
            v = val.view(torch.uint16) if val.dtype == torch.bfloat16 else val
            st = lib().mbx_spmv_bf16(*args, _dptr(v, torch.uint16, "val"), alpha,
                                     _dptr(x, xdt, "x"), beta, _dptr(y, xdt, "y"), s)
        check(st, f"mbx_spmv_{kind}")
        return y

    def pagerank(self, row_ptr: torch.Tensor, col: torch.Tensor, val: torch.Tensor,
                 damp: float = 0.85, iters: int = 20, x0: torch.Tensor | None = None,
                 y: torch.Tensor | None = None, stream=None) -> torch.Tensor:
        """PageRank power iterations (mbx_pagerank_f64) on this square plan's
        FP64 matrix, the CSR of P^T with values 1/out-degree
        (PARAMETER::damp of include/class.cuh:119-146 put to use):
        y = damp*A*x, x' = y + (sum(x) - sum(y))/n, `iters` times from x0
        (default uniform 1/n).  Returns the last iterate."""
        if self.nrow != self.ncol:
            raise ValueError("PageRank needs a square matrix")
        dev = row_ptr.device
        if x0 is None:
            x0 = torch.full((self.nrow,), 1.0 / max(self.nrow, 1), dtype=torch.float64,
                            device=dev)
        if x0.numel() != self.nrow or col.numel() != self.nnz or val.numel() != self.nnz:
            raise ValueError("arrays do not match the plan")
        if y is None:
            y = torch.empty(self.nrow, dtype=torch.float64, device=dev)
        nbytes = int(lib().mbx_pagerank_workspace_bytes(self.nrow))
        work = torch.empty(nbytes, dtype=torch.uint8, device=dev)
        check(lib().mbx_pagerank_f64(self.handle, _dptr(row_ptr, torch.int32, "row_ptr"),
                                     _dptr(col, torch.int32, "col"),
                                     _dptr(val, torch.float64, "val"), float(damp), int(iters),
                                     _dptr(x0, torch.float64, "x0"),
                                     _dptr(y, torch.float64, "y"), work.data_ptr(), nbytes,
                                     _stream_handle(stream)), "mbx_pagerank_f64")
        return y

    @staticmethod
    def _layout(t: torch.Tensor, name: str):
        """(layout, leading dimension) of a 2-D tensor viewed as n x k."""
        from ._lib import MBX_COL_MAJOR, MBX_ROW_MAJOR
        if t.dim() != 2:
            raise ValueError(f"{name} must be 2-D")
        st = t.stride()
        if st[1] == 1 and (t.shape[0] <= 1 or st[0] >= t.shape[1]):
            return MBX_ROW_MAJOR, st[0] if t.shape[0] > 1 else t.shape[1]
        if st[0] == 1 and (t.shape[1] <= 1 or st[1] >= t.shape[0]):
            return MBX_COL_MAJOR, st[1] if t.shape[1] > 1 else t.shape[0]
        raise ValueError(f"{name} must be row- or column-major")

    def spmm(self, row_ptr: torch.Tensor, col: torch.Tensor, val: torch.Tensor,
             X: torch.Tensor, Y: torch.Tensor | None = None, alpha: float = 1.0,
             beta: float = 0.0, workspace: torch.Tensor | None = None,
             stream=None) -> torch.Tensor:
        """Y = alpha*A*X + beta*Y for FP32 X (ncol x k) and Y (nrow x k), each
        row- or column-major (read from the strides).  Y defaults to a new
        column-major tensor."""
        if X.dtype != torch.float32 or val.dtype != torch.float32:
            raise TypeError("spmm takes FP32 values and right-hand sides")
        if X.shape[0] != self.ncol:
            raise ValueError("X has the wrong number of rows")
        k = int(X.shape[1])
        if Y is None:
            Y = torch.empty((k, self.nrow), dtype=torch.float32, device=X.device).t()
            beta = 0.0
        if tuple(Y.shape) != (self.nrow, k):
            raise ValueError("Y has the wrong shape")
        xl, ldx = self._layout(X, "X")
        yl, ldy = self._layout(Y, "Y")
        need = lib().mbx_spmm_workspace_bytes(self.handle, k, X.data_ptr(), ldx, xl)
        if workspace is None or workspace.numel() < need:
            workspace = torch.empty(max(need, 1), dtype=torch.uint8, device=X.device)
        check(lib().mbx_spmm_f32(self.handle, _dptr(row_ptr, torch.int32, "row_ptr"),
                                 _dptr(col, torch.int32, "col"),
                                 _dptr(val, torch.float32, "val"), k, alpha, X.data_ptr(),
                                 ldx, xl, beta, Y.data_ptr(), ldy, yl,
                                 workspace.data_ptr(), workspace.numel(),
                                 _stream_handle(stream)), "mbx_spmm_f32")
        return Y

    def spmm_workspace_bytes(self, X: torch.Tensor) -> int:
        xl, ldx = self._layout(X, "X")
        return int(lib().mbx_spmm_workspace_bytes(self.handle, int(X.shape[1]), X.data_ptr(),
                                                  ldx, xl))

    def clone(self) -> "MerbitPlan":
        """A plan object on the same partition with carry slots of its own
        (mbx_plan_clone): products on a plan and on its clones may run
        concurrently on different streams when their y differ."""
        h = ctypes.c_void_p()
        check(lib().mbx_plan_clone(self.handle, ctypes.byref(h)), "mbx_plan_clone")
        c = MerbitPlan.__new__(MerbitPlan)
        c.nrow, c.ncol, c.nnz = self.nrow, self.ncol, self.nnz
        c.ntiles, c.tile_items, c.wg_size = self.ntiles, self.tile_items, self.wg_size
        c.items_per_thread = self.items_per_thread
        c._h = h
        return c

    def close(self) -> None:
        if getattr(self, "_h", None) is not None:
            lib().mbx_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DeviceCsr:
    """A CSR matrix resident in HBM with its plan (GRAPH + TILE together)."""

    def __init__(self, A, device="cuda", value_dtype=None, items_per_thread: int | None = None,
                 wg_size: int = 256):
        """items_per_thread None: the launch rule picks the plan (see MerbitPlan)."""
        self.nrow, self.ncol, self.nnz = A.nrow, A.ncol, A.nnz
        self.row_ptr = torch.from_numpy(np.ascontiguousarray(A.row_ptr)).to(device)
        self.col = torch.from_numpy(np.ascontiguousarray(A.col)).to(device)
        val = np.ascontiguousarray(A.val)
        self.val = torch.from_numpy(val).to(device)
        if value_dtype is not None and self.val.dtype != value_dtype:
            self.val = self.val.to(value_dtype)
        self.plan = MerbitPlan(self.row_ptr, self.ncol, self.nnz, items_per_thread, wg_size,
                               col=self.col, value_dtype=self.val.dtype)

    def spmv(self, x, y=None, alpha=1.0, beta=0.0, stream=None):
        return self.plan.spmv(self.row_ptr, self.col, self.val, x, y, alpha, beta, stream)


_KIND_CODES = {"f64": 0, "f32": 1, "bf16": 2}


class ColumnBlockedCsr:
    """Column-blocked form of one CSR pattern (mbx_colblock_*): the columns
    split into ``nblocks`` ranges, A = A_0 + ... + A_{B-1} with A_j keeping
    the columns of range j, and a product chaining the block products so each
    gathers from a slice of x that stays in one XCD's L2.  For patterns whose
    x gathers spread evenly over a long x (random graphs); values are moved
    into the block-major order once with ``values`` and reused."""

    def __init__(self, row_ptr: torch.Tensor, col: torch.Tensor, ncol: int,
                 nblocks: int = 0, items_per_thread: int = 8, stream=None):
        self.nrow = int(row_ptr.numel()) - 1
        self.ncol = int(ncol)
        self.nnz = int(col.numel())
        h = ctypes.c_void_p()
        check(lib().mbx_colblock_create(_dptr(row_ptr, torch.int32, "row_ptr"),
                                        _dptr(col, torch.int32, "col"), self.nrow, self.ncol,
                                        self.nnz, nblocks, items_per_thread,
                                        _stream_handle(stream), ctypes.byref(h)),
              "mbx_colblock_create")
        self._h = h
        nb, slots = ctypes.c_int32(), ctypes.c_int64()
        bounds = (ctypes.c_int32 * 17)()
        check(lib().mbx_colblock_info(h, ctypes.byref(nb), ctypes.byref(slots), bounds),
              "mbx_colblock_info")
        self.nblocks = nb.value
        self.value_slots = slots.value
        self.col_bounds = [bounds[j] for j in range(self.nblocks + 1)]

    @property
    def handle(self) -> ctypes.c_void_p:
        if self._h is None:
            raise RuntimeError("column-blocked pattern is closed")
        return self._h

    def values(self, val: torch.Tensor, out: torch.Tensor | None = None,
               stream=None) -> torch.Tensor:
        """``val`` (CSR order) moved into the block-major order."""
        kind = _VALUE_KINDS.get(val.dtype)
        if kind is None:
            raise TypeError(f"unsupported value dtype {val.dtype}")
        if val.numel() != self.nnz:
            raise ValueError("val does not match the pattern")
        if out is None:
            out = torch.empty(max(self.value_slots, 1), dtype=val.dtype, device=val.device)
        if out.numel() < self.value_slots or out.dtype != val.dtype:
            raise ValueError("out is too short or of another dtype")
        check(lib().mbx_colblock_values(self.handle, _KIND_CODES[kind],
                                        _dptr(val, val.dtype, "val"),
                                        _dptr(out, out.dtype, "out"), _stream_handle(stream)),
              "mbx_colblock_values")
        return out

    def values_to_csr(self, val_blocked: torch.Tensor, out: torch.Tensor | None = None,
                      stream=None) -> torch.Tensor:
        """Block-major values moved back into CSR order (nnz entries)."""
        kind = _VALUE_KINDS.get(val_blocked.dtype)
        if kind is None:
            raise TypeError(f"unsupported value dtype {val_blocked.dtype}")
        if val_blocked.numel() < self.value_slots:
            raise ValueError("val_blocked is shorter than the blocked pattern")
        if out is None:
            out = torch.empty(max(self.nnz, 1), dtype=val_blocked.dtype,
                              device=val_blocked.device)
        if out.numel() < self.nnz or out.dtype != val_blocked.dtype:
            raise ValueError("out is too short or of another dtype")
        check(lib().mbx_colblock_values_to_csr(self.handle, _KIND_CODES[kind],
                                               _dptr(val_blocked, val_blocked.dtype, "val_blocked"),
                                               _dptr(out, out.dtype, "out"),
                                               _stream_handle(stream)),
              "mbx_colblock_values_to_csr")
        return out[:self.nnz] if out.numel() != self.nnz else out

    def spmv(self, val_blocked: torch.Tensor, x: torch.Tensor, y: torch.Tensor | None = None,
             alpha: float = 1.0, beta: float = 0.0, stream=None) -> torch.Tensor:
        """y = alpha*A*x + beta*y on values in block-major order."""
        kind = _VALUE_KINDS.get(val_blocked.dtype)
        if kind is None:
            raise TypeError(f"unsupported value dtype {val_blocked.dtype}")
        if val_blocked.numel() < self.value_slots:
            raise ValueError("val_blocked is shorter than the blocked pattern")
        xdt = torch.float64 if kind == "f64" else torch.float32
        if x.numel() != self.ncol:
            raise ValueError("x has the wrong length")
        if y is None:
            y = torch.empty(self.nrow, dtype=xdt, device=x.device)
            beta = 0.0
        if y.numel() != self.nrow:
            raise ValueError("y has the wrong length")
        check(lib().mbx_colblock_spmv(self.handle, _KIND_CODES[kind],
                                      _dptr(val_blocked, val_blocked.dtype, "val_blocked"),
                                      alpha, _dptr(x, xdt, "x"), beta, _dptr(y, xdt, "y"),
                                      _stream_handle(stream)),
              "mbx_colblock_spmv")
        return y

    def set_stream(self, mode: str) -> None:
        """Stream cache policy of every block's product (see MerbitPlan)."""
        check(lib().mbx_colblock_set_stream(self.handle, STREAM_MODES[mode]),
              "mbx_colblock_set_stream")

    def set_residency(self, wg_per_cu: int) -> None:
        """Tiles per CU of every block's product (0: no cap)."""
        check(lib().mbx_colblock_set_residency(self.handle, wg_per_cu),
              "mbx_colblock_set_residency")

    def sort_cols(self, enable: bool = True, stream=None) -> bool:
        """Tile-sorted gathers in every block product
        (mbx_colblock_sort_cols); False when the blocks' shape rules them
        out.  Results are bit-identical either way."""
        st = lib().mbx_colblock_sort_cols(self.handle, int(bool(enable)), _stream_handle(stream))
        if st == -3:  # MBX_EUNSUPPORTED
            return False
        check(st, "mbx_colblock_sort_cols")
        return bool(enable)

    def prepare_adjoint(self, row_ptr: torch.Tensor, col: torch.Tensor, nblocks: int = 0,
                        items_per_thread: int = 8, stream=None) -> int:
        """Build the column-blocked form of A^T (blocks over A's rows) from the
        CSR pattern this form was made from; returns its block count."""
        nb = ctypes.c_int32()
        check(lib().mbx_colblock_adjoint_prepare(self.handle,
                                                 _dptr(row_ptr, torch.int32, "row_ptr"),
                                                 _dptr(col, torch.int32, "col"), nblocks,
                                                 items_per_thread, _stream_handle(stream),
                                                 ctypes.byref(nb)),
              "mbx_colblock_adjoint_prepare")
        self.adjoint_blocks = nb.value
        return nb.value

    def adjoint_x(self, val_blocked: torch.Tensor, g: torch.Tensor,
                  dx: torch.Tensor | None = None, alpha: float = 1.0, beta: float = 0.0,
                  stream=None) -> torch.Tensor:
        """dx = alpha*A^T*g + beta*dx from values in block-major order."""
        kind = _VALUE_KINDS.get(val_blocked.dtype)
        if kind is None:
            raise TypeError(f"unsupported value dtype {val_blocked.dtype}")
        if val_blocked.numel() < self.value_slots:
            raise ValueError("val_blocked is shorter than the blocked pattern")
        gdt = torch.float64 if kind == "f64" else torch.float32
        if g.numel() != self.nrow:
            raise ValueError("g has the wrong length")
        if dx is None:
            dx = torch.empty(self.ncol, dtype=gdt, device=g.device)
            beta = 0.0
        if dx.numel() != self.ncol:
            raise ValueError("dx has the wrong length")
        check(lib().mbx_colblock_dx(self.handle, _KIND_CODES[kind],
                                    _dptr(val_blocked, val_blocked.dtype, "val_blocked"),
                                    alpha, _dptr(g, gdt, "g"), beta, _dptr(dx, gdt, "dx"),
                                    _stream_handle(stream)), "mbx_colblock_dx")
        return dx

    def sddmm(self, g: torch.Tensor, x: torch.Tensor, out: torch.Tensor | None = None,
              stream=None) -> torch.Tensor:
        """dval = g[row]*x[col] in block-major order (FP64 or FP32)."""
        if g.dtype not in (torch.float64, torch.float32) or x.dtype != g.dtype:
            raise TypeError("sddmm takes FP64 or FP32 g and x of one dtype")
        if g.numel() != self.nrow or x.numel() != self.ncol:
            raise ValueError("g or x has the wrong length")
        if out is None:
            out = torch.zeros(max(self.value_slots, 1), dtype=g.dtype, device=g.device)
        if out.numel() < self.value_slots or out.dtype != g.dtype:
            raise ValueError("out is too short or of another dtype")
        kind = 0 if g.dtype == torch.float64 else 1
        check(lib().mbx_colblock_sddmm(self.handle, kind, _dptr(g, g.dtype, "g"),
                                       _dptr(x, x.dtype, "x"), _dptr(out, out.dtype, "out"),
                                       _stream_handle(stream)), "mbx_colblock_sddmm")
        return out

    def close(self) -> None:
        if getattr(self, "_h", None) is not None:
            lib().mbx_colblock_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
