"""Differentiable merge-path SpMV: y = A(val) x with a hand-written adjoint.

``MerbitOperator`` holds one sparsity pattern in HBM with its forward plan and
(lazily) its transposed plan.  ``spmv(op, val, x)`` is a torch.autograd
function: the backward pass computes

* dx   = A^T g        (tile SpMV over the transposed pattern), and
* dval = g[row]*x[col] (SDDMM over the pattern, forward plan's tiles),

only for the inputs that require gradients.  FP64 and FP32 values are
supported (FP32 sums are TwoSum-compensated).  All device work runs in
libmerbit_amd.so.
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from ._lib import MBX_VALUE_F32, MBX_VALUE_F64, check, lib
from .spmv import MerbitPlan, _dptr, _stream_handle

_KIND = {torch.float64: MBX_VALUE_F64, torch.float32: MBX_VALUE_F32}


class MerbitOperator:
    def __init__(self, row_ptr: torch.Tensor, col: torch.Tensor, ncol: int,
                 items_per_thread: int | None = None, adjoint_blocks: int | None = None,
                 blocks: int | None = None, value_dtype=torch.float32):
        """adjoint_blocks: 0 for dx by one merge-path product on A^T; -1 or
        1..16 for dx through a column-blocked A^T (mbx_adjoint_set_blocks; -1
        picks the count), for graphs whose g outgrows one XCD's L2.
        blocks: -1 or 1..16 runs the forward product on a column-blocked
        form too (values stay in CSR order and are moved into block order per
        forward call), and sets adjoint_blocks to -1 unless it is given.  dval
        stays the CSR SDDMM: moving a blocked dval back to CSR order costs
        more than the blocked SDDMM saves.
        None (the defaults): the launch rule (mbx_launch_rule, for values of
        value_dtype) picks the forward plan, and both blocked forms when it
        finds spread columns over an x larger than L2 (random graphs).
        Otherwise, with items_per_thread left to the rule as well, the rule
        run on the transposed pattern (mbx_adjoint_auto) picks the plan of
        A^T: its hot-column split when A has a few long rows, the blocked
        A^T for spread rows; ``adjoint_choice`` holds that choice once the
        adjoint exists.  A given items_per_thread is kept for A^T too."""
        spread = False
        if blocks is None or (adjoint_blocks is None and items_per_thread is not None):
            from .spmv import launch_rule
            spread = launch_rule(row_ptr, col, ncol, value_dtype)["rule"] == "spread_columns"
        if blocks is None:
            blocks = -1 if spread else 0
        if adjoint_blocks is None and (blocks or items_per_thread is not None):
            adjoint_blocks = -1 if (blocks or spread) else 0
        if blocks and not adjoint_blocks:
            adjoint_blocks = -1
        self.adjoint_blocks = adjoint_blocks  # None: the rule on A^T (mbx_adjoint_auto)
        self.adjoint_choice = None
        from .spmv import _KIND_OF_DTYPE
        self._kind = _KIND_OF_DTYPE[value_dtype]
        self.row_ptr = row_ptr
        self.col = col
        self.nrow = row_ptr.numel() - 1
        self.ncol = int(ncol)
        self.nnz = col.numel()
        self.plan = MerbitPlan(row_ptr, ncol, self.nnz, items_per_thread, col=col,
                               value_dtype=value_dtype)
        self.items_per_thread = self.plan.items_per_thread
        # the SDDMM keeps a tile's row ends and rows in LDS, so it runs on a
        # partition of 8 items per thread where the plan's tiles are larger
        # (R-MAT FP64 dval 101 us on the rule's 16, 90 us on 8)
        self._sddmm_plan = MerbitPlan(row_ptr, ncol, self.nnz, 8) \
            if self.items_per_thread > 12 else None
        self._adj = None
        self._cb = None
        if blocks:
            from .spmv import ColumnBlockedCsr
            self._cb = ColumnBlockedCsr(row_ptr, col, ncol, 0 if blocks < 0 else blocks,
                                        8 if self.items_per_thread not in (8, 12) else
                                        self.items_per_thread)
        self._vb = None

    @property
    def adjoint_handle(self) -> ctypes.c_void_p:
        if self._adj is None:
            h = ctypes.c_void_p()
            check(lib().mbx_adjoint_create(_dptr(self.row_ptr, torch.int32, "row_ptr"),
                                           _dptr(self.col, torch.int32, "col"), self.nrow,
                                           self.ncol, self.nnz, self.items_per_thread,
                                           _stream_handle(None), ctypes.byref(h)),
                  "mbx_adjoint_create")
            self._adj = h
            if self.adjoint_blocks is None:
                from ._lib import LaunchChoice
                from .spmv import _choice_dict
                c = LaunchChoice()
                check(lib().mbx_adjoint_auto(h, self._kind, _stream_handle(None),
                                             ctypes.byref(c)), "mbx_adjoint_auto")
                self.adjoint_choice = _choice_dict(c)
            elif self.adjoint_blocks:
                check(lib().mbx_adjoint_set_blocks(h, self.adjoint_blocks, _stream_handle(None)),
                      "mbx_adjoint_set_blocks")
        return self._adj

    def transpose(self):
        """Host copies (col_ptr, row_idx, perm) of the transposed pattern."""
        cp = np.empty(self.ncol + 1, np.int32)
        ri = np.empty(max(self.nnz, 1), np.int32)
        pm = np.empty(max(self.nnz, 1), np.int32)
        check(lib().mbx_adjoint_transpose(self.adjoint_handle, cp.ctypes.data, ri.ctypes.data,
                                          pm.ctypes.data), "mbx_adjoint_transpose")
        return cp, ri[:self.nnz], pm[:self.nnz]

    def forward(self, val: torch.Tensor, x: torch.Tensor) -> torch.Tensor:
        if self._cb is not None:
            if self._vb is None or self._vb.dtype != val.dtype:
                self._vb = torch.empty(max(self._cb.value_slots, 1), dtype=val.dtype,
                                       device=val.device)
            return self._cb.spmv(self._cb.values(val, out=self._vb), x)
        return self.plan.spmv(self.row_ptr, self.col, val, x)

    def adjoint_x(self, val: torch.Tensor, g: torch.Tensor) -> torch.Tensor:
        kind = _KIND[val.dtype]
        dx = torch.empty(self.ncol, dtype=val.dtype, device=val.device)
        check(lib().mbx_adjoint_dx(self.adjoint_handle, kind, _dptr(val, val.dtype, "val"), 1.0,
                                   _dptr(g, val.dtype, "g"), 0.0, _dptr(dx, val.dtype, "dx"),
                                   _stream_handle(None)), "mbx_adjoint_dx")
        return dx

    def adjoint_val(self, g: torch.Tensor, x: torch.Tensor) -> torch.Tensor:
        kind = _KIND[x.dtype]
        dval = torch.empty(self.nnz, dtype=x.dtype, device=x.device)
        plan = self._sddmm_plan or self.plan
        check(lib().mbx_sddmm(plan.handle, kind, _dptr(self.row_ptr, torch.int32, "row_ptr"),
                              _dptr(self.col, torch.int32, "col"), _dptr(g, x.dtype, "g"),
                              _dptr(x, x.dtype, "x"), _dptr(dval, x.dtype, "dval"),
                              _stream_handle(None)), "mbx_sddmm")
        return dval

    def close(self) -> None:
        if self._adj is not None:
            lib().mbx_adjoint_destroy(self._adj)
            self._adj = None
        if self._cb is not None:
            self._cb.close()
            self._cb = None
        if self._sddmm_plan is not None:
            self._sddmm_plan.close()
            self._sddmm_plan = None
        self.plan.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class _SpMV(torch.autograd.Function):
    @staticmethod
    def forward(ctx, val, x, op):
        if val.dtype not in _KIND or x.dtype != val.dtype:
            raise TypeError("values and x must both be FP64 or both FP32")
        ctx.op = op
        ctx.save_for_backward(val, x)
        return op.forward(val.contiguous(), x.contiguous())

    @staticmethod
    def backward(ctx, gy):
        val, x = ctx.saved_tensors
        op = ctx.op
        gy = gy.contiguous().to(val.dtype)
        dval = op.adjoint_val(gy, x) if ctx.needs_input_grad[0] else None
        dx = op.adjoint_x(val, gy) if ctx.needs_input_grad[1] else None
        return dval, dx, None


class ColumnBlockedOperator:
    """The operator of ``MerbitOperator`` on the column-blocked form of the
    pattern (``ColumnBlockedCsr``), for random graphs whose x and g outgrow
    L2.  Its values live in the form's block-major order: convert a CSR-order
    parameter once with ``to_blocked`` and train that tensor; dval comes back
    in the same order (zero in the padding slots).  Use with ``spmv``."""

    def __init__(self, row_ptr: torch.Tensor, col: torch.Tensor, ncol: int,
                 nblocks: int = 0, adjoint_blocks: int = 0, items_per_thread: int = 8):
        from .spmv import ColumnBlockedCsr
        self.nrow = row_ptr.numel() - 1
        self.ncol = int(ncol)
        self.form = ColumnBlockedCsr(row_ptr, col, ncol, nblocks, items_per_thread)
        self.form.prepare_adjoint(row_ptr, col, adjoint_blocks, items_per_thread)
        self.value_slots = self.form.value_slots

    def to_blocked(self, val: torch.Tensor) -> torch.Tensor:
        return self.form.values(val)

    def forward(self, val: torch.Tensor, x: torch.Tensor) -> torch.Tensor:
        return self.form.spmv(val, x)

    def adjoint_x(self, val: torch.Tensor, g: torch.Tensor) -> torch.Tensor:
        return self.form.adjoint_x(val, g)

    def adjoint_val(self, g: torch.Tensor, x: torch.Tensor) -> torch.Tensor:
        return self.form.sddmm(g, x)

    def close(self) -> None:
        self.form.close()


def spmv(op, val: torch.Tensor, x: torch.Tensor) -> torch.Tensor:
    """Differentiable y = A x with A's pattern from `op` (a MerbitOperator,
    values in CSR order, or a ColumnBlockedOperator, values in its order)."""
    return _SpMV.apply(val, x, op)
