"""Reference-semantics graph ingest (numpy front end of include/merbit_host.h).

``load_graph(config, gid)`` returns the matrix the reference's Fun_Merbit
operates on: the CSR of A^T of the loaded MTX graph with values
1/out-degree(row) (src/class.cu:22-34,60-177; include/tile.cuh:571-574).
"""
from __future__ import annotations

import ctypes

import numpy as np

from ._lib import MerbitError, check, lib
from .gen import Csr


def config_lookup(config_file: str, gid: int) -> str:
    buf = ctypes.create_string_buffer(4096)
    check(lib().mbh_config_lookup(config_file.encode(), int(gid), buf, len(buf)),
          f"config ID {gid} in {config_file}")
    return buf.value.decode()


def read_mtx(path: str):
    """(nrow, row_ind, col_ind, val) with the reference's mirroring rules."""
    n = ctypes.c_int32()
    m = ctypes.c_int64()
    p = path.encode()
    check(lib().mbh_mtx_read(p, ctypes.byref(n), ctypes.byref(m), None, None, None),
          f"read {path}")
    ri = np.empty(m.value, np.int32)
    ci = np.empty(m.value, np.int32)
    v = np.empty(m.value, np.float64)
    check(lib().mbh_mtx_read(p, ctypes.byref(n), ctypes.byref(m), ri.ctypes.data,
                             ci.ctypes.data, v.ctypes.data), f"read {path}")
    return n.value, ri, ci, v


def ref_ingest(nrow: int, row_ind, col_ind) -> Csr:
    ri = np.ascontiguousarray(row_ind, np.int32)
    ci = np.ascontiguousarray(col_ind, np.int32)
    m = len(ri)
    rp = np.empty(nrow + 1, np.int32)
    col = np.empty(max(m, 1), np.int32)
    val = np.empty(max(m, 1), np.float64)
    check(lib().mbh_ref_ingest(nrow, m, ri.ctypes.data, ci.ctypes.data, rp.ctypes.data,
                               col.ctypes.data, val.ctypes.data), "ingest")
    return Csr(rp, col[:m], val[:m], nrow)


def device_ingest(nrow: int, row_ind, col_ind, dtype=None):
    """The same matrix built on the GPU (mbx_ingest_ref): torch tensors
    (row_ptr, col, val) of the CSR of A^T with values 1/out-degree(row).
    row_ind / col_ind: int32 device tensors; dtype float64 (default) or
    float32."""
    import torch
    from ._lib import MBX_VALUE_F32, MBX_VALUE_F64
    from .spmv import _stream_handle
    dtype = dtype or torch.float64
    kind = {torch.float64: MBX_VALUE_F64, torch.float32: MBX_VALUE_F32}[dtype]
    if row_ind.dtype != torch.int32 or col_ind.dtype != torch.int32 or not row_ind.is_cuda \
            or row_ind.shape != col_ind.shape:
        raise ValueError("row_ind and col_ind must be int32 device tensors of one length")
    ri, ci = row_ind.contiguous(), col_ind.contiguous()
    m = ri.numel()
    dev = ri.device
    rp = torch.empty(nrow + 1, dtype=torch.int32, device=dev)
    col = torch.empty(max(m, 1), dtype=torch.int32, device=dev)
    val = torch.empty(max(m, 1), dtype=dtype, device=dev)
    check(lib().mbx_ingest_ref(nrow, m, ri.data_ptr(), ci.data_ptr(), kind, rp.data_ptr(),
                               col.data_ptr(), val.data_ptr(), _stream_handle(None)),
          "device ingest")
    return rp, col[:m], val[:m]


def load_graph(config_file: str, gid: int) -> Csr:
    n, ri, ci, _ = read_mtx(config_lookup(config_file, gid))
    return ref_ingest(n, ri, ci)


__all__ = ["config_lookup", "read_mtx", "ref_ingest", "device_ingest", "load_graph",
           "MerbitError"]
