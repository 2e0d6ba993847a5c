"""Synthetic workloads for the merge-path SpMV (numpy front end of
include/merbit_gen.h).  All generators are bit-reproducible from their seed.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from ._lib import check, lib


@dataclass
class Csr:
    """0-based CSR: int32 row_ptr[nrow+1], int32 col[nnz], val[nnz]."""
    row_ptr: np.ndarray
    col: np.ndarray
    val: np.ndarray | None
    ncol: int

    @property
    def nrow(self) -> int:
        return len(self.row_ptr) - 1

    @property
    def nnz(self) -> int:
        return int(self.row_ptr[-1])


def _ptr(a: np.ndarray) -> int:
    return a.ctypes.data


def uniform(n: int, lo: float, hi: float, seed: int, dtype=np.float64) -> np.ndarray:
    out = np.empty(int(n), dtype=dtype)
    if out.dtype == np.float64:
        check(lib().mbg_fill_uniform_f64(_ptr(out), n, lo, hi, seed), "uniform")
    elif out.dtype == np.float32:
        check(lib().mbg_fill_uniform_f32(_ptr(out), n, lo, hi, seed), "uniform")
    else:
        raise TypeError(dtype)
    return out


def normal(n: int, seed: int, dtype=np.float64) -> np.ndarray:
    out = np.empty(int(n), dtype=dtype)
    if out.dtype == np.float64:
        check(lib().mbg_fill_normal_f64(_ptr(out), n, seed), "normal")
    elif out.dtype == np.float32:
        check(lib().mbg_fill_normal_f32(_ptr(out), n, seed), "normal")
    else:
        raise TypeError(dtype)
    return out


def to_bf16(a: np.ndarray) -> np.ndarray:
    """float32 -> bfloat16 bit patterns (uint16), round to nearest even."""
    src = np.ascontiguousarray(a, dtype=np.float32)
    out = np.empty(src.shape, dtype=np.uint16)
    check(lib().mbg_f32_to_bf16(_ptr(src), _ptr(out), src.size), "to_bf16")
    return out


def bf16_to_f32(h: np.ndarray) -> np.ndarray:
    return (np.ascontiguousarray(h, dtype=np.uint16).astype(np.uint32) << 16).view(np.float32)


def laplacian5(g: int) -> Csr:
    nnz = lib().mbg_laplacian5_nnz(g)
    rp = np.empty(g * g + 1, np.int32)
    col = np.empty(nnz, np.int32)
    val = np.empty(nnz, np.float64)
    check(lib().mbg_laplacian5(g, _ptr(rp), _ptr(col), _ptr(val)), "laplacian5")
    return Csr(rp, col, val, g * g)


def rmat(scale: int, edges: int, a: float = 0.57, b: float = 0.19, c: float = 0.19,
         seed: int = 0) -> Csr:
    n = 1 << scale
    rp = np.empty(n + 1, np.int32)
    col = np.empty(edges, np.int32)
    check(lib().mbg_rmat(scale, edges, a, b, c, seed, _ptr(rp), _ptr(col)), "rmat")
    return Csr(rp, col, None, n)


def banded(n: int, h: int) -> Csr:
    nnz = lib().mbg_banded_nnz(n, h)
    rp = np.empty(n + 1, np.int32)
    col = np.empty(nnz, np.int32)
    check(lib().mbg_banded(n, h, _ptr(rp), _ptr(col)), "banded")
    return Csr(rp, col, None, n)


def poisson_rows(n: int, ncol: int, lam: float, seed: int) -> Csr:
    rp = np.empty(n + 1, np.int32)
    check(lib().mbg_poisson_rows(n, ncol, lam, seed, _ptr(rp), None), "poisson_rows")
    col = np.empty(int(rp[-1]), np.int32)
    check(lib().mbg_poisson_rows(n, ncol, lam, seed, _ptr(rp), _ptr(col)), "poisson_rows")
    return Csr(rp, col, None, ncol)
