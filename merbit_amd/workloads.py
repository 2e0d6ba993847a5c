"""Named synthetic workloads of the merge-path SpMV benchmarks and tests.

Each builder returns host (numpy) arrays; seeds are fixed so every run and
every test sees the same matrix.  Sizes default to the full benchmark shapes
and can be scaled down for tests.
"""
from __future__ import annotations

import numpy as np

from . import gen
from .gen import Csr

SEED = {
    "laplacian_x": 101,
    "rmat_pattern": 202, "rmat_val": 203, "rmat_x": 204,
    "banded_val": 302, "banded_x": 303,
    "batch_shape": 401, "batch_pattern": 410, "batch_val": 430, "batch_x": 450,
    "graph_pattern": 501, "graph_val": 502, "graph_x": 503, "graph_g": 504,
    "spmm_x": 601,
}


def laplacian(grid: int = 256):
    """5-point Laplacian on grid x grid (4 on the diagonal, -1 off it), x ~ U[-1,1)."""
    A = gen.laplacian5(grid)
    x = gen.uniform(A.nrow, -1.0, 1.0, SEED["laplacian_x"])
    return A, x


def rmat(scale: int = 20, edges: int | None = None):
    """R-MAT power law (a=.57, b=c=.19), duplicates kept, values N(0,1), x ~ U[-1,1)."""
    if edges is None:
        edges = 16 << scale  # 2^24 entries at scale 20
    A = gen.rmat(scale, edges, 0.57, 0.19, 0.19, SEED["rmat_pattern"])
    A.val = gen.normal(A.nnz, SEED["rmat_val"])
    x = gen.uniform(A.nrow, -1.0, 1.0, SEED["rmat_x"])
    return A, x


def banded(n: int = 4_000_000, half: int = 13):
    """Banded n x n, 2*half+1 diagonals, values U[-1,1) as FP32, x ~ U[-1,1) FP32."""
    A = gen.banded(n, half)
    A.val = gen.uniform(A.nnz, -1.0, 1.0, SEED["banded_val"], np.float32)
    x = gen.uniform(n, -1.0, 1.0, SEED["banded_x"], np.float32)
    return A, x


def uneven_batch(count: int = 64, rows_lo: int = 1000, rows_hi: int = 200_000,
                 lam_lo: float = 5.0, lam_hi: float = 60.0):
    """`count` square CSR matrices, rows ~ U{rows_lo..rows_hi}, row length
    ~ Poisson(lambda_m), lambda_m ~ U[lam_lo, lam_hi], distinct uniform
    columns, FP32 values U[-1,1); one FP32 x per matrix."""
    u = gen.uniform(2 * count, 0.0, 1.0, SEED["batch_shape"])
    mats, xs = [], []
    for m in range(count):
        n = rows_lo + int(u[2 * m] * (rows_hi - rows_lo + 1))
        lam = lam_lo + (lam_hi - lam_lo) * u[2 * m + 1]
        A = gen.poisson_rows(n, n, lam, SEED["batch_pattern"] + m)
        A.val = gen.uniform(A.nnz, -1.0, 1.0, SEED["batch_val"] + m, np.float32)
        mats.append(A)
        xs.append(gen.uniform(n, -1.0, 1.0, SEED["batch_x"] + m, np.float32))
    return mats, xs


def graph(n: int = 2_400_000, degree: float = 25.0):
    """Random graph: out-degree ~ Poisson(degree), uniform distinct targets,
    FP32 values U[0,1), x and upstream g ~ N(0,1) FP32."""
    A = gen.poisson_rows(n, n, degree, SEED["graph_pattern"])
    A.val = gen.uniform(A.nnz, 0.0, 1.0, SEED["graph_val"], np.float32)
    x = gen.normal(n, SEED["graph_x"], np.float32)
    g = gen.normal(n, SEED["graph_g"], np.float32)
    return A, x, g


def algorithmic_bytes(nrow: int, ncol: int, nnz: int, s_val: int, s_x: int, s_y: int,
                      k: int = 1, beta_nonzero: bool = False) -> int:
    """Compulsory CSR traffic: row_ptr, col and values once, x read and y
    written once per right-hand side (y read too when beta != 0)."""
    b = 4 * (nrow + 1) + nnz * (4 + s_val) + k * (ncol * s_x + nrow * s_y)
    if beta_nonzero:
        b += k * nrow * s_y
    return b


__all__ = ["Csr", "laplacian", "rmat", "banded", "uneven_batch", "graph",
           "algorithmic_bytes", "SEED"]
