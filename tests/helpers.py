"""Shared checkers for the parity tests (condition-scaled error metrics)."""
from __future__ import annotations

import numpy as np


def cond_err(y, y_ref, absr) -> float:
    """max_i |y_i - y_ref_i| / (|A||x|)_i; rows with (|A||x|)_i == 0 must be
    exactly zero in both (else inf)."""
    y = np.asarray(y, np.float64)
    y_ref = np.asarray(y_ref, np.float64)
    absr = np.asarray(absr, np.float64)
    d = np.abs(y - y_ref)
    zero = absr == 0
    if np.any(d[zero] != 0) or np.any(y[zero] != 0):
        return float("inf")
    if not np.any(~zero):
        return 0.0
    return float(np.max(d[~zero] / absr[~zero]))


def norm_err(y, y_ref) -> float:
    y = np.asarray(y, np.float64)
    y_ref = np.asarray(y_ref, np.float64)
    den = np.linalg.norm(y_ref)
    return float(np.linalg.norm(y - y_ref) / (den if den > 0 else 1.0))


def random_csr(rng, nrow, ncol, density=None, lens=None):
    """Random CSR with sorted (possibly repeated) columns."""
    if lens is None:
        lens = rng.poisson(density, size=nrow)
    lens = np.asarray(lens, np.int64)
    rp = np.zeros(nrow + 1, np.int32)
    rp[1:] = np.cumsum(lens)
    col = np.concatenate([np.sort(rng.integers(0, ncol, size=int(l))) for l in lens]) \
        if nrow > 0 and rp[-1] > 0 else np.zeros(0, np.int64)
    return rp, col.astype(np.int32)


def edge_case_patterns(T: int):
    """(name, row lengths, ncol) cases around a tile of T merge items."""
    return [
        ("one_row", [7], 9),
        ("one_empty_row", [0], 1),
        ("all_empty", [0] * 5000, 100),
        ("no_nnz_rows_only", [0] * (3 * T + 5), 1),
        ("giant_row", [10 * T + 3], 1000),
        ("giant_row_mid", [3, 0, 5 * T, 0, 0, 2], 500),
        ("row_exact_tile", [T - 1] * 4, 300),
        ("row_tile_plus_end", [T] * 3, 300),
        ("alternating", [0, T + 1] * 3 + [0], 200),
        ("many_short", [1] * (2 * T), 50),
        ("tail_rows", [0] * (T - 1) + [5], 10),
        # rows crossing tile boundaries close to either side of them
        ("boundary_mix", [255, 256, 257, 511, 1, 0, 254] * 12, 700),
        # the fixup sums runs of up to 64 carries in one thread, longer runs
        # across the wave: rows just under, at and over that length
        ("run_64_edges", [1, 63 * T + 5, 0, 64 * T - 3, 2, 65 * T + 1, 3], 800),
        ("run_long", [2, 150 * T + 17, 1], 900),
    ]
