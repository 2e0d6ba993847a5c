"""The no-allocation FP64 product (mbx_spmv_f64_noalloc) through the C ABI,
against the planned product it is a variant of (mbx_plan_create with 8 items
per thread + mbx_spmv_f64): bit-identical results on consecutive calls, the
second on a smaller input (the scratch keeps the first call's contents), and
within 1e-12 of the oracle."""
from __future__ import annotations

import ctypes

import numpy as np
import pytest
import torch

from helpers import cond_err, random_csr

pytestmark = pytest.mark.gpu


def dev(a, device):
    return torch.from_numpy(np.ascontiguousarray(a)).to(device)


def planned(L, rp, col, val, x, y, ncol, alpha, beta):
    nrow, nnz = rp.numel() - 1, col.numel()
    s = torch.cuda.current_stream().cuda_stream
    h = ctypes.c_void_p()
    assert L.mbx_plan_create(rp.data_ptr(), nrow, ncol, nnz, 8, 256, s, ctypes.byref(h)) == 0
    assert L.mbx_spmv_f64(h, rp.data_ptr(), col.data_ptr(), val.data_ptr(), alpha,
                          x.data_ptr(), beta, y.data_ptr(), s) == 0
    torch.cuda.synchronize()
    L.mbx_plan_destroy(h)
    return y.cpu().numpy()


def noalloc(L, rp, col, val, x, y, ncol, alpha, beta):
    nrow, nnz = rp.numel() - 1, col.numel()
    s = torch.cuda.current_stream().cuda_stream
    assert L.mbx_spmv_f64_noalloc(rp.data_ptr(), col.data_ptr(), val.data_ptr(), nrow, ncol,
                                  nnz, alpha, x.data_ptr(), beta, y.data_ptr(), s) == 0
    torch.cuda.synchronize()
    return y.cpu().numpy()


def case(rng, lens, ncol):
    rp, col = random_csr(rng, len(lens), ncol, lens=lens)
    val = rng.standard_normal(int(rp[-1]))
    x = rng.uniform(-1, 1, ncol)
    y0 = rng.uniform(-1, 1, len(lens))
    return rp, col, val, x, y0


def test_noalloc_matches_planned_product_on_consecutive_calls(orc, device):
    from merbit_amd import _lib
    L = _lib.lib()
    rng = np.random.default_rng(2024)
    T = 2048
    big = np.concatenate([(rng.pareto(1.1, 150000) * 6).astype(int),
                          [70 * T + 11, 0, 3, T - 1, T, T + 1]])  # a row over 64 tiles
    small = np.concatenate([[0] * 100, (rng.pareto(1.3, 4000) * 3).astype(int), [3 * T + 5]])
    for name, lens, ncol, alpha, beta in (("big", big, 90000, 1.0, 0.0),
                                          ("small", small, 3000, -0.5, 0.25),
                                          ("small beta 0", small, 3000, 2.0, 0.0)):
        print(f"[case] noalloc {name}: {len(lens)} rows", flush=True)
        rp, col, val, x, y0 = case(rng, lens, ncol)
        d_rp, d_col, d_val, d_x = dev(rp, device), dev(col, device), dev(val, device), \
            dev(x, device)
        ya = noalloc(L, d_rp, d_col, d_val, d_x, dev(y0, device), ncol, alpha, beta)
        yb = planned(L, d_rp, d_col, d_val, d_x, dev(y0, device), ncol, alpha, beta)
        assert np.array_equal(ya.view(np.uint64), yb.view(np.uint64)), name
        yr, absr = orc.spmv_accurate(rp, col, val, x)
        want = alpha * yr + beta * y0
        den = abs(alpha) * absr + abs(beta) * np.abs(y0)
        assert cond_err(ya, want, den) <= 1e-12, name


def test_noalloc_empty_and_limits(device):
    from merbit_amd import _lib
    L = _lib.lib()
    s = torch.cuda.current_stream().cuda_stream
    # nothing to do: no kernel, MBX_OK
    assert L.mbx_spmv_f64_noalloc(None, None, None, 0, 0, 0, 1.0, None, 0.0, None, s) == 0
    # the scratch limit, refused before any launch
    assert L.mbx_spmv_f64_noalloc(None, None, None, 100, 100, 268433408, 1.0, None, 0.0,
                                  None, s) == -6
    # rows only (every row empty): y = beta*y
    rp = torch.zeros(5001, dtype=torch.int32, device=device)
    col = torch.zeros(0, dtype=torch.int32, device=device)
    y = torch.ones(5000, dtype=torch.float64, device=device)
    assert L.mbx_spmv_f64_noalloc(rp.data_ptr(), None, None, 5000, 7, 0, 1.0, None, 0.5,
                                  y.data_ptr(), s) == 0
    torch.cuda.synchronize()
    assert torch.all(y == 0.5)
