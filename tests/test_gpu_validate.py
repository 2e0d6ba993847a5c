"""mbx_csr_validate on well-formed and malformed CSR patterns, and the
entry points that refuse malformed input (plan creation, the adjoint, the
column-blocked adjoint).  Every malformed case here stays within the arrays
it names: the checks only read row_ptr[0..nrow] and col[0..nnz)."""
from __future__ import annotations

import ctypes

import numpy as np
import pytest
import torch

from helpers import random_csr

pytestmark = pytest.mark.gpu


def dev(a, device):
    return torch.from_numpy(np.ascontiguousarray(a)).to(device)


def test_valid_patterns(device):
    from merbit_amd import workloads
    from merbit_amd.spmv import validate_csr
    A, _ = workloads.laplacian(64)
    rep = validate_csr(dev(A.row_ptr, device), dev(A.col, device), A.ncol)
    assert rep["ok"] and rep["row_ptr_errors"] == 0 and rep["col_errors"] == 0
    assert rep["unsorted_pairs"] == 0 and rep["nnz_mismatch"] == 0 and rep["first_bad_col"] == -1
    rng = np.random.default_rng(3)
    rp, col = random_csr(rng, 5000, 700, density=9.0)
    rep = validate_csr(dev(rp, device), dev(col, device), 700)
    assert rep["ok"] and rep["unsorted_pairs"] == 0
    # empty rows only
    rep = validate_csr(dev(np.zeros(101, np.int32), device),
                       torch.zeros(0, dtype=torch.int32, device=device), 5)
    assert rep["ok"]


def test_malformed_patterns(device):
    from merbit_amd.spmv import validate_csr
    rng = np.random.default_rng(4)
    rp, col = random_csr(rng, 3000, 500, density=6.0)
    nnz = int(rp[-1])
    # columns out of range: one past the end and one negative
    bad = col.copy()
    bad[17] = 500
    bad[nnz - 2] = -3
    rep = validate_csr(dev(rp, device), dev(bad, device), 500)
    assert not rep["ok"] and rep["col_errors"] == 2 and rep["first_bad_col"] == 17
    # a decreasing row pointer (within [0, nnz]) and a non-zero start
    r2 = rp.copy()
    r = 1000 + int(np.argmax(np.diff(rp)[1000:] > 0))  # a non-empty row
    r2[r], r2[r + 1] = r2[r + 1], r2[r]
    rep = validate_csr(dev(r2, device), dev(col, device), 500)
    assert not rep["ok"] and rep["row_ptr_errors"] >= 1
    r3 = rp.copy()
    r3[0] = 1
    rep = validate_csr(dev(r3, device), dev(col, device), 500)
    assert not rep["ok"] and rep["row_ptr_errors"] == 1
    # row_ptr[nrow] != nnz (the column array holds one entry more)
    col_long = np.concatenate([col, np.zeros(1, np.int32)])
    rep = validate_csr(dev(rp, device), dev(col_long, device), 500)
    assert not rep["ok"] and rep["nnz_mismatch"] == 1
    # descending columns inside rows: reported, not an error
    unsorted = col.copy()
    r = int(np.argmax(np.diff(rp)))
    seg = slice(int(rp[r]), int(rp[r + 1]))
    unsorted[seg] = np.sort(unsorted[seg])[::-1]
    expect = int(np.sum(np.diff(unsorted[seg].astype(np.int64)) < 0))
    rep = validate_csr(dev(rp, device), dev(unsorted, device), 500)
    assert rep["ok"] and rep["unsorted_pairs"] == expect and expect > 0


def test_plan_refuses_row_ptr_not_ending_at_nnz(device):
    from merbit_amd import MerbitError
    from merbit_amd.spmv import MerbitPlan
    rp = dev(np.array([0, 2, 5, 5], np.int32), device)
    MerbitPlan(rp, 4, 5).close()
    with pytest.raises(MerbitError):
        MerbitPlan(rp, 4, 6)
    with pytest.raises(MerbitError):
        MerbitPlan(dev(np.array([1, 2, 5, 5], np.int32), device), 4, 5)


def test_adjoint_refuses_out_of_range_columns(device):
    from merbit_amd import _lib
    L = _lib.lib()
    rng = np.random.default_rng(5)
    rp, col = random_csr(rng, 2000, 300, density=5.0)
    bad = col.copy()
    bad[11] = 300  # one past the last column
    d_rp, d_bad = dev(rp, device), dev(bad, device)
    h = ctypes.c_void_p()
    s = torch.cuda.current_stream().cuda_stream
    st = L.mbx_adjoint_create(d_rp.data_ptr(), d_bad.data_ptr(), 2000, 300, int(rp[-1]), 8, s,
                              ctypes.byref(h))
    assert st == -1 and not h.value
    d_col = dev(col, device)
    assert L.mbx_adjoint_create(d_rp.data_ptr(), d_col.data_ptr(), 2000, 300, int(rp[-1]), 8, s,
                                ctypes.byref(h)) == 0
    L.mbx_adjoint_destroy(h)


def test_colblock_adjoint_refuses_another_pattern(device):
    from merbit_amd import MerbitError
    from merbit_amd.spmv import ColumnBlockedCsr
    rng = np.random.default_rng(6)
    rp, col = random_csr(rng, 4000, 900, density=7.0)
    d_rp, d_col = dev(rp, device), dev(col, device)
    cb = ColumnBlockedCsr(d_rp, d_col, 900, nblocks=3)
    # same nnz and row lengths, one column moved
    other = col.copy()
    other[5] = (other[5] + 1) % 900
    with pytest.raises(MerbitError):
        cb.prepare_adjoint(d_rp, dev(other, device), nblocks=2)
    # same nnz, different row lengths
    rp2 = rp.copy()
    k = int(np.argmax(np.diff(rp) > 1))
    rp2[k + 1] -= 1
    with pytest.raises(MerbitError):
        cb.prepare_adjoint(dev(rp2, device), d_col, nblocks=2)
    assert cb.prepare_adjoint(d_rp, d_col, nblocks=2) == 2
    cb.close()


def test_sddmm_out_of_range_column_reads_zero(device):
    """The SDDMM gathers x through a bounded buffer resource, like the
    forward product: an out-of-range column gives 0 for that entry."""
    from merbit_amd.spmv import MerbitPlan
    from merbit_amd import _lib
    L = _lib.lib()
    rng = np.random.default_rng(7)
    rp, col = random_csr(rng, 1000, 200, density=4.0)
    bad = col.copy()
    bad[3] = 200 + 1000000
    nnz = int(rp[-1])
    d_rp = dev(rp, device)
    plan = MerbitPlan(d_rp, 200, nnz)
    g = torch.randn(1000, dtype=torch.float64, device=device)
    x = torch.randn(200, dtype=torch.float64, device=device)
    out = torch.empty(nnz, dtype=torch.float64, device=device)
    s = torch.cuda.current_stream().cuda_stream
    d_bad = dev(bad, device)
    assert L.mbx_sddmm(plan.handle, 0, d_rp.data_ptr(), d_bad.data_ptr(), g.data_ptr(),
                       x.data_ptr(), out.data_ptr(), s) == 0
    torch.cuda.synchronize()
    rows = np.repeat(np.arange(1000), np.diff(rp))
    ref = g.cpu().numpy()[rows] * x.cpu().numpy()[np.minimum(bad, 199)]
    ref[3] = 0.0
    assert np.array_equal(out.cpu().numpy(), ref)
    plan.close()
