"""GPU parity of the batched SpMM (mbx_spmm_f32) against the oracle, column
by column.  Bar: FP32 accumulation with FP64 carries; condition-scaled error
max |y - y_ref| / (|A||x|) <= 1e-5 against the Dot2-accurate FP64 oracle on the
same FP32 values and right-hand sides."""
from __future__ import annotations

import numpy as np
import pytest
import torch

from helpers import cond_err, edge_case_patterns, random_csr

pytestmark = pytest.mark.gpu

SPMM_TOL = 1e-5


def say(msg):
    print(f"[case] {msg}", flush=True)


def dev(a, device):
    return torch.from_numpy(np.ascontiguousarray(a)).to(device)


def check_spmm(orc, rp, col, val, X, Y):
    for j in range(X.shape[1]):
        ya, absr = orc.spmv_accurate(rp, col, val, np.ascontiguousarray(X[:, j]))
        e = cond_err(Y[:, j].astype(np.float64), ya, absr)
        assert e <= SPMM_TOL, (j, e)


@pytest.mark.parametrize("ipt", [2, 4])
@pytest.mark.parametrize("k", [1, 3, 8, 13, 32])
def test_spmm_edge_cases(orc, device, ipt, k):
    from merbit_amd.spmv import MerbitPlan
    rng = np.random.default_rng(100 * ipt + k)
    for name, lens, ncol in edge_case_patterns(256 * ipt):
        say(f"spmm {name} ipt={ipt} k={k}")
        rp, col = random_csr(rng, len(lens), ncol, lens=lens)
        val = rng.uniform(-1, 1, int(rp[-1])).astype(np.float32)
        X = rng.uniform(-1, 1, (ncol, k)).astype(np.float32)
        d_rp, d_col, d_val = dev(rp, device), dev(col, device), dev(val, device)
        plan = MerbitPlan(d_rp, ncol, int(rp[-1]), items_per_thread=ipt)
        Xc = dev(X.T.copy(), device).t()           # column-major
        Y = plan.spmm(d_rp, d_col, d_val, Xc)      # column-major result
        check_spmm(orc, rp, col, val, X, Y.cpu().numpy())
        Xr = dev(X, device)                         # row-major
        Yr = torch.empty((len(lens), k), dtype=torch.float32, device=device)
        plan.spmm(d_rp, d_col, d_val, Xr, Yr)
        assert np.array_equal(Yr.cpu().numpy(), Y.cpu().numpy()), name
        plan.close()


@pytest.mark.parametrize("ipt", [2, 4])
@pytest.mark.parametrize("k", [6, 21])
def test_spmm_alpha_beta_and_ld(orc, device, ipt, k):
    from merbit_amd.spmv import MerbitPlan
    rng = np.random.default_rng(4 + k)
    rp, col = random_csr(rng, 3000, 2500, lens=(rng.pareto(1.2, 3000) * 5).astype(int))
    val = rng.uniform(-1, 1, int(rp[-1])).astype(np.float32)
    X = rng.uniform(-1, 1, (2500, k)).astype(np.float32)
    Y0 = rng.uniform(-1, 1, (3000, k)).astype(np.float32)
    d_rp, d_col, d_val = dev(rp, device), dev(col, device), dev(val, device)
    plan = MerbitPlan(d_rp, 2500, int(rp[-1]), items_per_thread=ipt)
    # leading dimensions larger than the matrix: views into bigger buffers
    Xbig = torch.zeros((k + 3, 2600), dtype=torch.float32, device=device)
    Xbig[:k, :2500] = dev(X.T.copy(), device)
    Xv = Xbig[:k, :2500].t()
    Ybig = torch.zeros((3000, k + 5), dtype=torch.float32, device=device)
    Ybig[:, :k] = dev(Y0, device)
    Yv = Ybig[:, :k]
    say(f"spmm alpha beta ld ipt={ipt} k={k}")
    plan.spmm(d_rp, d_col, d_val, Xv, Yv, alpha=-2.0, beta=0.5)
    Y = Yv.cpu().numpy().astype(np.float64)
    for j in range(k):
        ya, absr = orc.spmv_accurate(rp, col, val, np.ascontiguousarray(X[:, j]))
        want = -2.0 * ya + 0.5 * Y0[:, j]
        den = 2.0 * absr + 0.5 * np.abs(Y0[:, j]) + 1e-30
        assert np.max(np.abs(Y[:, j] - want) / den) <= SPMM_TOL
    assert torch.all(Ybig[:, k:] == 0)
    plan.close()


@pytest.fixture(scope="module")
def rmat_spmm_case(orc):
    """The batched workload's R-MAT at full size (2^20 rows, 2^24 entries,
    FP32 values) with 32 right-hand sides and the oracle's accurate product
    for each of them (computed once for every k and tile shape)."""
    from merbit_amd import gen, workloads
    A, _ = workloads.rmat()
    val = A.val.astype(np.float32)
    X = gen.uniform(A.ncol * 32, -1, 1, 77, np.float32).reshape(32, A.ncol).T
    refs = [orc.spmv_accurate(A.row_ptr, A.col, val, np.ascontiguousarray(X[:, j]))
            for j in range(32)]
    return A, val, X, refs


@pytest.mark.slow
@pytest.mark.parametrize("ipt", [2, 4])
@pytest.mark.parametrize("k", [8, 32])
def test_spmm_rmat(rmat_spmm_case, device, k, ipt):
    """R-MAT (2^20 rows) times k = 8 and 32 column-major right-hand sides,
    every column against the oracle."""
    from merbit_amd.spmv import MerbitPlan
    A, val, X32, refs = rmat_spmm_case
    say(f"spmm rmat scale 20 k={k} ipt={ipt}")
    X = X32[:, :k]
    d_rp, d_col, d_val = dev(A.row_ptr, device), dev(A.col, device), dev(val, device)
    plan = MerbitPlan(d_rp, A.ncol, A.nnz, items_per_thread=ipt)
    Y = plan.spmm(d_rp, d_col, d_val, dev(X.T.copy(), device).t()).cpu().numpy()
    plan.close()
    for j in range(k):
        ya, absr = refs[j]
        e = cond_err(Y[:, j].astype(np.float64), ya, absr)
        assert e <= SPMM_TOL, (j, e)


@pytest.mark.parametrize("ipt", [2, 4])
@pytest.mark.parametrize("k", [5, 64, 70])
def test_spmm_padded_columns(orc, device, k, ipt):
    """Column-major X whose leading dimension is a multiple of 4 but whose
    column count is not (the 16-byte repack path with a ragged tail), and
    more right-hand sides than one repack block holds."""
    from merbit_amd.spmv import MerbitPlan
    rng = np.random.default_rng(70 + k)
    ncol, ldx = 2501, 2504
    rp, col = random_csr(rng, 1700, ncol, lens=(rng.pareto(1.1, 1700) * 4).astype(int))
    val = rng.uniform(-1, 1, int(rp[-1])).astype(np.float32)
    X = rng.uniform(-1, 1, (ncol, k)).astype(np.float32)
    say(f"spmm padded columns k={k} ipt={ipt}")
    Xbig = torch.full((k, ldx), float("nan"), dtype=torch.float32, device=device)
    Xbig[:, :ncol] = dev(X.T.copy(), device)
    plan = MerbitPlan(dev(rp, device), ncol, int(rp[-1]), items_per_thread=ipt)
    Y = plan.spmm(dev(rp, device), dev(col, device), dev(val, device), Xbig[:, :ncol].t())
    check_spmm(orc, rp, col, val, X, Y.cpu().numpy())
    plan.close()
