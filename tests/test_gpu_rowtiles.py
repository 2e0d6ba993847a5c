"""Row-aligned tiles (mbx_plan_create_rows): tile t starts at the start of
the row open at merge diagonal t*(T - L) (L the longest row), so no row
crosses a tile boundary and products skip the carry fixup.  The partition
is checked against the oracle's merge search snapped to the row start, the
products against the oracle (1e-12 condition-scaled for FP64, 1e-6 for
FP32/BF16 values) and against the plain plan; the hot split on a row plan
is bit-identical to the row plan's plain product."""
from __future__ import annotations

import ctypes

import numpy as np
import pytest
import torch

from helpers import cond_err, random_csr

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[case] {msg}", flush=True)


def dev(a, device):
    return torch.from_numpy(np.ascontiguousarray(a)).to(device)


def bits(t):
    a = t.cpu().numpy()
    return a.view(np.uint64 if a.dtype == np.float64 else np.uint32)


def short_row_cases(rng, T):
    """(name, row lengths, ncol) with every row at most T/16 long."""
    L = T // 16
    lens = rng.integers(0, L + 1, 30000)
    lens[rng.random(30000) < 0.2] = 0
    return [
        ("one_row", [min(7, L)], 9),
        ("one_empty_row", [0], 1),
        ("all_empty", [0] * 5000, 100),
        ("rows_at_limit", [L] * (3 * 16 + 5), 400),
        ("many_short", [1] * (2 * T + 3), 50),
        ("tail_rows", [0] * (T - 1) + [5], 10),
        ("random_short", lens, 20000),
        ("stencil_5", [5] * 40000, 40000),
    ]


def expected_partition(orc, rp, T):
    """The row-aligned coordinates from the oracle's merge search."""
    nrow, nnz = len(rp) - 1, int(rp[-1])
    L = int(np.max(np.diff(rp))) if nrow else 0
    S = T - L
    path = nrow + nnz
    nt = (path + S - 1) // S
    rows = np.empty(nt + 1, np.int64)
    for t in range(nt + 1):
        _, r = orc.merge_search(rp, nnz, min(t * S, path))
        rows[t] = r
    return rows, rp[rows].astype(np.int64)


@pytest.mark.parametrize("ipt", [2, 4, 8, 12])
def test_row_partition(orc, device, ipt):
    from merbit_amd.spmv import MerbitPlan
    rng = np.random.default_rng(400 + ipt)
    T = 256 * ipt
    for name, lens, ncol in short_row_cases(rng, T):
        rp, col = random_csr(rng, len(lens), ncol, lens=lens)
        nnz = int(rp[-1])
        say(f"row partition ipt {ipt} {name}")
        plan = MerbitPlan(dev(rp, device), ncol, nnz, items_per_thread=ipt, row_tiles=True)
        rows, ks = plan.partition()
        want_rows, want_ks = expected_partition(orc, rp, T)
        assert np.array_equal(rows, want_rows) and np.array_equal(ks, want_ks), name
        items = np.diff(rows.astype(np.int64) + ks)
        assert items.size == 0 or (items.min() >= 1 and items.max() <= T), name
        assert rows[-1] == len(lens) and ks[-1] == nnz
        assert np.array_equal(rp[rows], ks)  # every tile starts at a row start
        st = plan.stats(dev(rp, device))
        assert st["split_tiles"] == 0, name
        plan.close()


def test_row_plan_refusals(device):
    from merbit_amd._lib import lib
    T = 512
    h = ctypes.c_void_p()
    for L, want in ((T // 16, 0), (T // 16 + 1, -3)):
        rp = np.array([0, 3, 3 + L, 3 + L + 1], np.int32)
        d = dev(rp, device)
        st = lib().mbx_plan_create_rows(d.data_ptr(), 3, 10, int(rp[-1]), 2, 256,
                                        torch.cuda.current_stream().cuda_stream,
                                        ctypes.byref(h))
        assert st == want, (L, st)
        if st == 0:
            lib().mbx_plan_destroy(h)
    # row_ptr not running from 0 to nnz
    d = dev(np.array([1, 2, 3], np.int32), device)
    assert lib().mbx_plan_create_rows(d.data_ptr(), 2, 4, 3, 2, 256, None,
                                      ctypes.byref(h)) == -1
    d = dev(np.array([0, 2, 3], np.int32), device)
    assert lib().mbx_plan_create_rows(d.data_ptr(), 2, 4, 4, 2, 256, None,
                                      ctypes.byref(h)) == -1
    assert lib().mbx_plan_create_rows(None, 0, 0, 0, 8, 256, None, ctypes.byref(h)) == 0
    lib().mbx_plan_destroy(h)


def _vals(rng, kind, nnz, ncol):
    from merbit_amd import gen
    v = rng.standard_normal(nnz)
    x = rng.uniform(-1, 1, ncol)
    if kind == "f64":
        return v, x
    v32 = v.astype(np.float32)
    return (v32 if kind == "f32" else gen.to_bf16(v32)), x.astype(np.float32)


@pytest.mark.parametrize("kind", ["f64", "f32", "bf16"])
@pytest.mark.parametrize("ipt", [4, 8, 12])
def test_row_plan_products(orc, device, kind, ipt):
    """Products on row-aligned plans against the oracle and the plain plan,
    with every plan option that applies: stream policies, residency, FP64
    sums, 16-bit columns, alpha/beta, a clone and a graph replay."""
    from merbit_amd.spmv import MerbitPlan
    rng = np.random.default_rng(500 + ipt)
    T = 256 * ipt
    tol = 1e-12 if kind == "f64" else 1e-6
    for name, lens, ncol in short_row_cases(rng, T):
        rp, col = random_csr(rng, len(lens), ncol, lens=lens)
        nnz = int(rp[-1])
        vals, x = _vals(rng, kind, nnz, ncol)
        ya, absr = orc.spmv_accurate(rp, col, vals, x)
        d_rp, d_col, d_val, d_x = dev(rp, device), dev(col, device), dev(vals, device), \
            dev(x, device)
        plan = MerbitPlan(d_rp, ncol, nnz, items_per_thread=ipt, row_tiles=True)
        plain = MerbitPlan(d_rp, ncol, nnz, items_per_thread=ipt)
        ref = plain.spmv(d_rp, d_col, d_val, d_x).double().cpu().numpy()
        y0 = rng.uniform(-1, 1, len(lens))
        y0 = y0 if kind == "f64" else y0.astype(np.float32)
        accs = ["compensated"] if kind == "f64" else ["compensated", "f64", "f32prod"]
        for acc in accs:
            plan.set_accumulation(acc)
            for c16 in (False, True):
                if c16 and not plan.compress_cols(d_col):
                    continue
                for mode in ("cached", "nontemporal"):
                    plan.set_stream(mode)
                    for wg in (0, 3):
                        say(f"rows {kind} ipt {ipt} {name} {acc} col16 {c16} {mode} wg {wg}")
                        plan.set_residency(wg)
                        y = plan.spmv(d_rp, d_col, d_val, d_x).double().cpu().numpy()
                        assert cond_err(y, ya, absr) <= tol, name
                        assert cond_err(y, ref, absr) <= 2 * tol, name
                y = plan.spmv(d_rp, d_col, d_val, d_x, dev(y0, device), -1.5, 0.5)
                want = -1.5 * ya + 0.5 * y0.astype(np.float64)
                assert cond_err(y.double().cpu().numpy(), want,
                                1.5 * absr + 0.5 * np.abs(y0)) <= tol, name
                plan.compress_cols(None)
                plan.set_stream("auto")
                plan.set_residency(0)
        plan.set_accumulation("compensated")
        y1 = plan.spmv(d_rp, d_col, d_val, d_x)
        c = plan.clone()
        assert np.array_equal(bits(c.spmv(d_rp, d_col, d_val, d_x)), bits(y1))
        yg = torch.empty_like(y1)
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            plan.spmv(d_rp, d_col, d_val, d_x, yg)
        for _ in range(3):
            yg.fill_(7.0)
            torch.cuda.synchronize()
            g.replay()
            torch.cuda.synchronize()
            assert np.array_equal(bits(yg), bits(y1)), name
        del g
        c.close()
        plan.close()
        plain.close()


@pytest.mark.parametrize("kind", ["f64", "f32"])
def test_row_plan_hot_split(device, kind):
    """Short rows with one column in every row: the hot split on a row plan
    gives the row plan's plain bits."""
    from merbit_amd.spmv import MerbitPlan
    rng = np.random.default_rng(77)
    n, ncol = 60000, 50000
    lens = rng.integers(1, 20, n)
    rp, col = random_csr(rng, n, ncol, lens=lens)
    col[rp[:-1]] = 0  # column 0 first in every row (rows stay sorted)
    vals, x = _vals(rng, kind, int(rp[-1]), ncol)
    d_rp, d_col, d_val, d_x = dev(rp, device), dev(col, device), dev(vals, device), \
        dev(x, device)
    for ipt in (8, 12):
        plan = MerbitPlan(d_rp, ncol, int(rp[-1]), items_per_thread=ipt, row_tiles=True)
        ref = plan.spmv(d_rp, d_col, d_val, d_x)
        for nhot in (1, 512):
            say(f"rows hot {kind} ipt {ipt} nhot {nhot}")
            plan.split_hot(d_col, nhot)
            assert np.array_equal(bits(plan.spmv(d_rp, d_col, d_val, d_x)), bits(ref))
        plan.close()


def test_row_plan_spmm_and_sddmm(orc, device):
    """SpMM (k = 3, 8, 32) and the SDDMM on a row-aligned plan against the
    oracle column by column and the exact products g[row]*x[col]."""
    from merbit_amd._lib import lib
    from merbit_amd.spmv import MerbitPlan
    rng = np.random.default_rng(91)
    n, ncol = 40000, 30000
    rp, col = random_csr(rng, n, ncol, lens=rng.integers(0, 30, n))
    nnz = int(rp[-1])
    v = rng.standard_normal(nnz).astype(np.float32)
    d_rp, d_col, d_val = dev(rp, device), dev(col, device), dev(v, device)
    for ipt in (2, 4):
        plan = MerbitPlan(d_rp, ncol, nnz, items_per_thread=ipt, row_tiles=True)
        for k in (3, 8, 32):
            say(f"rows spmm ipt {ipt} k {k}")
            X = rng.uniform(-1, 1, (ncol, k)).astype(np.float32)
            Y = plan.spmm(d_rp, d_col, d_val, dev(X, device)).cpu().numpy()
            for j in range(k):
                ya, absr = orc.spmv_accurate(rp, col, v, X[:, j])
                assert cond_err(Y[:, j].astype(np.float64), ya, absr) <= 1e-5, (k, j)
        plan.close()
    plan = MerbitPlan(d_rp, ncol, nnz, items_per_thread=8, row_tiles=True)
    g = rng.standard_normal(n)
    x = rng.standard_normal(ncol)
    dval = torch.empty(nnz, dtype=torch.float64, device=device)
    assert lib().mbx_sddmm(plan.handle, 0, d_rp.data_ptr(), d_col.data_ptr(),
                           dev(g, device).data_ptr(), dev(x, device).data_ptr(),
                           dval.data_ptr(), torch.cuda.current_stream().cuda_stream) == 0
    rows = np.repeat(np.arange(n), np.diff(rp))
    assert np.array_equal(dval.cpu().numpy(), g[rows] * x[col])
    plan.close()


def test_rule_picks_row_tiles(orc, device):
    """The launch rule gives short-row patterns row-aligned tiles (and
    reports the longest row), long-row patterns merge-path tiles."""
    from merbit_amd.spmv import MerbitPlan, launch_rule
    from merbit_amd import workloads
    A, x = workloads.laplacian(256)
    d_rp, d_col = dev(A.row_ptr, device), dev(A.col, device)
    c = launch_rule(d_rp, d_col, A.ncol)
    assert c["max_row"] == 5 and c["row_tiles"] == 1
    plan = MerbitPlan(d_rp, A.ncol, A.nnz, col=d_col)
    assert plan.choice["row_tiles"] == 1
    assert plan.stats(d_rp)["split_tiles"] == 0
    ya, absr = orc.spmv_accurate(A.row_ptr, A.col, A.val, x)
    y = plan.spmv(d_rp, d_col, dev(A.val, device), dev(x, device)).cpu().numpy()
    assert cond_err(y, ya, absr) <= 1e-12
    plan.close()
    # a local (banded) pattern: row tiles at the 8 items per thread of FP64
    # values, not at the 12 that FP32 and BF16 values get
    from merbit_amd import gen
    B = gen.banded(320000, 13)
    d_rp, d_col = dev(B.row_ptr, device), dev(B.col, device)
    c64 = launch_rule(d_rp, d_col, B.ncol, value_dtype=torch.float64)
    c32 = launch_rule(d_rp, d_col, B.ncol, value_dtype=torch.float32)
    c16 = launch_rule(d_rp, d_col, B.ncol, value_dtype=torch.bfloat16)
    assert c64["rule"] == "local" and c64["items_per_thread"] == 8 and c64["row_tiles"] == 1
    for c in (c32, c16):
        assert c["rule"] == "local" and c["items_per_thread"] == 12 and c["row_tiles"] == 0
        assert c["accumulation"] == 2  # FP32 products, FP64 sums
    assert c64["max_row"] == 27
    # one row longer than 1/16 of the chosen tile: merge-path tiles
    rng = np.random.default_rng(5)
    lens = np.full(20000, 3)
    lens[7] = 5000
    rp, col = random_csr(rng, 20000, 20000, lens=lens)
    c = launch_rule(dev(rp, device), dev(col, device), 20000)
    assert c["max_row"] == 5000 and c["row_tiles"] == 0
    # no columns at all (every index out of range): the features see no
    # rows, plan creation falls back to merge-path tiles for the long row
    p0 = MerbitPlan(dev(rp, device), 0, int(rp[-1]), col=dev(col, device))
    assert p0.choice["row_tiles"] == 0 and p0.stats(dev(rp, device))["split_tiles"] > 0
    p0.close()
