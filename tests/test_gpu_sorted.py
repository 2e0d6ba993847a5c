"""Tile-sorted gathers (mbx_plan_sort_cols, mbx_batch_sort_cols): products
on a sorted plan gather x in each tile's sorted column order and put every
value back into its element's product slot, so they must be bit-identical
to the plain product of the same plan shape (itself checked against the
oracle elsewhere, and here on the full R-MAT)."""
from __future__ import annotations

import numpy as np
import pytest
import torch

from helpers import cond_err, edge_case_patterns, random_csr

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[case] {msg}", flush=True)


def dev(a, device):
    return torch.from_numpy(np.ascontiguousarray(a)).to(device)


def values(rng, kind, nnz, ncol):
    from merbit_amd import gen
    v = rng.standard_normal(nnz)
    x = rng.uniform(-1, 1, ncol)
    if kind == "f64":
        return v, x
    v32 = v.astype(np.float32)
    return (v32 if kind == "f32" else gen.to_bf16(v32)), x.astype(np.float32)


def bits(t):
    a = t.cpu().numpy()
    return a.view(np.uint64 if a.dtype == np.float64 else np.uint32)


def uniform_csr(rng, nrow, ncol, lam):
    """Rows of Poisson(lam) entries with uniform columns (sorted per row)."""
    return random_csr(rng, nrow, ncol, lens=rng.poisson(lam, nrow))


@pytest.mark.parametrize("kind", ["f64", "f32", "bf16"])
@pytest.mark.parametrize("ipt", [8, 12, 16, 32])
def test_sorted_bit_identical(orc, device, kind, ipt):
    """32 items per thread is the 8192-item tile of 512 threads x 16 items,
    whose plain product is checked against the oracle here too."""
    from merbit_amd.spmv import MerbitPlan
    rng = np.random.default_rng(71 + ipt)
    T = 256 * ipt
    cases = edge_case_patterns(T) + [
        ("pareto", (rng.pareto(1.05, 40000) * 4).astype(int), 30000),
        ("uniform", rng.poisson(20, 30000), 25000),
        ("one_column", [3] * 5000, 1),
    ]
    accs = ["compensated", "f64", "f32prod"] if kind != "f64" else [None]
    for name, lens, ncol in cases:
        rp, col = random_csr(rng, len(lens), ncol, lens=lens)
        nnz = int(rp[-1])
        vals, x = values(rng, kind, nnz, ncol)
        d_rp, d_col, d_val, d_x = dev(rp, device), dev(col, device), dev(vals, device), \
            dev(x, device)
        y0 = rng.uniform(-1, 1, len(lens)).astype(np.float64 if kind == "f64" else np.float32)
        plan = MerbitPlan(d_rp, ncol, nnz, items_per_thread=ipt)
        for acc in accs:
            if acc:
                plan.set_accumulation(acc)
            plan.sort_cols(None)
            ref = plan.spmv(d_rp, d_col, d_val, d_x)
            ref_ab = plan.spmv(d_rp, d_col, d_val, d_x, dev(y0, device), -1.5, 0.5)
            say(f"sorted {kind} ipt {ipt} {name} acc {acc}")
            if ipt == 32:
                ya, absr = orc.spmv_accurate(rp, col, vals, x)
                assert cond_err(ref.double().cpu().numpy(), ya, absr) <= \
                    (1e-12 if kind == "f64" else 2.5e-7), (name, acc)
            assert plan.sort_cols(d_col)
            for mode in ("cached", "nontemporal"):
                plan.set_stream(mode)
                for wg in (0, 3):
                    plan.set_residency(wg)
                    y = plan.spmv(d_rp, d_col, d_val, d_x)
                    assert np.array_equal(bits(y), bits(ref)), (name, acc, mode, wg)
            plan.set_stream("auto")
            plan.set_residency(0)
            y = plan.spmv(d_rp, d_col, d_val, d_x, dev(y0, device), -1.5, 0.5)
            assert np.array_equal(bits(y), bits(ref_ab)), (name, acc)
        plan.sort_cols(None)
        assert np.array_equal(bits(plan.spmv(d_rp, d_col, d_val, d_x)), bits(ref))
        plan.close()


@pytest.mark.parametrize("ipt", [8])
def test_sorted_row_tiles(device, ipt):
    """Row-aligned tiles hold at most a tile of items and need no fixup; a
    sorted plan over them is bit-identical to the same row-tile plan."""
    from merbit_amd.spmv import MerbitPlan
    rng = np.random.default_rng(5)
    for n, lam, ncol in ((20000, 9, 20000), (3000, 30, 90000)):
        rp, col = uniform_csr(rng, n, ncol, lam)
        nnz = int(rp[-1])
        for kind in ("f64", "f32"):
            say(f"sorted row tiles {kind} n {n}")
            vals, x = values(rng, kind, nnz, ncol)
            d_rp, d_col, d_val, d_x = dev(rp, device), dev(col, device), dev(vals, device), \
                dev(x, device)
            plan = MerbitPlan(d_rp, ncol, nnz, items_per_thread=ipt, row_tiles=True)
            ref = plan.spmv(d_rp, d_col, d_val, d_x)
            assert plan.sort_cols(d_col)
            assert np.array_equal(bits(plan.spmv(d_rp, d_col, d_val, d_x)), bits(ref))
            plan.close()


def test_sorted_refusals_and_exclusions(device):
    """Shapes and sizes the sorted form does not take leave the plan as it
    was (False); the hot split and 16-bit columns exclude it both ways; a
    column outside [0, ncol) is an error that keeps no copy; clones read the
    caller's columns; the SDDMM of a sorted plan is unchanged."""
    from merbit_amd._lib import MerbitError
    from merbit_amd.spmv import MerbitPlan
    rng = np.random.default_rng(17)
    n = 6000
    rp, col = uniform_csr(rng, n, n, 12)
    nnz = int(rp[-1])
    v, x = values(rng, "f64", nnz, n)
    d_rp, d_col, d_val, d_x = dev(rp, device), dev(col, device), dev(v, device), dev(x, device)
    say("sorted refusals")
    p4 = MerbitPlan(d_rp, n, nnz, items_per_thread=4)
    assert not p4.sort_cols(d_col)
    p4.close()
    # ncol limit: 2^20 at 12 items per thread (columns drawn below n)
    big = MerbitPlan(d_rp, (1 << 20) + 1, nnz, items_per_thread=12)
    assert not big.sort_cols(d_col)
    big.close()
    ok = MerbitPlan(d_rp, 1 << 20, nnz, items_per_thread=12)
    assert ok.sort_cols(d_col)
    ok.close()
    plan = MerbitPlan(d_rp, n, nnz, items_per_thread=12)
    ref = plan.spmv(d_rp, d_col, d_val, d_x)
    plan.split_hot(d_col, 256)
    assert not plan.sort_cols(d_col)
    plan.split_hot(None, 0)
    assert plan.compress_cols(d_col)
    assert not plan.sort_cols(d_col)
    plan.compress_cols(None)
    assert plan.sort_cols(d_col)
    assert not plan.compress_cols(d_col)
    with pytest.raises(MerbitError):
        plan.split_hot(d_col, 256)
    assert np.array_equal(bits(plan.spmv(d_rp, d_col, d_val, d_x)), bits(ref))
    clone = plan.clone()
    assert np.array_equal(bits(clone.spmv(d_rp, d_col, d_val, d_x)), bits(ref))
    clone.close()
    # the SDDMM reads the caller's columns
    from merbit_amd._lib import lib
    g = torch.from_numpy(rng.uniform(-1, 1, n)).to(device)
    s = torch.cuda.current_stream().cuda_stream
    p2 = MerbitPlan(d_rp, n, nnz, items_per_thread=12)

    def dval():
        out = torch.zeros(nnz, dtype=torch.float64, device=device)
        assert lib().mbx_sddmm(p2.handle, 0, d_rp.data_ptr(), d_col.data_ptr(), g.data_ptr(),
                               d_x.data_ptr(), out.data_ptr(), s) == 0
        return out
    dv_ref = dval()
    assert p2.sort_cols(d_col)
    assert torch.equal(dval().view(torch.uint8), dv_ref.view(torch.uint8))
    p2.close()
    # an out-of-range column: an error, and the plan keeps no sorted copy
    bad = col.copy()
    bad[nnz // 2] = n + 3
    d_bad = dev(bad, device)
    with pytest.raises(MerbitError):
        plan.sort_cols(d_bad)
    assert np.array_equal(bits(plan.spmv(d_rp, d_col, d_val, d_x)), bits(ref))
    plan.close()
    # an empty pattern has nothing to sort
    e_rp = dev(np.zeros(101, np.int32), device)
    pe = MerbitPlan(e_rp, 50, 0, items_per_thread=8)
    assert pe.sort_cols(dev(np.zeros(0, np.int32), device))
    y = pe.spmv(e_rp, dev(np.zeros(0, np.int32), device),
                dev(np.zeros(0, np.float64), device), dev(np.ones(50, np.float64), device))
    assert not torch.any(y != 0)
    pe.close()


def test_sorted_graph_replays(device):
    """A captured product of a sorted plan replays bit-identically."""
    from merbit_amd.batch import spmv_graph
    from merbit_amd.spmv import MerbitPlan
    rng = np.random.default_rng(19)
    n = 30000
    rp, col = random_csr(rng, n, n, lens=(rng.pareto(1.1, n) * 6).astype(int))
    nnz = int(rp[-1])
    v, x = values(rng, "f64", nnz, n)
    d_rp, d_col, d_val, d_x = dev(rp, device), dev(col, device), dev(v, device), dev(x, device)
    plan = MerbitPlan(d_rp, n, nnz, items_per_thread=12)
    ref = plan.spmv(d_rp, d_col, d_val, d_x)
    assert plan.sort_cols(d_col)
    y = torch.zeros(n, dtype=torch.float64, device=device)
    g = spmv_graph(plan, d_rp, d_col, d_val, d_x, y, repeats=3)
    for _ in range(3):
        y.zero_()
        torch.cuda.synchronize()
        g.launch()
        torch.cuda.synchronize()
        assert np.array_equal(bits(y), bits(ref))
    g.close()
    plan.close()


@pytest.mark.parametrize("kind", ["f64", "f32", "bf16"])
@pytest.mark.parametrize("ipt", [8, 12, 16])
def test_batch_sorted_bit_identical(device, kind, ipt):
    """A batch with its members' columns tile-sorted gives the plain batch
    product's bits, for every member (empty ones included), with both
    stream policies, a residency cap and alpha/beta; dropping the copies
    restores the plain launch."""
    from merbit_amd.batch import MerbitBatch
    rng = np.random.default_rng(101 + ipt)
    sizes = [(0, 5, 1), (1, 50, 3), (1200, 1200, 9), (40000, 40000, 25), (777, 5000, 40),
             (25000, 600, 14), (3000, 3000, 0), (90000, 90000, 6)]
    items, ys = [], []
    ydt = torch.float64 if kind == "f64" else torch.float32
    for nrow, ncol, lam in sizes:
        rp, col = uniform_csr(rng, nrow, ncol, lam) if nrow else (np.zeros(1, np.int32),
                                                                  np.zeros(0, np.int32))
        v, x = values(rng, kind, int(rp[-1]), ncol)
        y = torch.full((nrow,), float("nan"), dtype=ydt, device=device)
        items.append((dev(rp, device), dev(col, device), dev(v, device), dev(x, device), y, ncol))
        ys.append(y)
    say(f"batch sorted {kind} ipt {ipt}")
    b = MerbitBatch(items, items_per_thread=ipt)
    b.spmv()
    torch.cuda.synchronize()
    ref = [bits(y) for y in ys]
    y0 = [y.clone() for y in ys]
    b.spmv(-1.5, 0.5)
    torch.cuda.synchronize()
    ref_ab = [bits(y) for y in ys]
    assert b.sort_cols(True)
    for mode in ("cached", "nontemporal"):
        b.set_stream(mode)
        for wg in (0, 3):
            b.set_residency(wg)
            b.spmv()
            torch.cuda.synchronize()
            for i, y in enumerate(ys):
                assert np.array_equal(bits(y), ref[i]), (i, mode, wg)
    b.set_stream("auto")
    b.set_residency(0)
    for y, yy in zip(ys, y0):
        y.copy_(yy)
    b.spmv(-1.5, 0.5)
    torch.cuda.synchronize()
    for i, y in enumerate(ys):
        assert np.array_equal(bits(y), ref_ab[i]), i
    assert not b.sort_cols(False)
    b.spmv()
    torch.cuda.synchronize()
    for i, y in enumerate(ys):
        assert np.array_equal(bits(y), ref[i]), i
    b.close()


@pytest.mark.parametrize("kind", ["f64", "f32", "bf16"])
def test_batch_wide_tiles(orc, device, kind):
    """Batches of 8192-item tiles (items_per_thread 32: 512 threads of 16
    items), plain and with sorted gathers: every member within its bar of
    the oracle (empty members included), the sorted product bit-identical
    to the plain one, with and without a residency cap; members beyond 2^19
    columns refuse the sort."""
    from merbit_amd.batch import MerbitBatch
    rng = np.random.default_rng(131)
    sizes = [(0, 5, 1), (1, 50, 3), (900, 900, 9), (60000, 60000, 25), (777, 5000, 40),
             (25000, 600, 14), (3000, 3000, 0), (120000, 150000, 6)]
    items, host = [], []
    ydt = torch.float64 if kind == "f64" else torch.float32
    for nrow, ncol, lam in sizes:
        rp, col = uniform_csr(rng, nrow, ncol, lam) if nrow else (np.zeros(1, np.int32),
                                                                  np.zeros(0, np.int32))
        v, x = values(rng, kind, int(rp[-1]), ncol)
        y = torch.full((nrow,), float("nan"), dtype=ydt, device=device)
        items.append((dev(rp, device), dev(col, device), dev(v, device), dev(x, device), y, ncol))
        host.append((rp, col, v, x))
    tol = 1e-12 if kind == "f64" else 1e-6
    say(f"batch wide tiles {kind}")
    b = MerbitBatch(items, items_per_thread=32)
    assert b.tile_items == 8192
    b.spmv()
    torch.cuda.synchronize()
    ref = []
    for (rp, col, v, x), it in zip(host, items):
        ya, absr = orc.spmv_accurate(rp, col, v, x)
        got = it[4].double().cpu().numpy()
        assert cond_err(got, ya, absr) <= tol
        ref.append(bits(it[4]))
    assert b.sort_cols(True)
    for wg in (0, 1):
        b.set_residency(wg)
        b.spmv()
        torch.cuda.synchronize()
        for i, it in enumerate(items):
            assert np.array_equal(bits(it[4]), ref[i]), (i, wg)
    b.close()
    # a member of more than 2^19 columns: no sorted form at this shape
    big = (items[3][0], items[3][1], items[3][2],
           torch.zeros((1 << 19) + 1, dtype=items[3][3].dtype, device=device), items[3][4],
           (1 << 19) + 1)
    b = MerbitBatch([items[2], big], items_per_thread=32)
    assert not b.sort_cols(True)
    b.close()


def test_batch_sorted_refusals(device):
    from merbit_amd.batch import MerbitBatch
    rng = np.random.default_rng(3)
    rp, col = uniform_csr(rng, 2000, 2000, 8)
    nnz = int(rp[-1])
    v, x = values(rng, "f32", nnz, 2000)
    item = (dev(rp, device), dev(col, device), dev(v, device), dev(x, device),
            torch.empty(2000, dtype=torch.float32, device=device), 2000)
    say("batch sorted refusals")
    b = MerbitBatch([item], items_per_thread=4)
    assert not b.sort_cols(True)
    b.close()
    big = (item[0], item[1], item[2], torch.zeros((1 << 20) + 1, dtype=torch.float32,
                                                  device=device), item[4], (1 << 20) + 1)
    b = MerbitBatch([item, big], items_per_thread=16)
    assert not b.sort_cols(True)
    b.close()


@pytest.mark.parametrize("kind", ["f64", "f32"])
def test_colblock_sorted_bit_identical(device, kind):
    """The column-blocked form with tile-sorted block products gives the
    plain form's bits for the forward product and for dx through its A^T
    form, whether sorted before or after the A^T form is prepared; IPT 4
    blocks refuse; dropping restores the plain gathers."""
    from merbit_amd.spmv import ColumnBlockedCsr
    rng = np.random.default_rng(61)
    n = 60000
    rp, col = uniform_csr(rng, n, n, 14)
    nnz = int(rp[-1])
    v, x = values(rng, kind, nnz, n)
    g = rng.uniform(-1, 1, n).astype(v.dtype)
    d_rp, d_col, d_val, d_x, d_g = (dev(a, device) for a in (rp, col, v, x, g))
    dt = torch.float64 if kind == "f64" else torch.float32
    for order in ("before", "after"):
        say(f"colblock sorted {kind} {order} prepare")
        cb = ColumnBlockedCsr(d_rp, d_col, n, nblocks=3, items_per_thread=12)
        vb = cb.values(d_val)
        y, dx = torch.empty(n, dtype=dt, device=device), torch.empty(n, dtype=dt, device=device)
        cb.prepare_adjoint(d_rp, d_col, nblocks=2)
        cb.spmv(vb, d_x, y)
        cb.adjoint_x(vb, d_g, dx)
        y_ref, dx_ref = bits(y), bits(dx)
        cb.close()
        cb = ColumnBlockedCsr(d_rp, d_col, n, nblocks=3, items_per_thread=12)
        if order == "before":
            assert cb.sort_cols(True)
        cb.prepare_adjoint(d_rp, d_col, nblocks=2)
        if order == "after":
            assert cb.sort_cols(True)
        for wg in (0, 3):
            cb.set_residency(wg)
            cb.spmv(vb, d_x, y)
            cb.adjoint_x(vb, d_g, dx)
            assert np.array_equal(bits(y), y_ref) and np.array_equal(bits(dx), dx_ref), wg
        assert not cb.sort_cols(False)
        cb.spmv(vb, d_x, y)
        assert np.array_equal(bits(y), y_ref)
        cb.close()
    cb4 = ColumnBlockedCsr(d_rp, d_col, n, nblocks=2, items_per_thread=4)
    assert not cb4.sort_cols(True)
    cb4.close()


def test_rule_takes_sorted_gathers(orc, device):
    """The launch rule takes tile-sorted gathers for non-local patterns whose
    columns fit the sorted words: at 32 items per thread (8192-item tiles)
    for uniform random patterns whose x fits L2 (with few tiles or more), at
    16 for one with a hot column set; not for a
    local pattern, a pattern too small for 2 tiles per CU at 16 items per
    thread, or one with more than 2^20 columns (which keeps the hot split).
    The rule's plans match the oracle."""
    from merbit_amd.spmv import MerbitPlan, launch_rule
    rng = np.random.default_rng(41)
    say("rule sorted: uniform, x fits")
    rp, col = uniform_csr(rng, 300000, 300000, 40)
    c = launch_rule(dev(rp, device), dev(col, device), 300000, value_dtype=torch.float32)
    assert c["rule"] == "x_fits" and c["sort_cols"] == 1 and c["items_per_thread"] == 32, c
    say("rule sorted: uniform, few tiles")
    rp, col = uniform_csr(rng, 200000, 200000, 30)
    d_rp, d_col = dev(rp, device), dev(col, device)
    c = launch_rule(d_rp, d_col, 200000)
    assert c["rule"] == "few_tiles" and c["sort_cols"] == 1 and c["items_per_thread"] == 32, c
    v, x = values(rng, "f64", int(rp[-1]), 200000)
    plan = MerbitPlan(d_rp, 200000, int(rp[-1]), col=d_col)
    assert plan.choice["sort_cols"] == 1
    y = plan.spmv(d_rp, d_col, dev(v, device), dev(x, device)).cpu().numpy()
    ya, absr = orc.spmv_accurate(rp, col, v, x)
    assert cond_err(y, ya, absr) <= 1e-12
    plan.close()
    say("rule sorted: small and local patterns")
    rp, col = uniform_csr(rng, 1631, 1631, 30)
    c = launch_rule(dev(rp, device), dev(col, device), 1631)
    assert c["rule"] == "few_tiles" and c["sort_cols"] == 0 and c["items_per_thread"] == 2, c
    n = 400000
    rp = np.arange(0, 3 * n + 1, 3, dtype=np.int32)
    col = np.stack([np.maximum(np.arange(n) - 1, 0), np.arange(n),
                    np.minimum(np.arange(n) + 1, n - 1)], 1).reshape(-1).astype(np.int32)
    c = launch_rule(dev(rp, device), dev(col, device), n)  # few tiles, all local
    assert c["max_tile_span"] < 65536 and c["sort_cols"] == 0, c
    say("rule sorted: hot columns beyond 2^20 columns")
    ncol = (1 << 20) + 5
    nrow = 400000
    lens = rng.poisson(20, nrow)
    rp = np.zeros(nrow + 1, np.int32)
    rp[1:] = np.cumsum(lens)
    nnz = int(rp[-1])
    hot = rng.random(nnz) < 0.5
    col = np.where(hot, rng.integers(0, 64, nnz), rng.integers(0, ncol, nnz)).astype(np.int32)
    d_rp, d_col = dev(rp, device), dev(col, device)
    c = launch_rule(d_rp, d_col, ncol)
    assert c["rule"] == "hot_columns" and c["sort_cols"] == 0 and c["hot_columns"] > 0, c
    col2 = np.minimum(col, (1 << 20) - 1)  # the same pattern within 2^20 columns
    c = launch_rule(d_rp, dev(col2, device), 1 << 20)
    assert c["rule"] == "hot_columns" and c["sort_cols"] == 1 and c["hot_columns"] == 0, c


def test_autotune_form(device):
    """The timing autotuner over launch shapes and sorted gathers returns a
    shape the plan takes, whose product matches the plain product of that
    shape bit for bit."""
    from merbit_amd.batch import autotune_spmv_form
    from merbit_amd.spmv import MerbitPlan
    rng = np.random.default_rng(23)
    rp, col = uniform_csr(rng, 60000, 60000, 20)
    nnz = int(rp[-1])
    v, x = values(rng, "f64", nnz, 60000)
    d_rp, d_col, d_val, d_x = dev(rp, device), dev(col, device), dev(v, device), dev(x, device)
    y = torch.empty(60000, dtype=torch.float64, device=device)
    say("autotune form")
    ipt, mode, wg, srt, us = autotune_spmv_form(d_rp, d_col, d_val, d_x, y, reps=3)
    assert ipt in (2, 4, 6, 8, 10, 12, 14, 16) and mode in ("cached", "nontemporal")
    assert wg in (0, 3, 4, 5) and us > 0 and (not srt or ipt in (8, 12, 16))
    ref = MerbitPlan(d_rp, 60000, nnz, items_per_thread=ipt)
    plan = MerbitPlan(d_rp, 60000, nnz, items_per_thread=ipt)
    plan.set_stream(mode)
    plan.set_residency(wg)
    if srt:
        assert plan.sort_cols(d_col)
    assert np.array_equal(bits(plan.spmv(d_rp, d_col, d_val, d_x)),
                          bits(ref.spmv(d_rp, d_col, d_val, d_x)))
    plan.close()
    ref.close()


@pytest.mark.slow
@pytest.mark.parametrize("kind", ["f64", "f32"])
def test_sorted_rmat_full(orc, device, kind):
    from merbit_amd import workloads
    from merbit_amd.spmv import MerbitPlan
    say(f"sorted rmat scale 20 {kind}")
    A, x = workloads.rmat()
    vals = A.val if kind == "f64" else A.val.astype(np.float32)
    xs = x if kind == "f64" else x.astype(np.float32)
    d_rp, d_col, d_val, d_x = dev(A.row_ptr, device), dev(A.col, device), dev(vals, device), \
        dev(xs, device)
    plan = MerbitPlan(d_rp, A.ncol, A.nnz, items_per_thread=12)
    ref = plan.spmv(d_rp, d_col, d_val, d_x)
    assert plan.sort_cols(d_col)
    plan.set_stream("nontemporal")
    y = plan.spmv(d_rp, d_col, d_val, d_x)
    assert np.array_equal(bits(y), bits(ref))
    ya, absr = orc.spmv_accurate(A.row_ptr, A.col, vals, xs)
    assert cond_err(y.cpu().numpy().astype(np.float64), ya, absr) <= \
        (1e-12 if kind == "f64" else 1e-6)
    plan.close()
