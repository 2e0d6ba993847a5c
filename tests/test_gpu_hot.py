"""The hot-column split (mbx_plan_split_hot): products on a split plan run a
persistent tile kernel that reads the hot columns' x values from LDS; they
must be bit-identical to the plain product of the same plan shape, and the
plain product is itself checked against the oracle elsewhere (and here on
the full R-MAT)."""
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


@pytest.mark.parametrize("kind", ["f64", "f32", "bf16"])
@pytest.mark.parametrize("ipt", [8, 12])
def test_hot_split_bit_identical(device, kind, ipt):
    from merbit_amd.spmv import MerbitPlan
    rng = np.random.default_rng(11 + ipt)
    T = 256 * ipt
    cases = edge_case_patterns(T) + [
        ("pareto", (rng.pareto(1.05, 40000) * 4).astype(int), 30000)]
    for name, lens, ncol in cases:
        rp, col = random_csr(rng, len(lens), ncol, lens=lens)
        nnz = int(rp[-1])
        vals, x = values(rng, kind, nnz, ncol)
        d_rp, d_col, d_val, d_x = dev(rp, device), dev(col, device), dev(vals, device), \
            dev(x, device)
        plan = MerbitPlan(d_rp, ncol, nnz, items_per_thread=ipt)
        y0 = rng.uniform(-1, 1, len(lens)).astype(np.float64 if kind == "f64" else np.float32)
        ref = plan.spmv(d_rp, d_col, d_val, d_x)
        ref_ab = plan.spmv(d_rp, d_col, d_val, d_x, dev(y0, device), -1.5, 0.5)
        for nhot in (1, 64, 2048):
            say(f"hot {kind} ipt {ipt} {name} nhot {nhot}")
            plan.split_hot(d_col, nhot)
            for mode in ("cached", "nontemporal"):
                plan.set_stream(mode)
                for wg in (0, 3):
                    plan.set_residency(wg)
                    y = plan.spmv(d_rp, d_col, d_val, d_x)
                    assert np.array_equal(bits(y), bits(ref)), (name, nhot, mode, wg)
            y = plan.spmv(d_rp, d_col, d_val, d_x, dev(y0, device), -1.5, 0.5)
            assert np.array_equal(bits(y), bits(ref_ab)), (name, nhot)
            plan.set_stream("auto")
            plan.set_residency(0)
        plan.split_hot(None, 0)
        assert np.array_equal(bits(plan.spmv(d_rp, d_col, d_val, d_x)), bits(ref))
        plan.close()


def test_hot_split_coverage_and_fallback(device):
    """The reported share of the entries in the kept columns matches a count
    on the host (most frequent first, ties by column index); a hot set too
    large for one workgroup's LDS with FP64 x falls back to the plain
    kernel (same bits); clones do not inherit the split."""
    from merbit_amd.spmv import MerbitPlan
    rng = np.random.default_rng(5)
    rp, col = random_csr(rng, 20000, 40000, lens=(rng.pareto(1.1, 20000) * 5).astype(int))
    nnz = int(rp[-1])
    cnt = np.bincount(col, minlength=40000)
    order = np.lexsort((np.arange(40000), -cnt))  # count descending, column ascending
    d_rp, d_col = dev(rp, device), dev(col, device)
    plan = MerbitPlan(d_rp, 40000, nnz)
    for nhot in (1, 100, 5000):
        cov = plan.split_hot(d_col, nhot)
        assert cov == pytest.approx(cnt[order[:nhot]].sum() / nnz, rel=0, abs=1e-12)
    v = rng.standard_normal(nnz)
    x = rng.uniform(-1, 1, 40000)
    d_val, d_x = dev(v, device), dev(x, device)
    plan.split_hot(None, 0)
    ref = plan.spmv(d_rp, d_col, d_val, d_x)
    plan.split_hot(d_col, 32768)  # 256 KB of FP64 x: does not fit LDS
    assert np.array_equal(bits(plan.spmv(d_rp, d_col, d_val, d_x)), bits(ref))
    plan.split_hot(d_col, 4096)
    clone = plan.clone()
    assert np.array_equal(bits(clone.spmv(d_rp, d_col, d_val, d_x)), bits(ref))
    clone.close()
    plan.close()


@pytest.mark.parametrize("kind", ["f64", "f32"])
def test_hot_split_out_of_range_columns_read_zero(device, kind):
    """Column indices outside [0, ncol), negative ones included, read x as 0
    in the plain product (bounded gathers); a hot-column plan must do the
    same (a negative index must not be taken for a hot slot), so both
    products are bit-identical and match a host product that drops those
    entries."""
    from merbit_amd.spmv import MerbitPlan
    rng = np.random.default_rng(17)
    n = 20000
    rp, col = random_csr(rng, n, n, lens=(rng.pareto(1.1, n) * 5).astype(int))
    nnz = int(rp[-1])
    bad = col.astype(np.int64)
    picks = rng.choice(nnz, 64, replace=False)
    bad[picks[:32]] = -1 - rng.integers(0, 5000, 32)       # negative: would alias hot slots
    bad[picks[32:]] = n + rng.integers(0, 1 << 20, 32)      # past the end
    bad = bad.astype(np.int32)
    v, x = values(rng, kind, nnz, n)
    d_rp, d_bad, d_val, d_x = dev(rp, device), dev(bad, device), dev(v, device), dev(x, device)
    plan = MerbitPlan(d_rp, n, nnz, items_per_thread=12)
    ref = plan.spmv(d_rp, d_bad, d_val, d_x)
    say(f"out-of-range columns {kind}")
    plan.split_hot(d_bad, 4096)
    y = plan.spmv(d_rp, d_bad, d_val, d_x)
    assert np.array_equal(bits(y), bits(ref))
    rows = np.repeat(np.arange(n), np.diff(rp))
    ok = (bad >= 0) & (bad < n)
    vv = v.astype(np.float64) if kind == "f64" else v.astype(np.float32).astype(np.float64)
    host = np.zeros(n)
    np.add.at(host, rows[ok], vv[ok] * np.asarray(x, np.float64)[bad[ok]])
    absrow = np.zeros(n)
    np.add.at(absrow, rows[ok], np.abs(vv[ok] * np.asarray(x, np.float64)[bad[ok]]))
    tol = 1e-12 if kind == "f64" else 1e-6
    assert np.all(np.abs(y.double().cpu().numpy() - host) <= tol * absrow + 1e-300)
    plan.close()


@pytest.mark.parametrize("nnz,ncol", [(4095, 50000), (4096, 50000), (4097, 50000),
                                      (3 * 4096 - 1, 7), (3 * 4096 + 1, 1), (40000, 200000)])
def test_hot_split_counts_at_block_edges(device, nnz, ncol):
    """The column counts behind the split (device radix sort of the column
    indices, run-length encoded): one column holding 40 % of the entries,
    a long run of another, nnz at and around multiples of 4096, one-column
    and few-column patterns.  The covered share for every nhot prefix must
    match a host count."""
    from merbit_amd.spmv import MerbitPlan
    rng = np.random.default_rng(nnz + ncol)
    col = rng.integers(0, ncol, nnz)
    col[: nnz * 2 // 5] = 0                      # one column with 40 % of the entries
    col[nnz // 2: nnz // 2 + 300] = min(3, ncol - 1)  # a run of 300 entries
    rng.shuffle(col[nnz // 3:])
    col = col.astype(np.int32)
    rp = np.minimum(np.arange(0, nnz + 8, 8), nnz).astype(np.int32)
    if rp[-1] != nnz:
        rp = np.append(rp, nnz).astype(np.int32)
    nrow = len(rp) - 1
    cnt = np.sort(np.bincount(col, minlength=ncol))[::-1]
    d_rp, d_col = dev(rp, device), dev(col, device)
    plan = MerbitPlan(d_rp, ncol, nnz)
    for nhot in (1, 2, 3, 4, 10, 100, 1000):
        say(f"counts nnz {nnz} ncol {ncol} nrow {nrow} nhot {nhot}")
        cov = plan.split_hot(d_col, nhot)
        assert cov == pytest.approx(cnt[:nhot].sum() / nnz, rel=0, abs=1e-12)
    plan.close()


def test_hot_split_graph_replays(device):
    """A captured product of a split plan replays bit-identically."""
    from merbit_amd.batch import spmv_graph
    from merbit_amd.spmv import MerbitPlan
    rng = np.random.default_rng(9)
    rp, col = random_csr(rng, 30000, 30000, lens=(rng.pareto(1.1, 30000) * 6).astype(int))
    nnz = int(rp[-1])
    v, x = values(rng, "f64", nnz, 30000)
    d_rp, d_col, d_val, d_x = dev(rp, device), dev(col, device), dev(v, device), dev(x, device)
    plan = MerbitPlan(d_rp, 30000, nnz)
    ref = plan.spmv(d_rp, d_col, d_val, d_x)
    plan.split_hot(d_col, 4096)
    y = torch.zeros(30000, dtype=torch.float64, device=device)
    g = spmv_graph(plan, d_rp, d_col, d_val, d_x, y, repeats=3)
    for _ in range(3):
        y.zero_()
        torch.cuda.synchronize()
        g.launch()
        torch.cuda.synchronize()
        assert np.array_equal(bits(y), bits(ref))
    g.close()
    plan.close()


def _timed_call(plan, fn):
    b, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    b.record()  # the underlying events exist once recorded
    e.record()
    torch.cuda.synchronize()
    plan.set_events(b, e)
    out = fn()
    torch.cuda.synchronize()
    return out, b.elapsed_time(e)


def test_timing_events_on_tile_kernel(device):
    """Events set with set_events are attached to the next call's tile kernel
    (plain and hot-column FP64/FP32 plans, and an SpMM): that call's result is
    bit-identical to a call without them, the interval is positive, and the
    events are cleared after the one call (the next call leaves them as they
    were)."""
    from merbit_amd.spmv import MerbitPlan
    rng = np.random.default_rng(21)
    n = 40000
    rp, col = random_csr(rng, n, n, lens=(rng.pareto(1.1, n) * 5).astype(int))
    nnz = int(rp[-1])
    d_rp, d_col = dev(rp, device), dev(col, device)
    for kind in ("f64", "f32"):
        v, x = values(rng, kind, nnz, n)
        d_val, d_x = dev(v, device), dev(x, device)
        for nhot in (0, 2048):
            say(f"events {kind} nhot {nhot}")
            plan = MerbitPlan(d_rp, n, nnz, items_per_thread=12)
            if nhot:
                plan.split_hot(d_col, nhot)
            ref = plan.spmv(d_rp, d_col, d_val, d_x)
            y, ms = _timed_call(plan, lambda: plan.spmv(d_rp, d_col, d_val, d_x))
            assert np.array_equal(bits(y), bits(ref))
            assert 0.0 < ms < 1000.0
            b2, e2 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            b2.record()
            e2.record()
            torch.cuda.synchronize()
            before = b2.elapsed_time(e2)
            y2 = plan.spmv(d_rp, d_col, d_val, d_x)
            torch.cuda.synchronize()
            assert np.array_equal(bits(y2), bits(ref))
            assert b2.elapsed_time(e2) == before  # not attached to this call
            plan.close()
    say("events spmm k 8")
    v, _ = values(rng, "f32", nnz, n)
    d_val = dev(v, device)
    X = torch.from_numpy(rng.uniform(-1, 1, (n, 8)).astype(np.float32)).to(device)
    plan = MerbitPlan(d_rp, n, nnz, items_per_thread=4)
    ref = plan.spmm(d_rp, d_col, d_val, X)
    Y, ms = _timed_call(plan, lambda: plan.spmm(d_rp, d_col, d_val, X))
    assert np.array_equal(bits(Y.contiguous()), bits(ref.contiguous()))
    assert 0.0 < ms < 1000.0
    plan.close()


@pytest.mark.slow
@pytest.mark.parametrize("kind", ["f64", "f32"])
def test_hot_split_rmat_full(orc, device, kind):
    from merbit_amd import workloads
    from merbit_amd.spmv import MerbitPlan
    say(f"hot rmat scale 20 {kind}")
    A, x = workloads.rmat()
    vals = A.val if kind == "f64" else A.val.astype(np.float32)
    xs = x if kind == "f64" else x.astype(np.float32)
    d_rp, d_col, d_val, d_x = dev(A.row_ptr, device), dev(A.col, device), dev(vals, device), \
        dev(xs, device)
    plan = MerbitPlan(d_rp, A.ncol, A.nnz)
    ref = plan.spmv(d_rp, d_col, d_val, d_x)
    cov = plan.split_hot(d_col, 4096 if kind == "f64" else 5800)
    assert cov > 0.35
    plan.set_stream("nontemporal")
    y = plan.spmv(d_rp, d_col, d_val, d_x)
    assert np.array_equal(bits(y), bits(ref))
    ya, absr = orc.spmv_accurate(A.row_ptr, A.col, vals, xs)
    assert cond_err(y.cpu().numpy().astype(np.float64), ya, absr) <= \
        (1e-12 if kind == "f64" else 1e-6)
    plan.close()


@pytest.mark.parametrize("kind", ["f64", "f32"])
@pytest.mark.parametrize("ipt", [8, 12])
def test_sddmm_on_hot_split_plan_bit_identical(device, kind, ipt):
    """mbx_sddmm on a plan with a hot-column split gathers x for every entry
    on the caller's columns, as without the split (an LDS copy of the hot
    values cost the SDDMM more occupancy than it saved); dval must be
    bit-identical to the same plan's SDDMM without the split, also with
    column indices outside [0, ncol) (read as 0)."""
    from merbit_amd._lib import lib
    from merbit_amd.spmv import MerbitPlan
    rng = np.random.default_rng(51 + ipt)
    T = 256 * ipt
    cases = edge_case_patterns(T) + [
        ("pareto", (rng.pareto(1.05, 40000) * 4).astype(int), 30000)]
    dt = torch.float64 if kind == "f64" else torch.float32
    code = 0 if kind == "f64" else 1
    s = torch.cuda.current_stream().cuda_stream
    for name, lens, ncol in cases:
        rp, col = random_csr(rng, len(lens), ncol, lens=lens)
        nnz = int(rp[-1])
        if nnz > 8:
            col[3] = -2
            col[nnz - 2] = ncol + 5
        d_rp, d_col = dev(rp, device), dev(col, device)
        g = torch.from_numpy(rng.uniform(-1, 1, len(lens))).to(device, dt)
        x = torch.from_numpy(rng.uniform(-1, 1, ncol)).to(device, dt)
        plan = MerbitPlan(d_rp, ncol, nnz, items_per_thread=ipt)

        def dval():
            out = torch.zeros(max(nnz, 1), dtype=dt, device=device)
            assert lib().mbx_sddmm(plan.handle, code, d_rp.data_ptr(), d_col.data_ptr(),
                                   g.data_ptr(), x.data_ptr(), out.data_ptr(), s) == 0
            return out
        ref = dval()
        for nhot in (1, 64, 2048):
            say(f"sddmm hot {kind} ipt {ipt} {name} nhot {nhot}")
            plan.split_hot(d_col, nhot)
            assert torch.equal(dval().view(torch.uint8), ref.view(torch.uint8)), (name, nhot)
        plan.close()
