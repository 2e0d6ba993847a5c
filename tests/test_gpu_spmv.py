"""GPU parity of the merge-path SpMV (libmerbit_amd.so through the C ABI)
against the CPU oracle.  FP64 bar: condition-scaled error <= 1e-12 against
the Dot2-accurate oracle and against the reference-order emulator; mixed
precision bar: <= 1e-6 against the FP64 oracle on the same stored values."""
from __future__ import annotations

import math

import numpy as np
import pytest
import torch

from helpers import cond_err, edge_case_patterns, norm_err, random_csr

pytestmark = pytest.mark.gpu

F64_TOL = 1e-12
MIXED_TOL = 1e-6


def say(msg):
    print(f"[case] {msg}", flush=True)


def dev(a, device):
    return torch.from_numpy(np.ascontiguousarray(a)).to(device)


def run_f64(rp, col, val, x, device, ipt=8, alpha=1.0, beta=0.0, y0=None):
    from merbit_amd.spmv import MerbitPlan
    d_rp, d_col, d_val, d_x = (dev(rp, device), dev(col, device), dev(val, device),
                               dev(x, device))
    plan = MerbitPlan(d_rp, len(x), int(rp[-1]), items_per_thread=ipt)
    y = None if y0 is None else dev(y0, device)
    out = plan.spmv(d_rp, d_col, d_val, d_x, y, alpha, beta)
    torch.cuda.synchronize()
    res = out.cpu().numpy()
    plan.close()
    return res


@pytest.mark.parametrize("ipt", [4, 8, 16])
def test_partition_bit_exact(orc, device, ipt):
    from merbit_amd.spmv import MerbitPlan
    rng = np.random.default_rng(ipt)
    for name, lens, ncol in edge_case_patterns(256 * ipt) + [
            ("random", rng.poisson(6, 5000), 5000)]:
        say(f"partition {name} ipt={ipt}")
        rp, _ = random_csr(rng, len(lens), ncol, lens=lens)
        plan = MerbitPlan(dev(rp, device), ncol, int(rp[-1]), items_per_thread=ipt)
        rows, nnz = plan.partition()
        T = plan.tile_items
        path = len(rp) - 1 + int(rp[-1])
        for t in range(plan.ntiles + 1):
            assert (nnz[t], rows[t]) == orc.merge_search(rp, int(rp[-1]), min(t * T, path))
        plan.close()


@pytest.mark.parametrize("ipt", [4, 6, 8, 10, 12, 14, 16])
def test_f64_edge_cases(orc, device, ipt):
    rng = np.random.default_rng(1000 + ipt)
    for name, lens, ncol in edge_case_patterns(256 * ipt):
        say(f"f64 edge {name} ipt={ipt}")
        rp, col = random_csr(rng, len(lens), ncol, lens=lens)
        val = rng.standard_normal(int(rp[-1]))
        x = rng.uniform(-1, 1, ncol)
        y = run_f64(rp, col, val, x, device, ipt)
        ya, absr = orc.spmv_accurate(rp, col, val, x)
        assert cond_err(y, ya, absr) <= F64_TOL, name
        if ncol == len(lens):
            ye = orc.merbit_emulate(rp, col, val, x, sigma=14)
            assert cond_err(y, ye, absr) <= F64_TOL, name


def test_f64_laplacian_closed_form(orc, device):
    from merbit_amd import workloads
    say("laplacian 256 x=1 closed form")
    A, x = workloads.laplacian(256)
    g = 256
    y = run_f64(A.row_ptr, A.col, A.val, np.ones(A.nrow), device)
    r, c = np.divmod(np.arange(g * g), g)
    want = (r == 0).astype(float) + (r == g - 1) + (c == 0) + (c == g - 1)
    assert np.array_equal(y, want)
    say("laplacian 256 x~U[-1,1)")
    y = run_f64(A.row_ptr, A.col, A.val, x, device)
    ya, absr = orc.spmv_accurate(A.row_ptr, A.col, A.val, x)
    assert cond_err(y, ya, absr) <= F64_TOL
    ye = orc.merbit_emulate(A.row_ptr, A.col, A.val, x, sigma=14)
    assert cond_err(y, ye, absr) <= F64_TOL


def test_f64_alpha_beta(orc, device):
    rng = np.random.default_rng(77)
    rp, col = random_csr(rng, 20000, 20000, lens=(rng.pareto(1.1, 20000) * 4).astype(int))
    val = rng.standard_normal(int(rp[-1]))
    x = rng.uniform(-1, 1, 20000)
    y0 = rng.uniform(-1, 1, 20000)
    say("alpha/beta")
    y = run_f64(rp, col, val, x, device, alpha=-1.5, beta=0.25, y0=y0)
    ya, absr = orc.spmv_accurate(rp, col, val, x)
    want = -1.5 * ya + 0.25 * y0
    assert np.max(np.abs(y - want) / (1.5 * absr + 0.25 * np.abs(y0) + 1e-300)) <= F64_TOL
    # beta == 0 must not read y: NaN garbage in y is overwritten
    y = run_f64(rp, col, val, x, device, alpha=2.0, beta=0.0,
                y0=np.full(20000, np.nan))
    assert cond_err(y, 2.0 * ya, 2.0 * absr) <= F64_TOL


@pytest.mark.parametrize("kind", ["f64", "f32", "bf16"])
def test_repeat_calls_bit_identical(device, kind):
    """The carry completion sums in a fixed order: repeated calls agree bit
    for bit, including rows that span thousands of tiles."""
    from merbit_amd import gen
    from merbit_amd.spmv import MerbitPlan
    rng = np.random.default_rng(31)
    cases = edge_case_patterns(2048) + [
        ("pareto", (rng.pareto(0.9, 40000) * 6).astype(int), 40000)]
    for name, lens, ncol in cases:
        say(f"repeat {kind} {name}")
        rp, col = random_csr(rng, len(lens), ncol, lens=lens)
        v = rng.standard_normal(int(rp[-1]))
        x = rng.uniform(-1, 1, ncol)
        if kind == "f64":
            vals, xs = v, x
        else:
            v32 = v.astype(np.float32)
            vals = v32 if kind == "f32" else gen.to_bf16(v32)
            xs = x.astype(np.float32)
        d_rp, d_col, d_val, d_x = (dev(rp, device), dev(col, device), dev(vals, device),
                                   dev(xs, device))
        plan = MerbitPlan(d_rp, ncol, int(rp[-1]))
        outs = [plan.spmv(d_rp, d_col, d_val, d_x).cpu().numpy() for _ in range(4)]
        plan.close()
        for o in outs[1:]:
            assert np.array_equal(o.view(np.uint8), outs[0].view(np.uint8)), name


@pytest.mark.slow
def test_f64_rmat_full(orc, device):
    from merbit_amd import workloads
    from merbit_amd.spmv import DeviceCsr
    say("rmat scale 20, 2^24 entries, FP64")
    A, x = workloads.rmat()
    M = DeviceCsr(A, device)
    d_x = dev(x, device)
    y1 = M.spmv(d_x).cpu().numpy()
    y2 = M.spmv(d_x).cpu().numpy()
    assert np.array_equal(y1, y2), "result must be deterministic"
    ya, absr = orc.spmv_accurate(A.row_ptr, A.col, A.val, x)
    assert cond_err(y1, ya, absr) <= F64_TOL
    ye = orc.merbit_emulate(A.row_ptr, A.col, A.val, x, sigma=14)
    assert cond_err(y1, ye, absr) <= F64_TOL
    M.plan.close()


@pytest.mark.slow
@pytest.mark.parametrize("kind", ["f32", "bf16"])
def test_rmat_mixed_full(orc, device, kind):
    """The R-MAT half of the mixed-precision workload at full size (2^20
    rows, 2^24 entries): values stored FP32 or bfloat16, x FP32, against the
    oracle's accurate FP64 product on the same stored values, with both the
    compensated FP32 sums and the FP64 sums.  Bar 1e-6, condition-scaled and
    relative on the rows whose result is not a cancellation
    (|y| > 1e-3 (|A||x|))."""
    from merbit_amd import gen, workloads
    from merbit_amd.spmv import MerbitPlan
    say(f"rmat scale 20, 2^24 entries, {kind} values")
    A, x = workloads.rmat()
    v32 = A.val.astype(np.float32)
    vals = v32 if kind == "f32" else gen.to_bf16(v32)
    xs = x.astype(np.float32)
    ya, absr = orc.spmv_accurate(A.row_ptr, A.col, vals, xs)
    big = np.abs(ya) > 1e-3 * absr
    d_rp, d_col, d_val, d_x = (dev(A.row_ptr, device), dev(A.col, device), dev(vals, device),
                               dev(xs, device))
    plan = MerbitPlan(d_rp, A.ncol, A.nnz)
    for acc in ("compensated", "f64"):
        plan.set_accumulation(acc)
        y = plan.spmv(d_rp, d_col, d_val, d_x).cpu().numpy().astype(np.float64)
        assert cond_err(y, ya, absr) <= MIXED_TOL, acc
        rel = np.max(np.abs(y[big] - ya[big]) / np.abs(ya[big]))
        assert rel <= MIXED_TOL, acc
    plan.close()


@pytest.mark.slow
@pytest.mark.parametrize("kind", ["f64", "f32", "bf16"])
def test_banded_full(orc, device, kind):
    """The banded workload at full size (4M rows, 27 diagonals, 108M entries)
    against the oracle's accurate product on the same stored values, and the
    FP64 result with x = 1 against the row sums (a checksum of checksums:
    sum(y) = sum of all values, accumulated accurately by the oracle)."""
    from merbit_amd import gen, workloads
    from merbit_amd.spmv import MerbitPlan
    say(f"banded 4M rows, 108M entries, {kind}")
    A, x32 = workloads.banded()
    if kind == "f64":
        vals, xs, tol = A.val.astype(np.float64), x32.astype(np.float64), F64_TOL
    else:
        vals = A.val if kind == "f32" else gen.to_bf16(A.val)
        xs, tol = x32, MIXED_TOL
    d_rp = dev(A.row_ptr, device)
    plan = MerbitPlan(d_rp, A.ncol, A.nnz, items_per_thread=12)
    d_col, d_val = dev(A.col, device), dev(vals, device)
    y = plan.spmv(d_rp, d_col, d_val, dev(xs, device)).cpu().numpy().astype(np.float64)
    ya, absr = orc.spmv_accurate(A.row_ptr, A.col, vals, xs)
    assert cond_err(y, ya, absr) <= tol
    if kind == "f64":
        ones = np.ones(A.ncol)
        y1 = plan.spmv(d_rp, d_col, d_val, dev(ones, device)).cpu().numpy()
        rs, rabs = orc.spmv_accurate(A.row_ptr, A.col, vals, ones)
        assert cond_err(y1, rs, rabs) <= F64_TOL
        assert abs(math.fsum(y1) - math.fsum(rs)) <= 1e-12 * math.fsum(rabs)
    plan.close()


@pytest.mark.parametrize("kind", ["f32", "bf16"])
def test_mixed_precision_small(orc, device, kind):
    from merbit_amd import gen
    from merbit_amd.spmv import MerbitPlan
    rng = np.random.default_rng(5)
    for name, lens, ncol in edge_case_patterns(2048) + [
            ("pareto", (rng.pareto(1.1, 30000) * 5).astype(int), 30000)]:
        say(f"{kind} {name}")
        rp, col = random_csr(rng, len(lens), ncol, lens=lens)
        v32 = rng.uniform(-1, 1, int(rp[-1])).astype(np.float32)
        x = rng.uniform(-1, 1, ncol).astype(np.float32)
        stored = v32 if kind == "f32" else gen.to_bf16(v32)
        d_rp = dev(rp, device)
        plan = MerbitPlan(d_rp, ncol, int(rp[-1]))
        y = plan.spmv(d_rp, dev(col, device), dev(stored, device), dev(x, device))
        y = y.cpu().numpy().astype(np.float64)
        plan.close()
        ya, absr = orc.spmv_accurate(rp, col, stored, x)
        # FP32 output: one rounding of the result plus compensated-sum error
        assert cond_err(y, ya, absr) <= MIXED_TOL, name
        big = np.abs(ya) > 1e-3 * absr
        if np.any(big):
            rel = np.max(np.abs(y[big] - ya[big]) / np.abs(ya[big]))
            assert rel <= MIXED_TOL, name


@pytest.mark.parametrize("kind", ["f64", "f32", "bf16"])
def test_unaligned_operands(orc, device, kind):
    """Views one element into larger buffers are not 16-byte aligned: the
    library stages them element by element; results match the oracle."""
    from merbit_amd import gen
    from merbit_amd.spmv import MerbitPlan
    rng = np.random.default_rng(17)
    for name, lens, ncol in edge_case_patterns(2048)[4:8]:
        say(f"unaligned {kind} {name}")
        rp, col = random_csr(rng, len(lens), ncol, lens=lens)
        v = rng.uniform(-1, 1, int(rp[-1]))
        x = rng.uniform(-1, 1, ncol)
        if kind == "f64":
            vals, xs, tol = v, x, F64_TOL
        else:
            v32 = v.astype(np.float32)
            vals = v32 if kind == "f32" else gen.to_bf16(v32)
            xs, tol = x.astype(np.float32), MIXED_TOL

        def shifted(a):
            t = torch.zeros(len(a) + 1, dtype=dev(a[:1], device).dtype, device=device)
            t[1:] = dev(a, device)
            return t[1:]
        d_rp, d_col, d_val = shifted(rp), shifted(col), shifted(vals)
        assert d_col.data_ptr() % 16 != 0
        plan = MerbitPlan(d_rp, ncol, int(rp[-1]))
        y = plan.spmv(d_rp, d_col, d_val, dev(xs, device)).cpu().numpy().astype(np.float64)
        plan.close()
        ya, absr = orc.spmv_accurate(rp, col, vals, xs)
        assert cond_err(y, ya, absr) <= tol, name


def test_pattern_stats(orc, device):
    """mbx_plan_stats against numpy moments of the row lengths and tile
    statistics recomputed from the partition coordinates."""
    from merbit_amd.spmv import MerbitPlan
    rng = np.random.default_rng(44)
    cases = [("pareto", (rng.pareto(1.0, 30000) * 4).astype(int), 30000)]
    cases += [c for c in edge_case_patterns(256 * 4) if c[0] in ("giant_row", "all_empty",
                                                                 "run_long", "many_short")]
    for name, lens, ncol in cases:
        say(f"stats {name}")
        rp, _ = random_csr(rng, len(lens), ncol, lens=lens)
        plan = MerbitPlan(dev(rp, device), ncol, int(rp[-1]), items_per_thread=4)
        st = plan.stats(dev(rp, device))
        ln = np.diff(rp).astype(np.float64)
        n = len(ln)
        assert (st["nrow"], st["nnz"], st["ntiles"]) == (n, int(rp[-1]), plan.ntiles)
        assert st["empty_rows"] == int((ln == 0).sum()) and st["max_row"] == int(ln.max())
        d = ln - ln.mean()
        var = (d ** 2).mean()
        assert abs(st["mean"] - ln.mean()) <= 1e-12 * max(1.0, ln.mean())
        assert abs(st["var"] - var) <= 1e-9 * max(1.0, var)
        if var > 0:
            assert abs(st["skew"] - (d ** 3).mean() / var ** 1.5) <= 1e-9 * max(1.0, abs(st["skew"]))
            assert abs(st["kurt"] - (d ** 4).mean() / var ** 2) <= 1e-9 * max(1.0, st["kurt"])
        rows, nnz = plan.partition()
        T = plan.tile_items
        nt = plan.ntiles
        per = rows[1:nt + 1] - rows[:nt]
        assert st["single_row_tiles"] == int((per == 0).sum())
        assert st["max_rows_per_tile"] == int(per.max())
        split = [t for t in range(1, nt) if per[t] > 0 and rp[rows[t]] < nnz[t]]
        assert st["split_tiles"] == len(split)
        assert st["long_rows"] == sum(1 for t in split if t - (rows[t] + rp[rows[t]]) // T > 64)
        plan.close()


def test_plan_clones_on_concurrent_streams(orc, device):
    """A plan and its clone share the partition and own their carry slots:
    products on two streams at once equal the products run one by one, and
    destroying the original leaves the clone working."""
    from merbit_amd.spmv import MerbitPlan
    say("plan clone, two streams")
    rng = np.random.default_rng(41)
    lens = (rng.pareto(1.0, 40000) * 6).astype(int)
    lens[7] = 20 * 2048  # a row across many tiles: the carry slots are in use
    rp, col = random_csr(rng, len(lens), 30000, lens=lens)
    val = rng.standard_normal(int(rp[-1]))
    xs = [rng.uniform(-1, 1, 30000) for _ in range(2)]
    d_rp, d_col, d_val = dev(rp, device), dev(col, device), dev(val, device)
    d_x = [dev(x, device) for x in xs]
    plan = MerbitPlan(d_rp, 30000, int(rp[-1]))
    twin = plan.clone()
    want = [plan.spmv(d_rp, d_col, d_val, d_x[i]).clone() for i in range(2)]
    streams = [torch.cuda.Stream(), torch.cuda.Stream()]
    ys = [torch.empty(len(lens), dtype=torch.float64, device=device) for _ in range(2)]
    torch.cuda.synchronize()
    for _ in range(20):
        for i, (p, s) in enumerate(((plan, streams[0]), (twin, streams[1]))):
            with torch.cuda.stream(s):
                p.spmv(d_rp, d_col, d_val, d_x[i], ys[i], stream=s)
    torch.cuda.synchronize()
    for i in range(2):
        assert torch.equal(ys[i], want[i]), i
    plan.close()
    y = twin.spmv(d_rp, d_col, d_val, d_x[1])
    assert torch.equal(y, want[1])
    ya, absr = orc.spmv_accurate(rp, col, val, xs[1])
    assert cond_err(y.cpu().numpy(), ya, absr) <= F64_TOL
    twin.close()


@pytest.mark.parametrize("kind", ["f64", "f32", "bf16"])
def test_stream_policies_identical(device, kind):
    """Cached, non-temporal and automatic stream loads change only cache
    hints, and a cap on tiles per CU only the LDS reserved per tile: the
    products are bit-identical, for a plan and for a batch."""
    from merbit_amd import gen
    from merbit_amd.batch import MerbitBatch
    from merbit_amd.spmv import MerbitPlan
    say(f"stream policies {kind}")
    rng = np.random.default_rng(43)
    lens = (rng.pareto(1.0, 30000) * 6).astype(int)
    rp, col = random_csr(rng, len(lens), 20000, lens=lens)
    v32 = rng.uniform(-1, 1, int(rp[-1])).astype(np.float32)
    if kind == "f64":
        val, x = v32.astype(np.float64), rng.uniform(-1, 1, 20000)
    else:
        val = v32 if kind == "f32" else gen.to_bf16(v32)
        x = rng.uniform(-1, 1, 20000).astype(np.float32)
    d_rp, d_col, d_val, d_x = (dev(a, device) for a in (rp, col, val, x))
    plan = MerbitPlan(d_rp, 20000, int(rp[-1]), items_per_thread=12)
    outs = {}
    for mode in ("auto", "cached", "nontemporal"):
        plan.set_stream(mode)
        outs[mode] = plan.spmv(d_rp, d_col, d_val, d_x).clone()
    for wg in (1, 3, 16):  # tiles per CU (2 and 1 need the over-64 KB LDS attribute)
        plan.set_residency(wg)
        outs[f"resid{wg}"] = plan.spmv(d_rp, d_col, d_val, d_x).clone()
    plan.set_residency(0)
    plan.close()
    ydt = outs["auto"].dtype
    mats = [(d_rp, d_col, d_val, d_x, torch.empty(len(lens), dtype=ydt, device=device), 20000)]
    b = MerbitBatch(mats, items_per_thread=12)
    bouts = {}
    for mode in ("cached", "nontemporal"):
        b.set_stream(mode)
        b.spmv()
        bouts[mode] = mats[0][4].clone()
    for wg in (1, 3):
        b.set_residency(wg)
        b.spmv()
        bouts[f"resid{wg}"] = mats[0][4].clone()
    b.close()
    for o in list(outs.values()) + list(bouts.values()):
        assert torch.equal(o.view(torch.uint8), outs["cached"].view(torch.uint8))


@pytest.mark.slow
def test_f64_near_int32_limit(device):
    """nrow + nnz just under 2^31 (2^21 ragged rows, 2.14e9 entries, 26 GB of
    CSR built on the device): with x[c] = c every row sum is an integer below
    2^53, so the FP64 result is exact in any order and must equal the closed
    form."""
    from merbit_amd.spmv import MerbitPlan
    n = 1 << 21
    base = 1022
    say(f"int32 limit: {n} rows, ~{n * base} entries")
    r = torch.arange(n, dtype=torch.int64, device=device)
    lens = base + (r % 5) - 2                      # ragged, same total
    rp = torch.zeros(n + 1, dtype=torch.int64, device=device)
    rp[1:] = torch.cumsum(lens, 0)
    nnz = int(rp[-1])
    assert n + nnz + 16 * 256 < 2**31 - 1
    start = (7 * r) % n                             # first column of each row
    # col[k] = (start[row] + k - rp[row]) mod n, built in chunks of rows
    col = torch.empty(nnz, dtype=torch.int32, device=device)
    step = 1 << 17
    for r0 in range(0, n, step):
        r1 = min(n, r0 + step)
        k0, k1 = int(rp[r0]), int(rp[r1])
        rows = torch.repeat_interleave(torch.arange(r0, r1, device=device), lens[r0:r1])
        k = torch.arange(k0, k1, dtype=torch.int64, device=device)
        col[k0:k1] = ((start[rows] + k - rp[rows]) % n).to(torch.int32)
        del rows, k
    val = torch.ones(nnz, dtype=torch.float64, device=device)
    x = torch.arange(n, dtype=torch.float64, device=device)
    rp32 = rp.to(torch.int32)
    plan = MerbitPlan(rp32, n, nnz, items_per_thread=12)
    y = plan.spmv(rp32, col, val, x)
    plan.close()
    # sum_j ((s + j) mod n), j < L, with at most one wrap (L < n)
    L = lens
    want = L * start + L * (L - 1) // 2 - n * torch.clamp(start + L - n, min=0)
    assert torch.equal(y, want.to(torch.float64))
