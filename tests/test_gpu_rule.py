"""The launch rule (mbx_launch_rule / mbx_plan_create_auto): which rule each
workload class meets, that the plan it makes gives the same results as a
plain plan within the bars, and that its launch is within 5 % of the
timing autotuner's best (mbx_autotune_spmv_form, which sweeps items per
thread x stream policy x tiles per CU x tile-sorted gathers) on the
benchmark workloads.  The
autotuner's pick is re-timed with the same method as the rule's plan (the
minimum of its 64 single timings is biased low), both as graph replays."""
from __future__ import annotations

import ctypes

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[case] {msg}", flush=True)


def dev(a, device):
    return torch.from_numpy(np.ascontiguousarray(a)).to(device)


def time_us(fn, reps, rounds=5):
    """Device time per call: `reps` calls captured as one HIP graph, the
    best of `rounds` replays (host launch overhead left out, which would
    dominate the microsecond products)."""
    fn()
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        for _ in range(reps):
            fn()
    torch.cuda.synchronize()
    best = float("inf")
    for _ in range(rounds):
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        g.replay()
        b.record()
        torch.cuda.synchronize()
        best = min(best, a.elapsed_time(b) * 1e3 / reps)
    del g
    return best


def operands(name):
    from merbit_amd import gen, workloads
    if name.startswith("rmat"):
        # rmat18_*: 2^18 rows and 2^22 entries, a power law in the few-tiles class
        A, x = workloads.rmat(18) if name.startswith("rmat18") else workloads.rmat()
        if name in ("rmat_f64", "rmat18_f64"):
            return A, A.val, x
        return A, A.val.astype(np.float32), x.astype(np.float32)
    if name.startswith("banded"):
        B, xb = workloads.banded()
        if name == "banded_f64":
            return B, B.val.astype(np.float64), xb.astype(np.float64)
        if name == "banded_bf16":
            return B, gen.to_bf16(B.val), xb
        return B, B.val, xb
    if name == "graph_f32":
        G, xg, _ = workloads.graph()
        return G, G.val, xg
    if name == "laplacian_f64":
        A, x = workloads.laplacian(256)
        return A, A.val, x
    mats, xs = workloads.uneven_batch()
    i = int(name.split("_")[1])
    return mats[i], mats[i].val, xs[i]


EXPECTED = {"rmat_f64": "hot_columns", "rmat_f32": "hot_columns", "banded_f64": "local",
            "banded_f32": "local", "banded_bf16": "local", "graph_f32": "spread_columns",
            "laplacian_f64": "few_tiles", "batch_21": "few_tiles", "batch_15": "few_tiles",
            "batch_38": "x_fits", "rmat18_f64": "few_tiles"}


@pytest.mark.slow
@pytest.mark.parametrize("name", list(EXPECTED))
def test_rule_within_5pct_of_timed_best(device, name):
    from merbit_amd._lib import lib
    from merbit_amd.spmv import MerbitPlan
    A, vals, x = operands(name)
    kind = {np.dtype(np.float64): 0, np.dtype(np.float32): 1, np.dtype(np.uint16): 2}[vals.dtype]
    d_rp, d_col, d_val, d_x = dev(A.row_ptr, device), dev(A.col, device), dev(vals, device), \
        dev(x, device)
    vdt = {0: torch.float64, 1: torch.float32, 2: torch.uint16}[kind]
    y = torch.empty(A.nrow, dtype=torch.float64 if kind == 0 else torch.float32, device=device)
    reps = max(5, min(200, int(4e7 / max(A.nnz, 1))))
    s = torch.cuda.current_stream().cuda_stream
    ipt, mode, wg, srt = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
    us = ctypes.c_float()
    assert lib().mbx_autotune_spmv_form(d_rp.data_ptr(), A.nrow, A.ncol, A.nnz,
                                        d_col.data_ptr(), d_val.data_ptr(), kind,
                                        d_x.data_ptr(), y.data_ptr(), reps, s,
                                        ctypes.byref(ipt), ctypes.byref(mode), ctypes.byref(wg),
                                        ctypes.byref(srt), ctypes.byref(us)) == 0
    tuned = MerbitPlan(d_rp, A.ncol, A.nnz, items_per_thread=ipt.value)
    tuned.set_stream({1: "cached", 2: "nontemporal"}[mode.value])
    tuned.set_residency(wg.value)
    if srt.value:
        assert tuned.sort_cols(d_col)
    rule = MerbitPlan(d_rp, A.ncol, A.nnz, col=d_col, value_dtype=vdt)
    c = rule.choice
    assert c["rule"] == EXPECTED[name], c
    ref = tuned.spmv(d_rp, d_col, d_val, d_x).double().cpu().numpy()
    got = rule.spmv(d_rp, d_col, d_val, d_x).double().cpu().numpy()
    tol = 1e-12 if kind == 0 else 1e-6
    assert np.linalg.norm(got - ref) <= tol * np.linalg.norm(ref)
    t_tuned = time_us(lambda: tuned.spmv(d_rp, d_col, d_val, d_x, y), reps)
    t_rule = time_us(lambda: rule.spmv(d_rp, d_col, d_val, d_x, y), reps)
    say(f"{name}: rule {c['rule']} ipt {c['items_per_thread']} wg {c['wg_per_cu']} "
        f"hot {c['hot_columns']} sort {c['sort_cols']} col16 {c['compress_cols']} "
        f"acc {c['accumulation']}: "
        f"{t_rule:.2f} us; tuned ipt {ipt.value} mode {mode.value} wg {wg.value} "
        f"sort {srt.value}: "
        f"{t_tuned:.2f} us (autotuner {us.value:.2f}); ratio {t_rule / t_tuned:.3f}")
    assert t_rule <= 1.05 * t_tuned, (t_rule, t_tuned)
    tuned.close()
    rule.close()


def test_rule_features_small(device):
    """Rule features on small patterns: a banded pattern is local, a
    pattern with one dense column has that column's share, nothing wider
    than the tile span is reported, and an empty pattern is valid."""
    from merbit_amd.spmv import launch_rule
    n = 50000
    lens = np.full(n, 3)
    rp = np.zeros(n + 1, np.int32)
    rp[1:] = np.cumsum(lens)
    col = np.stack([np.maximum(np.arange(n) - 1, 0), np.arange(n),
                    np.minimum(np.arange(n) + 1, n - 1)], 1).reshape(-1).astype(np.int32)
    c = launch_rule(dev(rp, device), dev(col, device), n)
    assert c["rule"] == "few_tiles" and c["max_tile_span"] < 2048 and c["ntiles8"] == 98
    # every row also references column 0: one column holds a quarter
    col2 = np.stack([np.zeros(n, np.int32), col[0::3], col[1::3], col[2::3]], 1).reshape(-1)
    rp2 = np.arange(0, 4 * n + 1, 4, dtype=np.int32)
    c2 = launch_rule(dev(rp2, device), dev(col2.astype(np.int32), device), n)
    cnt = np.bincount(col2, minlength=n)
    want = np.sort(cnt)[::-1][:4096].sum() / col2.size
    assert c2["hot_share"] == pytest.approx(want, abs=1e-12)
    assert c2["max_tile_span"] == n - 1
    c3 = launch_rule(dev(np.zeros(11, np.int32), device),
                     torch.zeros(0, dtype=torch.int32, device=device), 7)
    assert c3["rule"] == "few_tiles" and c3["hot_share"] == 0.0
    # fewer tiles at 4 items per thread than CUs: 2 items per thread
    for ch in (c, c3):
        assert ch["items_per_thread"] == (2 if 2 * ch["ntiles8"] < ch["cus"] else 4)
    assert c3["items_per_thread"] == 2
