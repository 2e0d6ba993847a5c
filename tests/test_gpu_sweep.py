"""Randomised sweep over the product's options in combination (seeded, so a
failure names a reproducible case): pattern shape and row-length
distribution, items per thread, value kind, stream policy, residency cap,
hot-column split, 16-bit columns, tile-sorted gathers, FP64 sums,
alpha/beta and value arrays that are not 16-byte aligned (the scalar
staging path).

Every plain result is checked against the oracle's accurate product (FP64
1e-12, FP32/BF16 1e-6, condition-scaled), the options that keep the
arithmetic (stream policy, residency, hot split, 16-bit columns, sorted
gathers) must give
the plan's plain product bit for bit, and the rest (FP64 sums, alpha/beta,
the scalar path) must stay within the bars."""
from __future__ import annotations

import numpy as np
import pytest
import torch

from helpers import cond_err, random_csr

pytestmark = pytest.mark.gpu

CASES = 160


def say(msg):
    print(f"[case] {msg}", flush=True)


def dev(a, device):
    return torch.from_numpy(np.ascontiguousarray(a)).to(device)


def pattern(rng):
    """(name, row_ptr, col, ncol): one of several row-length families."""
    fam = int(rng.integers(0, 6))
    nrow = int(rng.integers(1, 60000))
    ncol = int(rng.integers(1, 80000))
    if fam == 0:
        lens = rng.poisson(rng.uniform(0.5, 12.0), nrow)
        name = "poisson"
    elif fam == 1:
        lens = np.minimum((rng.pareto(1.1, nrow) * rng.uniform(1, 6)).astype(np.int64), 60000)
        name = "pareto"
    elif fam == 2:  # mostly empty rows
        lens = np.where(rng.random(nrow) < 0.05, rng.integers(1, 40, nrow), 0)
        name = "sparse_rows"
    elif fam == 3:  # a few giant rows among short ones
        lens = rng.integers(0, 4, nrow)
        lens[rng.integers(0, nrow, 3)] = rng.integers(5000, 40000, 3)
        name = "giant_rows"
    elif fam == 4:  # banded: local columns
        width = int(rng.integers(1, 9))
        ncol = nrow
        rows = np.repeat(np.arange(nrow), 2 * width + 1)
        cols = rows + np.tile(np.arange(-width, width + 1), nrow)
        keep = (cols >= 0) & (cols < ncol)
        rp = np.zeros(nrow + 1, np.int32)
        rp[1:] = np.cumsum(np.bincount(rows[keep], minlength=nrow))
        return "banded", rp, cols[keep].astype(np.int32), ncol
    else:  # one column holds a large share (power law in the columns)
        lens = rng.poisson(6.0, nrow)
        name = "hot_column"
    rp, col = random_csr(rng, nrow, ncol, lens=lens)
    if fam == 5 and col.size:
        hot = rng.random(col.size) < 0.3
        col[hot] = int(rng.integers(0, ncol))
        # keep the columns ascending within rows, as the reference's layout
        rows = np.repeat(np.arange(nrow), np.diff(rp))
        col = col[np.lexsort((col, rows))]
    return name, rp, col, ncol


def values(rng, kind, nnz, ncol):
    from merbit_amd import gen
    v = rng.standard_normal(nnz)
    x = rng.uniform(-1, 1, ncol)
    if kind == "f64":
        return v, x
    v32 = v.astype(np.float32)
    return (v32 if kind == "f32" else gen.to_bf16(v32)), x.astype(np.float32)


def test_random_option_sweep(orc, device):
    from merbit_amd.spmv import MerbitPlan
    rng = np.random.default_rng(2026)
    for case in range(CASES):
        kind = ["f64", "f32", "bf16"][int(rng.integers(0, 3))]
        ipt = int(rng.choice([2, 4, 6, 8, 10, 12, 14, 16]))
        name, rp, col, ncol = pattern(rng)
        nrow, nnz = len(rp) - 1, int(rp[-1])
        vals, x = values(rng, kind, nnz, ncol)
        tol = 1e-12 if kind == "f64" else 1e-6
        ydt = torch.float64 if kind == "f64" else torch.float32
        say(f"sweep {case}: {name} nrow {nrow} ncol {ncol} nnz {nnz} {kind} ipt {ipt}")
        d_rp, d_col, d_val, d_x = dev(rp, device), dev(col, device), dev(vals, device), \
            dev(x, device)
        plan = MerbitPlan(d_rp, ncol, nnz, items_per_thread=ipt)
        base = plan.spmv(d_rp, d_col, d_val, d_x)
        ya, absr = orc.spmv_accurate(rp, col, vals, x)
        assert cond_err(base.double().cpu().numpy(), ya, absr) <= tol, case
        # options that keep the arithmetic: bit-identical to the plain product
        mode = ["auto", "cached", "nontemporal"][int(rng.integers(0, 3))]
        wg = int(rng.integers(0, 5))
        plan.set_stream(mode)
        plan.set_residency(wg)
        opts = [mode, f"wg {wg}"]
        if ipt in (8, 12) and rng.random() < 0.6:
            nh = int(rng.choice([1, 64, 1000, 4096, 6900]))
            plan.split_hot(d_col, nh)
            opts.append(f"hot {nh}")
        if ipt in (8, 12) and rng.random() < 0.4 and plan.compress_cols(d_col):
            opts.append("col16")
        # tile-sorted gathers where neither of those took (an own draw, so
        # the cases above stay as they were)
        if ipt in (8, 12, 16) and len(opts) == 2 and \
                np.random.default_rng(7000 + case).random() < 0.7 and plan.sort_cols(d_col):
            opts.append("sorted")
        y = plan.spmv(d_rp, d_col, d_val, d_x)
        assert torch.equal(y.view(torch.uint8), base.view(torch.uint8)), (case, opts)
        # alpha / beta against the host's combination of the oracle's product
        alpha, beta = float(rng.uniform(-2, 2)), float(rng.choice([0.0, 1.0, -0.5]))
        y0 = rng.uniform(-1, 1, nrow).astype(np.float64 if kind == "f64" else np.float32)
        yab = plan.spmv(d_rp, d_col, d_val, d_x, dev(y0, device), alpha, beta)
        want = alpha * ya + beta * y0.astype(np.float64)
        scale = abs(alpha) * absr + abs(beta) * np.abs(y0.astype(np.float64))
        got = yab.double().cpu().numpy()
        bad = np.abs(got - want) > tol * scale + (4 * np.finfo(np.float32).eps
                                                  if kind != "f64" else 0) * np.abs(want)
        assert not np.any(bad & (scale > 0)), (case, opts, alpha, beta)
        assert np.all(got[scale == 0] == 0), (case, opts)
        # FP64 sums for the narrow value kinds, with exact or FP32 products
        r_acc = rng.random()
        if kind != "f64" and r_acc < 0.5:
            acc = "f64" if r_acc < 0.25 else "f32prod"
            plan.set_accumulation(acc)
            yf = plan.spmv(d_rp, d_col, d_val, d_x)
            assert cond_err(yf.double().cpu().numpy(), ya, absr) <= tol, (case, acc)
            plan.set_accumulation("compensated")
        # values one element into a larger buffer: not 16-byte aligned, so the
        # scalar staging path (the hot split and 16-bit columns need the
        # vector path and step aside)
        if nnz > 0 and rng.random() < 0.3:
            holder = torch.zeros(nnz + 1, dtype=d_val.dtype, device=device)
            holder[1:] = d_val
            ys = plan.spmv(d_rp, d_col, holder[1:], d_x)
            assert cond_err(ys.double().cpu().numpy(), ya, absr) <= tol, (case, "scalar path")
        assert y.dtype == ydt
        plan.close()


def test_random_row_tiles_sweep(orc, device):
    """Row-aligned plans (mbx_plan_create_rows) over random short-row
    patterns (every row at most 1/16 of a tile): against the oracle, the
    options that keep the arithmetic bit-identical to the row plan's plain
    product, the scalar staging path within the bar."""
    from merbit_amd.spmv import MerbitPlan
    rng = np.random.default_rng(77)
    for case in range(80):
        kind = ["f64", "f32", "bf16"][int(rng.integers(0, 3))]
        ipt = int(rng.choice([2, 4, 6, 8, 10, 12, 14, 16]))
        L = 16 * ipt  # longest row allowed: T / 16
        nrow = int(rng.integers(1, 50000))
        ncol = int(rng.integers(1, 60000))
        fam = int(rng.integers(0, 3))
        if fam == 0:
            lens = np.minimum(rng.poisson(rng.uniform(0.5, 20.0), nrow), L)
        elif fam == 1:
            lens = np.where(rng.random(nrow) < 0.1, rng.integers(1, L + 1, nrow), 0)
        else:
            lens = np.full(nrow, int(rng.integers(0, L + 1)))
        rp, col = random_csr(rng, nrow, ncol, lens=lens)
        nnz = int(rp[-1])
        vals, x = values(rng, kind, nnz, ncol)
        tol = 1e-12 if kind == "f64" else 1e-6
        say(f"row sweep {case}: fam {fam} nrow {nrow} ncol {ncol} nnz {nnz} {kind} ipt {ipt}")
        d_rp, d_col, d_val, d_x = dev(rp, device), dev(col, device), dev(vals, device), \
            dev(x, device)
        plan = MerbitPlan(d_rp, ncol, nnz, items_per_thread=ipt, row_tiles=True)
        base = plan.spmv(d_rp, d_col, d_val, d_x)
        ya, absr = orc.spmv_accurate(rp, col, vals, x)
        assert cond_err(base.double().cpu().numpy(), ya, absr) <= tol, case
        mode = ["auto", "cached", "nontemporal"][int(rng.integers(0, 3))]
        wg = int(rng.integers(0, 5))
        plan.set_stream(mode)
        plan.set_residency(wg)
        opts = [mode, f"wg {wg}"]
        if ipt in (8, 12) and rng.random() < 0.5:
            nh = int(rng.choice([1, 64, 1000, 4096]))
            plan.split_hot(d_col, nh)
            opts.append(f"hot {nh}")
        if ipt in (8, 12) and rng.random() < 0.5 and plan.compress_cols(d_col):
            opts.append("col16")
        if ipt in (8, 12, 16) and len(opts) == 2 and \
                np.random.default_rng(9000 + case).random() < 0.7 and plan.sort_cols(d_col):
            opts.append("sorted")
        y = plan.spmv(d_rp, d_col, d_val, d_x)
        assert torch.equal(y.view(torch.uint8), base.view(torch.uint8)), (case, opts)
        if nnz > 0 and rng.random() < 0.3:
            holder = torch.zeros(nnz + 1, dtype=d_val.dtype, device=device)
            holder[1:] = d_val
            ys = plan.spmv(d_rp, d_col, holder[1:], d_x)
            assert cond_err(ys.double().cpu().numpy(), ya, absr) <= tol, (case, "scalar path")
        assert plan.stats(d_rp)["split_tiles"] == 0, case
        plan.close()


def test_random_spmm_sweep(orc, device):
    """SpMM (FP32) over random patterns, k from 1 to 33, row- and
    column-major X and Y, 2 or 4 items per thread and alpha/beta: every
    column within 1e-5 (condition-scaled) of the oracle's accurate product."""
    from merbit_amd.spmv import MerbitPlan
    rng = np.random.default_rng(2027)
    for case in range(60):
        name, rp, col, ncol = pattern(rng)
        nrow, nnz = len(rp) - 1, int(rp[-1])
        k = int(rng.choice([1, 2, 3, 4, 5, 7, 8, 13, 16, 32, 33]))
        ipt = int(rng.choice([2, 4]))
        say(f"spmm sweep {case}: {name} nrow {nrow} ncol {ncol} nnz {nnz} k {k} ipt {ipt}")
        vals = rng.standard_normal(nnz).astype(np.float32)
        X = rng.uniform(-1, 1, (ncol, k)).astype(np.float32)
        d_rp, d_col, d_val = dev(rp, device), dev(col, device), dev(vals, device)
        dX = dev(X, device)
        if rng.random() < 0.5:  # column-major X
            dX = dX.t().contiguous().t()
        y0 = rng.uniform(-1, 1, (nrow, k)).astype(np.float32)
        dY = dev(y0, device)
        if rng.random() < 0.5:  # column-major Y
            dY = dY.t().contiguous().t()
        alpha, beta = float(rng.uniform(-2, 2)), float(rng.choice([0.0, 1.0, -0.5]))
        plan = MerbitPlan(d_rp, ncol, nnz, items_per_thread=ipt)
        Y = plan.spmm(d_rp, d_col, d_val, dX, dY, alpha, beta).cpu().numpy().astype(np.float64)
        for j in range(k):
            ya, absr = orc.spmv_accurate(rp, col, vals, np.ascontiguousarray(X[:, j]))
            want = alpha * ya + beta * y0[:, j].astype(np.float64)
            scale = abs(alpha) * absr + abs(beta) * np.abs(y0[:, j].astype(np.float64))
            bad = np.abs(Y[:, j] - want) > 1e-5 * scale
            assert not np.any(bad & (scale > 0)), (case, j)
            assert np.all(Y[scale == 0, j] == 0), (case, j)
        plan.close()


def test_random_batch_sweep(orc, device):
    """Uneven batches (1 to 40 matrices of the families above, cut to at most
    6000 rows, one value kind per batch) with random items per thread,
    stream policy, residency and alpha/beta: every member within its bar of
    the oracle."""
    from merbit_amd.batch import MerbitBatch
    rng = np.random.default_rng(2028)
    for case in range(30):
        kind = ["f64", "f32", "bf16"][int(rng.integers(0, 3))]
        ipt = int(rng.choice([2, 4, 8, 12, 16]))
        count = int(rng.integers(1, 41))
        tol = 1e-12 if kind == "f64" else 1e-6
        ydt = np.float64 if kind == "f64" else np.float32
        alpha, beta = float(rng.uniform(-2, 2)), float(rng.choice([0.0, 1.0]))
        say(f"batch sweep {case}: {count} matrices {kind} ipt {ipt}")
        mats, host = [], []
        for _ in range(count):
            name, rp, col, ncol = pattern(rng)
            r = min(len(rp) - 1, int(rng.integers(1, 6001)))
            rp, nnz = rp[:r + 1].copy(), int(rp[r])
            col = col[:nnz].copy()
            vals, x = values(rng, kind, nnz, ncol)
            y0 = rng.uniform(-1, 1, r).astype(ydt)
            d = (dev(rp, device), dev(col, device), dev(vals, device), dev(x, device),
                 dev(y0, device), ncol)
            mats.append(d)
            host.append((rp, col, vals, x, y0))
        b = MerbitBatch(mats, items_per_thread=ipt)
        b.set_stream(["auto", "cached", "nontemporal"][int(rng.integers(0, 3))])
        b.set_residency(int(rng.integers(0, 5)))
        if ipt in (8, 12, 16) and np.random.default_rng(8000 + case).random() < 0.6:
            b.sort_cols(True)  # tile-sorted gathers (an own draw: the cases stay)
        b.spmv(alpha, beta)
        torch.cuda.synchronize()
        for i, ((rp, col, vals, x, y0), m) in enumerate(zip(host, mats)):
            ya, absr = orc.spmv_accurate(rp, col, vals, x)
            want = alpha * ya + beta * y0.astype(np.float64)
            scale = abs(alpha) * absr + abs(beta) * np.abs(y0.astype(np.float64))
            got = m[4].double().cpu().numpy()
            bad = np.abs(got - want) > tol * scale
            assert not np.any(bad & (scale > 0)), (case, i)
            assert np.all(got[scale == 0] == 0), (case, i)
        b.close()
