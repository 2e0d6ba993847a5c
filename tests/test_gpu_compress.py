"""GPU parity of compressed column indices (mbx_plan_compress_cols).

Only the column encoding changes, so products on a compressed plan must be
bit-identical to products on the same plan without compression, for every
precision, and match the oracle with the plain product's bars.  Patterns
whose tiles span more than 16 bits of columns are refused and keep working
uncompressed."""
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


def values(kind, rng, n):
    from merbit_amd import gen
    v = rng.standard_normal(n)
    if kind == "f64":
        return v
    v32 = v.astype(np.float32)
    return v32 if kind == "f32" else gen.to_bf16(v32)


@pytest.mark.parametrize("ipt", [8, 12])
@pytest.mark.parametrize("kind", ["f64", "f32", "bf16"])
def test_compressed_bit_identical(orc, device, kind, ipt):
    from merbit_amd.spmv import MerbitPlan
    rng = np.random.default_rng(40 + ipt)
    cases = edge_case_patterns(256 * ipt) + [
        ("banded", [27] * 20000, 20000), ("random", rng.poisson(8, 9000), 60000)]
    for name, lens, ncol in cases:
        say(f"compress {kind} ipt={ipt} {name}")
        if name == "banded":
            n = len(lens)
            rows = np.repeat(np.arange(n), 27)
            cols = np.clip(rows + np.tile(np.arange(-13, 14), n), 0, n - 1)
            rp = np.arange(0, 27 * n + 1, 27, dtype=np.int32)
            col = cols.astype(np.int32)
        else:
            rp, col = random_csr(rng, len(lens), ncol, lens=lens)
        vals = values(kind, rng, int(rp[-1]))
        xs = rng.uniform(-1, 1, ncol)
        if kind != "f64":
            xs = xs.astype(np.float32)
        d_rp, d_col, d_val, d_x = dev(rp, device), dev(col, device), dev(vals, device), dev(xs, device)
        plan = MerbitPlan(d_rp, ncol, int(rp[-1]), items_per_thread=ipt)
        y0 = plan.spmv(d_rp, d_col, d_val, d_x).cpu().numpy()
        assert plan.compress_cols(d_col), name
        y1 = plan.spmv(d_rp, d_col, d_val, d_x).cpu().numpy()
        assert np.array_equal(y0.view(np.uint8), y1.view(np.uint8)), name
        ya, absr = orc.spmv_accurate(rp, col, vals, xs)
        assert cond_err(y1, ya, absr) <= (1e-12 if kind == "f64" else 1e-6), name
        # a clone reads the caller's columns; dropping returns to them too
        c = plan.clone()
        assert np.array_equal(c.spmv(d_rp, d_col, d_val, d_x).cpu().numpy().view(np.uint8),
                              y0.view(np.uint8)), name
        c.close()
        assert not plan.compress_cols(None)
        y2 = plan.spmv(d_rp, d_col, d_val, d_x).cpu().numpy()
        assert np.array_equal(y2.view(np.uint8), y0.view(np.uint8)), name
        plan.close()


def test_compress_refuses_wide_tiles_and_other_shapes(device):
    from merbit_amd.spmv import MerbitPlan
    say("compress refused")
    rng = np.random.default_rng(3)
    rp, col = random_csr(rng, 5000, 500000, density=10)
    val = rng.standard_normal(int(rp[-1]))
    x = rng.uniform(-1, 1, 500000)
    d_rp, d_col, d_val, d_x = dev(rp, device), dev(col, device), dev(val, device), dev(x, device)
    plan = MerbitPlan(d_rp, 500000, int(rp[-1]), items_per_thread=8)
    y0 = plan.spmv(d_rp, d_col, d_val, d_x).cpu().numpy()
    assert not plan.compress_cols(d_col)          # tiles span ~500k columns
    assert np.array_equal(plan.spmv(d_rp, d_col, d_val, d_x).cpu().numpy(), y0)
    plan.close()
    plan = MerbitPlan(d_rp, 500000, int(rp[-1]), items_per_thread=4)
    assert not plan.compress_cols(d_col)          # no 16-bit kernel at 4 items
    plan.close()


@pytest.mark.slow
def test_compressed_banded_full(device):
    """The benchmark's banded matrix (4M rows, 27 diagonals), FP32 values."""
    from merbit_amd import workloads
    from merbit_amd.spmv import MerbitPlan
    say("compress banded full")
    A, x = workloads.banded()
    d_rp, d_col = dev(A.row_ptr, device), dev(A.col, device)
    d_val, d_x = dev(A.val, device), dev(x, device)
    plan = MerbitPlan(d_rp, A.ncol, A.nnz, items_per_thread=8)
    y0 = plan.spmv(d_rp, d_col, d_val, d_x)
    assert plan.compress_cols(d_col)
    for mode in ("cached", "nontemporal"):
        plan.set_stream(mode)
        assert torch.equal(plan.spmv(d_rp, d_col, d_val, d_x), y0)
    plan.close()


@pytest.mark.parametrize("ipt", [4, 8, 12])
@pytest.mark.parametrize("kind", ["f32", "bf16"])
def test_f64_accumulation(orc, device, kind, ipt):
    """MBX_ACC_F64: exact FP64 products and FP64 sums, y rounded once to
    FP32, so the condition-scaled error is that one rounding (<= 2^-24 of
    |y| <= |A||x|); bar 1e-7.  Compressed columns change nothing."""
    from merbit_amd.spmv import MerbitPlan
    rng = np.random.default_rng(80 + ipt)
    cases = edge_case_patterns(256 * ipt) + [("random", rng.poisson(8, 9000), 60000)]
    for name, lens, ncol in cases:
        say(f"acc f64 {kind} ipt={ipt} {name}")
        rp, col = random_csr(rng, len(lens), ncol, lens=lens)
        vals = values(kind, rng, int(rp[-1]))
        xs = rng.uniform(-1, 1, ncol).astype(np.float32)
        d_rp, d_col, d_val, d_x = dev(rp, device), dev(col, device), dev(vals, device), dev(xs, device)
        plan = MerbitPlan(d_rp, ncol, int(rp[-1]), items_per_thread=ipt)
        plan.set_accumulation("f64")
        y = plan.spmv(d_rp, d_col, d_val, d_x).cpu().numpy()
        ya, absr = orc.spmv_accurate(rp, col, vals, xs)
        assert cond_err(y, ya, absr) <= 1e-7, name
        # alpha, beta
        y0 = rng.uniform(-1, 1, len(lens)).astype(np.float32)
        yb = plan.spmv(d_rp, d_col, d_val, d_x, dev(y0, device), alpha=-2.0, beta=0.5)
        want = -2.0 * ya + 0.5 * y0.astype(np.float64)
        den = 2.0 * absr + 0.5 * np.abs(y0) + 1e-300
        assert np.max(np.abs(yb.cpu().numpy() - want) / den, initial=0.0) <= 1e-7, name
        if ipt in (8, 12) and plan.compress_cols(d_col):
            y1 = plan.spmv(d_rp, d_col, d_val, d_x).cpu().numpy()
            assert np.array_equal(y1.view(np.uint8), y.view(np.uint8)), name
        plan.set_accumulation("compensated")
        yc = plan.spmv(d_rp, d_col, d_val, d_x).cpu().numpy()
        assert cond_err(yc, ya, absr) <= 1e-6, name
        plan.close()


@pytest.mark.parametrize("ipt", [4, 8, 12])
@pytest.mark.parametrize("kind", ["f32", "bf16"])
def test_f32_products_accumulation(orc, device, kind, ipt):
    """MBX_ACC_F32_PRODUCTS: each product rounded once to FP32, FP64 sums,
    a row's partial rounded to FP32 where it is stored in LDS and again where
    a carry completes it: at most 3 * 2^-24 (1.8e-7) of |A||x| in all; bar
    2.5e-7.  Compressed columns and row-aligned tiles change nothing beyond
    that; the launch rule's local plans use this mode."""
    from merbit_amd.spmv import MerbitPlan
    rng = np.random.default_rng(180 + ipt)
    cases = edge_case_patterns(256 * ipt) + [("random", rng.poisson(8, 9000), 60000)]
    for name, lens, ncol in cases:
        say(f"acc f32prod {kind} ipt={ipt} {name}")
        rp, col = random_csr(rng, len(lens), ncol, lens=lens)
        vals = values(kind, rng, int(rp[-1]))
        xs = rng.uniform(-1, 1, ncol).astype(np.float32)
        d_rp, d_col, d_val, d_x = dev(rp, device), dev(col, device), dev(vals, device), dev(xs, device)
        ya, absr = orc.spmv_accurate(rp, col, vals, xs)
        plan = MerbitPlan(d_rp, ncol, int(rp[-1]), items_per_thread=ipt)
        plan.set_accumulation("f32prod")
        y = plan.spmv(d_rp, d_col, d_val, d_x).cpu().numpy()
        assert cond_err(y, ya, absr) <= 2.5e-7, name
        y0 = rng.uniform(-1, 1, len(lens)).astype(np.float32)
        yb = plan.spmv(d_rp, d_col, d_val, d_x, dev(y0, device), alpha=-2.0, beta=0.5)
        want = -2.0 * ya + 0.5 * y0.astype(np.float64)
        den = 2.0 * absr + 0.5 * np.abs(y0) + 1e-300
        assert np.max(np.abs(yb.cpu().numpy() - want) / den, initial=0.0) <= 2.5e-7, name
        if ipt in (8, 12) and plan.compress_cols(d_col):
            y1 = plan.spmv(d_rp, d_col, d_val, d_x).cpu().numpy()
            assert np.array_equal(y1.view(np.uint8), y.view(np.uint8)), name
        plan.close()
        if max(lens) * 16 <= 256 * ipt:  # row-aligned tiles
            rplan = MerbitPlan(d_rp, ncol, int(rp[-1]), items_per_thread=ipt, row_tiles=True)
            rplan.set_accumulation("f32prod")
            yr = rplan.spmv(d_rp, d_col, d_val, d_x).cpu().numpy()
            assert cond_err(yr, ya, absr) <= 2.5e-7, name
            rplan.close()
