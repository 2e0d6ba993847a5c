"""GPU parity of the column-blocked product (mbx_colblock_*).

* The block-major layout is checked bit for bit: values moved by
  mbx_colblock_values must equal a numpy stable sort of the CSR entries by
  (column block, row, position), with block starts padded to 4 slots, 64
  spare slots after the last block and zeros in the padding.
* The chained product is checked against the oracle's accurate CSR product
  with the bars of the plain product: condition-scaled error <= 1e-12 for
  FP64 and <= 1e-6 for FP32/BF16 values (rows are split over the blocks, so
  the rounding differs from mbx_spmv_* but stays within them)."""
from __future__ import annotations

import numpy as np
import pytest
import torch

from helpers import cond_err, edge_case_patterns, random_csr

pytestmark = pytest.mark.gpu

F64_TOL = 1e-12
MIXED_TOL = 1e-6


def say(msg):
    print(f"[case] {msg}", flush=True)


def dev(a, device):
    return torch.from_numpy(np.ascontiguousarray(a)).to(device)


def expected_blocked(rp, col, ncol, nb, val):
    """Block-major values with padded block starts (numpy restatement)."""
    nrow = len(rp) - 1
    rows = np.repeat(np.arange(nrow), np.diff(rp))
    blk = (col.astype(np.int64) * nb) // max(ncol, 1)
    order = np.lexsort((np.arange(len(col)), rows, blk))
    counts = np.bincount(blk, minlength=nb) if len(col) else np.zeros(nb, np.int64)
    base = np.concatenate([[0], np.cumsum((counts + 3) // 4 * 4)])
    out = np.zeros(int(base[-1]) + 64, val.dtype)  # 64 spare slots after the last block
    start = 0
    for j in range(nb):
        out[base[j]:base[j] + counts[j]] = val[order[start:start + counts[j]]]
        start += counts[j]
    return out


def cases(rng):
    out = [(n, lens, ncol) for n, lens, ncol in edge_case_patterns(2048)]
    out.append(("random_graph", rng.poisson(25, 30000), 30000))
    out.append(("narrow", rng.poisson(4, 3000), 3))
    return out


@pytest.mark.parametrize("nb", [0, 1, 3, 16])
def test_colblock_layout_and_f64(orc, device, nb):
    from merbit_amd.spmv import ColumnBlockedCsr
    rng = np.random.default_rng(500 + nb)
    for name, lens, ncol in cases(rng):
        say(f"colblock f64 {name} nb={nb}")
        rp, col = random_csr(rng, len(lens), ncol, lens=lens)
        if name == "random_graph":  # columns need not be sorted within rows
            for r in range(0, len(lens), 7):
                rng.shuffle(col[rp[r]:rp[r + 1]])
        val = rng.standard_normal(int(rp[-1]))
        x = rng.uniform(-1, 1, ncol)
        cb = ColumnBlockedCsr(dev(rp, device), dev(col, device), ncol, nblocks=nb)
        assert 1 <= cb.nblocks <= 16 and cb.nblocks <= max(ncol, 1)
        if nb:
            assert cb.nblocks == min(nb, ncol)
        vb = cb.values(dev(val, device))
        want = expected_blocked(rp, col, ncol, cb.nblocks, val)
        assert cb.value_slots == len(want)
        assert np.array_equal(vb[:cb.value_slots].cpu().numpy().view(np.uint64),
                              want.view(np.uint64)), name
        y = cb.spmv(vb, dev(x, device)).cpu().numpy()
        ya, absr = orc.spmv_accurate(rp, col, val, x)
        assert cond_err(y, ya, absr) <= F64_TOL, name
        # alpha, beta; y read once
        y0 = rng.uniform(-1, 1, len(lens))
        yb = cb.spmv(vb, dev(x, device), dev(y0, device), alpha=-2.0, beta=0.5).cpu().numpy()
        den = 2.0 * absr + 0.5 * np.abs(y0) + 1e-300
        assert np.max(np.abs(yb - (-2.0 * ya + 0.5 * y0)) / den, initial=0.0) <= F64_TOL, name
        cb.close()


@pytest.mark.parametrize("kind", ["f32", "bf16"])
def test_colblock_mixed(orc, device, kind):
    from merbit_amd import gen
    from merbit_amd.spmv import ColumnBlockedCsr
    rng = np.random.default_rng(77)
    for name, lens, ncol in cases(rng):
        say(f"colblock {kind} {name}")
        rp, col = random_csr(rng, len(lens), ncol, lens=lens)
        v32 = rng.standard_normal(int(rp[-1])).astype(np.float32)
        vals = v32 if kind == "f32" else gen.to_bf16(v32)
        x = rng.uniform(-1, 1, ncol).astype(np.float32)
        cb = ColumnBlockedCsr(dev(rp, device), dev(col, device), ncol, nblocks=4)
        vb = cb.values(dev(vals, device))
        want = expected_blocked(rp, col, ncol, cb.nblocks, vals)
        assert np.array_equal(vb[:cb.value_slots].cpu().numpy(), want), name
        y = cb.spmv(vb, dev(x, device)).cpu().numpy()
        ya, absr = orc.spmv_accurate(rp, col, vals, x)
        assert cond_err(y, ya, absr) <= MIXED_TOL, name
        cb.close()


def test_colblock_matches_csr_product_on_a_graph(device):
    """On a random graph the blocked product agrees with the plain CSR
    product to FP64 rounding, and repeated calls are bit-identical under
    every stream policy and residency."""
    from merbit_amd import workloads
    from merbit_amd.spmv import ColumnBlockedCsr, MerbitPlan
    say("colblock graph 200k")
    A, x, _ = workloads.graph(200_000)
    d_rp, d_col = dev(A.row_ptr, device), dev(A.col, device)
    val = dev(A.val.astype(np.float64), device)
    xx = dev(x.astype(np.float64), device)
    plan = MerbitPlan(d_rp, A.ncol, A.nnz)
    y_csr = plan.spmv(d_rp, d_col, val, xx).cpu().numpy()
    cb = ColumnBlockedCsr(d_rp, d_col, A.ncol, nblocks=3)
    vb = cb.values(val)
    outs = [cb.spmv(vb, xx).cpu().numpy() for _ in range(3)]
    # the cache policy and tiles per CU change no arithmetic
    for mode, wg in (("nontemporal", 4), ("cached", 3), ("auto", 0)):
        cb.set_stream(mode)
        cb.set_residency(wg)
        outs.append(cb.spmv(vb, xx).cpu().numpy())
    for o in outs[1:]:
        assert np.array_equal(o, outs[0])
    absr = np.abs(A.val.astype(np.float64))
    s = np.zeros(A.nrow)
    np.add.at(s, np.repeat(np.arange(A.nrow), np.diff(A.row_ptr)),
              absr * np.abs(x.astype(np.float64))[A.col])
    assert cond_err(outs[0], y_csr, s) <= F64_TOL
    cb.close()
    plan.close()


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_colblock_sddmm(device, dtype):
    """dval = g[row]*x[col] in block-major order: one rounded product per
    entry, so it must equal numpy's bit for bit; padding slots untouched."""
    from merbit_amd.spmv import ColumnBlockedCsr
    rng = np.random.default_rng(91)
    for name, lens, ncol in cases(rng):
        say(f"colblock sddmm {np.dtype(dtype).name} {name}")
        rp, col = random_csr(rng, len(lens), ncol, lens=lens)
        g = rng.standard_normal(len(lens)).astype(dtype)
        x = rng.standard_normal(ncol).astype(dtype)
        cb = ColumnBlockedCsr(dev(rp, device), dev(col, device), ncol, nblocks=5)
        rows = np.repeat(np.arange(len(lens)), np.diff(rp))
        want = expected_blocked(rp, col, ncol, cb.nblocks, (g[rows] * x[col]).astype(dtype))
        got = cb.sddmm(dev(g, device), dev(x, device))[:cb.value_slots].cpu().numpy()
        assert np.array_equal(got, want), name
        cb.close()


def transpose_csr(rp, col, ncol, val):
    nrow = len(rp) - 1
    rows = np.repeat(np.arange(nrow), np.diff(rp))
    order = np.lexsort((rows, col))  # stable by column, rows ascending
    rpT = np.zeros(ncol + 1, np.int32)
    rpT[1:] = np.cumsum(np.bincount(col, minlength=ncol))
    return rpT, rows[order].astype(np.int32), val[order]


@pytest.mark.parametrize("kind", ["f64", "f32", "bf16"])
@pytest.mark.parametrize("nb,nbt", [(1, 1), (3, 4), (16, 0)])
def test_colblock_adjoint_dx(orc, device, kind, nb, nbt):
    """dx = A^T g from block-major values against the oracle's accurate
    product on the transposed CSR (bars: 1e-12 FP64, 1e-6 FP32/BF16)."""
    from merbit_amd import gen
    from merbit_amd.spmv import ColumnBlockedCsr
    rng = np.random.default_rng(700 + nb + 10 * nbt)
    for name, lens, ncol in cases(rng):
        say(f"colblock dx {kind} {name} nb={nb} nbt={nbt}")
        rp, col = random_csr(rng, len(lens), ncol, lens=lens)
        v = rng.standard_normal(int(rp[-1]))
        if kind == "f64":
            vals, g = v, rng.uniform(-1, 1, len(lens))
        else:
            v32 = v.astype(np.float32)
            vals = v32 if kind == "f32" else gen.to_bf16(v32)
            g = rng.uniform(-1, 1, len(lens)).astype(np.float32)
        d_rp, d_col = dev(rp, device), dev(col, device)
        cb = ColumnBlockedCsr(d_rp, d_col, ncol, nblocks=nb)
        nbt_got = cb.prepare_adjoint(d_rp, d_col, nblocks=nbt)
        assert 1 <= nbt_got <= 16
        vb = cb.values(dev(vals, device))
        dx = cb.adjoint_x(vb, dev(g, device)).cpu().numpy()
        rpT, colT, valT = transpose_csr(rp, col, ncol, vals)
        want, absr = orc.spmv_accurate(rpT, colT, valT, g)
        assert cond_err(dx, want, absr) <= (F64_TOL if kind == "f64" else MIXED_TOL), name
        # alpha, beta
        if kind == "f64":
            d0 = rng.uniform(-1, 1, ncol)
            dxb = cb.adjoint_x(vb, dev(g, device), dev(d0, device), alpha=3.0,
                               beta=-1.0).cpu().numpy()
            den = 3.0 * absr + np.abs(d0) + 1e-300
            assert np.max(np.abs(dxb - (3.0 * want - d0)) / den, initial=0.0) <= F64_TOL, name
        cb.close()


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_colblock_autograd_matches_csr_operator(device, dtype):
    """The column-blocked operator's forward and gradients equal the CSR
    operator's within the bars, dval in block-major order."""
    from merbit_amd import workloads
    from merbit_amd.autograd import ColumnBlockedOperator, MerbitOperator, spmv
    say(f"colblock autograd {dtype}")
    A, x, g = workloads.graph(150_000)
    d_rp, d_col = dev(A.row_ptr, device), dev(A.col, device)
    val = dev(A.val, device).to(dtype)
    xx = dev(x, device).to(dtype)
    gg = dev(g, device).to(dtype)
    csr = MerbitOperator(d_rp, d_col, A.ncol)
    blk = ColumnBlockedOperator(d_rp, d_col, A.ncol, nblocks=3, adjoint_blocks=3)
    v1 = val.clone().requires_grad_(True)
    x1 = xx.clone().requires_grad_(True)
    y1 = spmv(csr, v1, x1)
    y1.backward(gg)
    vb = blk.to_blocked(val).requires_grad_(True)
    x2 = xx.clone().requires_grad_(True)
    y2 = spmv(blk, vb, x2)
    y2.backward(gg)
    tol = 1e-12 if dtype == torch.float64 else 1e-6

    def rel(a, b):
        a, b = a.detach().double(), b.detach().double()
        return float((a - b).norm() / b.norm())
    assert rel(y2, y1) <= tol
    assert rel(x2.grad, x1.grad) <= tol
    # the blocked dval is the CSR dval moved into block order, bit for bit
    assert torch.equal(vb.grad[:blk.value_slots], blk.to_blocked(v1.grad)[:blk.value_slots])
    blk.close()
    csr.close()


def test_colblock_rejects_nnz_other_than_row_ptr(device):
    """nnz must equal row_ptr[nrow]: the block counts are checked against it
    (the arrays themselves are whole, so every read stays in bounds)."""
    import ctypes
    from merbit_amd._lib import lib
    say("colblock nnz mismatch")
    rng = np.random.default_rng(5)
    rp, col = random_csr(rng, 500, 400, density=6)
    d_rp, d_col = dev(rp, device), dev(col, device)
    h = ctypes.c_void_p()
    nnz = int(rp[-1])
    for bad in (nnz - 1, nnz + 5):
        st = lib().mbx_colblock_create(d_rp.data_ptr(), d_col.data_ptr(), 500, 400, bad, 3, 8,
                                       None, ctypes.byref(h))
        assert st == -1 and not h.value
    st = lib().mbx_colblock_create(d_rp.data_ptr(), d_col.data_ptr(), 500, 400, nnz, 3, 8, None,
                                   ctypes.byref(h))
    assert st == 0 and h.value
    lib().mbx_colblock_destroy(h)


@pytest.mark.parametrize("kind", ["f64", "f32", "bf16"])
@pytest.mark.parametrize("blocks", [-1, 1, 3, 16])
def test_csr_adjoint_through_blocked_transpose(orc, device, kind, blocks):
    """mbx_adjoint_set_blocks: dx from CSR-order values through a blocked
    A^T (one bucketed permutation whose padding slots read virtual zero
    sources), against the oracle's accurate product on the transposed CSR;
    back at 0 blocks the single merge-path product returns bit for bit."""
    from merbit_amd import gen
    from merbit_amd._lib import check, lib
    from merbit_amd.autograd import MerbitOperator
    from merbit_amd.spmv import _stream_handle
    rng = np.random.default_rng(900 + blocks + 17)
    for name, lens, ncol in cases(rng):
        say(f"adjoint blocks={blocks} {kind} {name}")
        rp, col = random_csr(rng, len(lens), ncol, lens=lens)
        v = rng.standard_normal(int(rp[-1]))
        if kind == "f64":
            vals, g = v, rng.uniform(-1, 1, len(lens))
        else:
            v32 = v.astype(np.float32)
            vals = v32 if kind == "f32" else gen.to_bf16(v32)
            g = rng.uniform(-1, 1, len(lens)).astype(np.float32)
        d_rp, d_col = dev(rp, device), dev(col, device)
        op = MerbitOperator(d_rp, d_col, ncol)
        d_val, d_g = dev(vals, device), dev(g, device)
        kcode = {"f64": 0, "f32": 1, "bf16": 2}[kind]
        gdt = torch.float64 if kind == "f64" else torch.float32

        def dx_now():
            out = torch.empty(ncol, dtype=gdt, device=device)
            check(lib().mbx_adjoint_dx(op.adjoint_handle, kcode, d_val.data_ptr(), 1.0,
                                       d_g.data_ptr(), 0.0, out.data_ptr(), _stream_handle(None)),
                  "mbx_adjoint_dx")
            return out.cpu().numpy()
        plain = dx_now()
        check(lib().mbx_adjoint_set_blocks(op.adjoint_handle, blocks, _stream_handle(None)),
              "mbx_adjoint_set_blocks")
        blocked = dx_now()
        rpT, colT, valT = transpose_csr(rp, col, ncol, vals)
        want, absr = orc.spmv_accurate(rpT, colT, valT, g)
        assert cond_err(blocked, want, absr) <= (F64_TOL if kind == "f64" else MIXED_TOL), name
        check(lib().mbx_adjoint_set_blocks(op.adjoint_handle, 0, _stream_handle(None)),
              "mbx_adjoint_set_blocks")
        assert np.array_equal(dx_now().view(np.uint8), plain.view(np.uint8)), name
        op.close()


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_autograd_with_blocked_transpose(device, dtype):
    """MerbitOperator(adjoint_blocks=3): the autograd x gradient through the
    blocked A^T equals the single-product one within the bars."""
    from merbit_amd import workloads
    from merbit_amd.autograd import MerbitOperator, spmv
    say(f"autograd blocked transpose {dtype}")
    A, x, g = workloads.graph(150_000)
    d_rp, d_col = dev(A.row_ptr, device), dev(A.col, device)
    val = dev(A.val, device).to(dtype)
    gg = dev(g, device).to(dtype)
    grads = []
    for blocks in (0, 3):
        op = MerbitOperator(d_rp, d_col, A.ncol, adjoint_blocks=blocks)
        xx = dev(x, device).to(dtype).requires_grad_(True)
        spmv(op, val, xx).backward(gg)
        grads.append(xx.grad.detach().double())
        op.close()
    tol = 1e-12 if dtype == torch.float64 else 1e-6
    assert float((grads[1] - grads[0]).norm() / grads[0].norm()) <= tol


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_csr_operator_on_blocked_forms(device, dtype):
    """MerbitOperator(blocks=3): values stay in CSR order; the forward runs
    on the column-blocked form and dx on the blocked transpose.  Forward and
    dx agree with the plain operator within the bars; dval and the CSR
    round trip of the values are bit-identical."""
    from merbit_amd import workloads
    from merbit_amd.autograd import MerbitOperator, spmv
    say(f"csr operator blocked {dtype}")
    A, x, g = workloads.graph(150_000)
    d_rp, d_col = dev(A.row_ptr, device), dev(A.col, device)
    val = dev(A.val, device).to(dtype)
    gg = dev(g, device).to(dtype)
    res = []
    for blocks in (0, 3):
        op = MerbitOperator(d_rp, d_col, A.ncol, blocks=blocks)
        v = val.clone().requires_grad_(True)
        xx = dev(x, device).to(dtype).requires_grad_(True)
        y = spmv(op, v, xx)
        y.backward(gg)
        res.append((y.detach().double(), xx.grad.detach().double(), v.grad.detach()))
        if blocks:
            back = op._cb.values_to_csr(op._cb.values(val))
            assert torch.equal(back, val)
        op.close()
    tol = 1e-12 if dtype == torch.float64 else 1e-6
    for i in (0, 1):
        assert float((res[1][i] - res[0][i]).norm() / res[0][i].norm()) <= tol
    assert torch.equal(res[1][2], res[0][2])


def test_colblock_products_replay_in_a_graph(device):
    """The blocked product and the blocked dx record no memset nodes: captured
    once as a HIP graph and replayed several times (eager work between the
    replays), every replay equals the eager result bit for bit."""
    from merbit_amd import workloads
    from merbit_amd.spmv import ColumnBlockedCsr
    say("colblock graph replay")
    A, x, g = workloads.graph(100_000)
    d_rp, d_col = dev(A.row_ptr, device), dev(A.col, device)
    cb = ColumnBlockedCsr(d_rp, d_col, A.ncol, nblocks=3)
    cb.prepare_adjoint(d_rp, d_col, nblocks=3)
    vb = cb.values(dev(A.val, device))
    xx, gg = dev(x, device), dev(g, device)
    y = torch.empty(A.nrow, dtype=torch.float32, device=device)
    dx = torch.empty(A.ncol, dtype=torch.float32, device=device)
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):  # warm the kernels outside the capture
        cb.spmv(vb, xx, y, stream=s)
        cb.adjoint_x(vb, gg, dx, stream=s)
    torch.cuda.synchronize()
    y_ref, dx_ref = y.clone(), dx.clone()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph, stream=s):
        cb.spmv(vb, xx, y, stream=s)
        cb.adjoint_x(vb, gg, dx, stream=s)
    for i in range(3):
        y.fill_(float(i))
        dx.fill_(-1.0)
        torch.cuda.synchronize()
        graph.replay()
        torch.cuda.synchronize()
        assert torch.equal(y, y_ref) and torch.equal(dx, dx_ref), i
    del graph
    cb.close()
