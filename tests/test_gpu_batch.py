"""GPU checks of the persistent multi-matrix batch, HIP-graph replay and the
autotuner.  The batch runs the single-matrix per-tile code, so its results
must equal one mbx_spmv_* call per matrix bit for bit; graph replays must
equal eager products bit for bit, replay after replay."""
from __future__ import annotations

import numpy as np
import pytest
import torch

from helpers import cond_err, random_csr

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[case] {msg}", flush=True)


def make_mats(device, count, kind, seed, rows_hi=20000):
    from merbit_amd import gen, workloads
    mats, xs = workloads.uneven_batch(count, rows_lo=1000, rows_hi=rows_hi)
    out = []
    for A, x in zip(mats, xs):
        if kind == "f64":
            v, xv = A.val.astype(np.float64), x.astype(np.float64)
        elif kind == "f32":
            v, xv = A.val, x
        else:
            v, xv = gen.to_bf16(A.val), x
        vdt = torch.float64 if kind == "f64" else torch.float32
        out.append((torch.from_numpy(A.row_ptr).to(device), torch.from_numpy(A.col).to(device),
                    torch.from_numpy(v).to(device), torch.from_numpy(xv).to(device),
                    torch.full((A.nrow,), float("nan"), dtype=vdt, device=device), A.ncol))
    return out


@pytest.mark.parametrize("kind", ["f64", "f32", "bf16"])
@pytest.mark.parametrize("ipt", [4, 8])
def test_batch_matches_single_calls(device, kind, ipt):
    from merbit_amd.batch import MerbitBatch
    from merbit_amd.spmv import MerbitPlan
    say(f"batch {kind} ipt={ipt}")
    mats = make_mats(device, 12, kind, 0)
    b = MerbitBatch(mats, items_per_thread=ipt)
    b.spmv()
    torch.cuda.synchronize()
    got = [m[4].clone() for m in mats]
    for (rp, col, val, x, y, ncol), g in zip(mats, got):
        plan = MerbitPlan(rp, ncol, col.numel(), items_per_thread=ipt)
        ref = plan.spmv(rp, col, val, x)
        plan.close()
        assert torch.equal(ref.view(torch.uint8), g.view(torch.uint8))
    b.close()


@pytest.mark.slow
def test_batch_full(orc, device):
    """The 64-matrix uneven batch at full size (6.3M rows, 214M entries,
    FP32): one batch product equals the 64 single-matrix products bit for
    bit, and every eighth matrix meets the mixed bar against the oracle."""
    from merbit_amd.batch import MerbitBatch
    from merbit_amd.spmv import MerbitPlan
    from merbit_amd import workloads
    say("batch 64 matrices full size, f32")
    mats_h, xs_h = workloads.uneven_batch()
    mats = [(torch.from_numpy(A.row_ptr).to(device), torch.from_numpy(A.col).to(device),
             torch.from_numpy(A.val).to(device), torch.from_numpy(x).to(device),
             torch.full((A.nrow,), float("nan"), dtype=torch.float32, device=device), A.ncol)
            for A, x in zip(mats_h, xs_h)]
    b = MerbitBatch(mats, items_per_thread=16)
    b.spmv()
    torch.cuda.synchronize()
    for i, ((rp, col, val, x, y, ncol), A, xh) in enumerate(zip(mats, mats_h, xs_h)):
        plan = MerbitPlan(rp, ncol, col.numel(), items_per_thread=16)
        ref = plan.spmv(rp, col, val, x)
        plan.close()
        assert torch.equal(ref.view(torch.uint8), y.view(torch.uint8)), i
        if i % 8 == 0:
            ya, absr = orc.spmv_accurate(A.row_ptr, A.col, A.val, xh)
            assert cond_err(y.cpu().numpy().astype(np.float64), ya, absr) <= 1e-6, i
    b.close()


@pytest.mark.parametrize("kind", ["f64", "f32"])
def test_batch_long_rows(orc, device, kind):
    """Rows spanning more than 64 tiles are completed by a whole wave of the
    fixup; in a batch that wave also holds tiles of other matrices."""
    from helpers import random_csr
    from merbit_amd.batch import MerbitBatch
    from merbit_amd.spmv import MerbitPlan
    say(f"batch long rows {kind}")
    rng = np.random.default_rng(77)
    T = 256 * 4
    dt = np.float64 if kind == "f64" else np.float32
    mats, host = [], []
    for lens in ([1, 66 * T + 5, 2], [3] * 50, [2, 70 * T + 1], [0, 5, 0], [1, 65 * T, 1]):
        ncol = 300
        rp, col = random_csr(rng, len(lens), ncol, lens=lens)
        val = rng.standard_normal(int(rp[-1])).astype(dt)
        x = rng.uniform(-1, 1, ncol).astype(dt)
        host.append((rp, col, val, x))
        mats.append((torch.from_numpy(rp).to(device), torch.from_numpy(col).to(device),
                     torch.from_numpy(val).to(device), torch.from_numpy(x).to(device),
                     torch.full((len(lens),), float("nan"), dtype=torch.from_numpy(x).dtype,
                                device=device), ncol))
    b = MerbitBatch(mats, items_per_thread=4)
    b.spmv()
    torch.cuda.synchronize()
    tol = 1e-12 if kind == "f64" else 1e-6
    for (rp, col, val, x, y, ncol), (hrp, hcol, hval, hx) in zip(mats, host):
        plan = MerbitPlan(rp, ncol, col.numel(), items_per_thread=4)
        ref = plan.spmv(rp, col, val, x)
        plan.close()
        assert torch.equal(ref.view(torch.uint8), y.view(torch.uint8))
        ya, absr = orc.spmv_accurate(hrp, hcol, hval.astype(np.float64), hx.astype(np.float64))
        assert cond_err(y.cpu().numpy().astype(np.float64), ya, absr) <= tol
    b.close()


def test_batch_against_oracle(orc, device):
    from merbit_amd.batch import MerbitBatch
    say("batch f32 vs oracle")
    mats = make_mats(device, 16, "f32", 1, rows_hi=60000)
    b = MerbitBatch(mats)
    b.spmv()
    torch.cuda.synchronize()
    for rp, col, val, x, y, ncol in mats:
        ya, absr = orc.spmv_accurate(rp.cpu().numpy(), col.cpu().numpy(), val.cpu().numpy(),
                                     x.cpu().numpy())
        assert cond_err(y.cpu().numpy().astype(np.float64), ya, absr) <= 1e-6
    b.close()


def test_batch_graph_replays(device):
    from merbit_amd.batch import MerbitBatch
    say("batch graph replays")
    mats = make_mats(device, 8, "f32", 2)
    b = MerbitBatch(mats, items_per_thread=8)
    b.spmv(alpha=0.5)
    torch.cuda.synchronize()
    eager = [m[4].clone() for m in mats]
    g = b.graph(alpha=0.5, repeats=3)
    for rep in range(4):
        for m in mats:
            m[4].fill_(float("nan"))        # eager work between replays
        torch.cuda.synchronize()
        g.launch()
        torch.cuda.synchronize()
        for m, e in zip(mats, eager):
            assert torch.equal(m[4].view(torch.uint8), e.view(torch.uint8)), rep
    g.close()
    b.close()


@pytest.mark.parametrize("kind", ["f64", "f32"])
def test_spmv_graph_replays(device, kind):
    from merbit_amd import workloads
    from merbit_amd.batch import spmv_graph
    from merbit_amd.spmv import MerbitPlan
    say(f"spmv graph {kind}")
    A, x = workloads.rmat(14)
    dt = torch.float64 if kind == "f64" else torch.float32
    rp, col = torch.from_numpy(A.row_ptr).to(device), torch.from_numpy(A.col).to(device)
    val = torch.from_numpy(A.val).to(device).to(dt)
    xx = torch.from_numpy(x).to(device).to(dt)
    y = torch.empty(A.nrow, dtype=dt, device=device)
    plan = MerbitPlan(rp, A.ncol, A.nnz)
    ref = plan.spmv(rp, col, val, xx).clone()
    g = spmv_graph(plan, rp, col, val, xx, y, repeats=5)
    for rep in range(3):
        y.zero_()
        torch.cuda.synchronize()
        g.launch()
        torch.cuda.synchronize()
        assert torch.equal(y.view(torch.uint8), ref.view(torch.uint8)), rep
    g.close()
    plan.close()


def test_autotuners(orc, device):
    from merbit_amd import workloads
    from merbit_amd.batch import autotune_batch, autotune_spmm, autotune_spmv
    from merbit_amd.spmv import MerbitPlan
    say("autotune spmv")
    A, x = workloads.rmat(15)
    rp, col = torch.from_numpy(A.row_ptr).to(device), torch.from_numpy(A.col).to(device)
    val = torch.from_numpy(A.val).to(device)
    xx = torch.from_numpy(x).to(device)
    y = torch.empty(A.nrow, dtype=torch.float64, device=device)
    ipt, us = autotune_spmv(rp, col, val, xx, y, reps=3)
    assert ipt in (2, 4, 6, 8, 10, 12, 14, 16) and us > 0
    plan = MerbitPlan(rp, A.ncol, A.nnz, items_per_thread=ipt)
    yy = plan.spmv(rp, col, val, xx).cpu().numpy()
    ya, absr = orc.spmv_accurate(A.row_ptr, A.col, A.val, x)
    assert cond_err(yy, ya, absr) <= 1e-12
    plan.close()
    say("autotune spmv launch shape")
    from merbit_amd.batch import autotune_spmv_launch
    ipt, mode, wg, us = autotune_spmv_launch(rp, col, val, xx, y, reps=2)
    assert ipt in (2, 4, 6, 8, 10, 12, 14, 16) and mode in ("cached", "nontemporal")
    assert wg in (0, 3, 4, 5) and us > 0
    plan = MerbitPlan(rp, A.ncol, A.nnz, items_per_thread=ipt)
    plan.set_stream(mode)
    plan.set_residency(wg)
    yy = plan.spmv(rp, col, val, xx).cpu().numpy()
    assert cond_err(yy, ya, absr) <= 1e-12
    plan.close()
    say("autotune spmm")
    X = torch.rand((A.ncol, 8), device=device)
    Y = torch.empty((A.nrow, 8), device=device)
    ipt, us = autotune_spmm(rp, col, val.float(), X, Y, reps=2)
    assert ipt in (2, 4) and us > 0
    say("autotune batch")
    mats = make_mats(device, 6, "f32", 3)
    ipt, us = autotune_batch(mats, reps=2)
    assert ipt in (2, 4, 6, 8, 10, 12, 14, 16) and us > 0


def test_batch_unaligned_member(device):
    """A batch with one matrix whose arrays are not 16-byte aligned runs the
    scalar staging path for every matrix; results still equal single calls
    bit for bit."""
    from merbit_amd.batch import MerbitBatch
    from merbit_amd.spmv import MerbitPlan
    say("batch with an unaligned member")
    mats = make_mats(device, 6, "f32", 3)
    rp, col, val, x, y, ncol = mats[2]
    # views starting one element in: 4-byte aligned, not 16-byte aligned
    rp_u = torch.empty(rp.numel() + 1, dtype=rp.dtype, device=device)[1:]
    rp_u.copy_(rp)
    col_u = torch.empty(col.numel() + 1, dtype=col.dtype, device=device)[1:]
    col_u.copy_(col)
    val_u = torch.empty(val.numel() + 1, dtype=val.dtype, device=device)[1:]
    val_u.copy_(val)
    mats[2] = (rp_u, col_u, val_u, x, y, ncol)
    b = MerbitBatch(mats, items_per_thread=8)
    b.spmv()
    torch.cuda.synchronize()
    got = [m[4].clone() for m in mats]
    for (rp, col, val, x, y, ncol), g in zip(mats, got):
        plan = MerbitPlan(rp, ncol, col.numel(), items_per_thread=8)
        ref = plan.spmv(rp, col, val, x)
        plan.close()
        assert torch.equal(ref.view(torch.uint8), g.view(torch.uint8))
    b.close()


@pytest.mark.parametrize("kind", ["f64", "bf16"])
def test_batch_degenerate_members(device, kind):
    """Members with no rows, with rows but no entries, a single entry, and a
    normal matrix between them: every member equals its single-matrix
    product bit for bit, and a member with no entries gets y = 0."""
    from merbit_amd import gen
    from merbit_amd.batch import MerbitBatch
    from merbit_amd.spmv import MerbitPlan
    say(f"batch degenerate members {kind}")
    rng = np.random.default_rng(31)
    shapes = [([], 5), ([0] * 1000, 40), ([1], 1), (list(rng.poisson(9, 3000)), 2500), ([], 0),
              ([0, 0, 3, 0], 7)]
    mats = []
    for lens, ncol in shapes:
        rp, col = random_csr(rng, len(lens), max(ncol, 1), lens=lens) if lens else \
            (np.zeros(1, np.int32), np.zeros(0, np.int32))
        v32 = rng.uniform(-1, 1, int(rp[-1])).astype(np.float32)
        if kind == "f64":
            v, x = v32.astype(np.float64), rng.uniform(-1, 1, ncol)
        else:
            v, x = gen.to_bf16(v32), rng.uniform(-1, 1, ncol).astype(np.float32)
        ydt = torch.float64 if kind == "f64" else torch.float32
        mats.append((torch.from_numpy(rp).to(device), torch.from_numpy(col).to(device),
                     torch.from_numpy(v).to(device), torch.from_numpy(x).to(device),
                     torch.full((len(lens),), float("nan"), dtype=ydt, device=device), ncol))
    b = MerbitBatch(mats, items_per_thread=8)
    b.spmv()
    torch.cuda.synchronize()
    for i, (rp, col, val, x, y, ncol) in enumerate(mats):
        if y.numel() == 0:
            continue
        plan = MerbitPlan(rp, ncol, col.numel(), items_per_thread=8)
        ref = plan.spmv(rp, col, val, x)
        plan.close()
        assert torch.equal(ref.view(torch.uint8), y.view(torch.uint8)), i
        if col.numel() == 0:
            assert torch.all(y == 0), i
    b.close()
