"""GPU checks of the adjoint (dx = A^T g, dval = g[row]*x[col]) and the
torch.autograd function: transposed pattern vs the oracle's transpose
(exact), dx vs the oracle on the transpose (condition-scaled 1e-12 FP64,
1e-6 FP32), dval bit-exact against the FP32/FP64 products, gradients against
a dense reference and against central finite differences (FP64)."""
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


def make_op(rp, col, ncol, device, ipt=8):
    from merbit_amd.autograd import MerbitOperator
    return MerbitOperator(dev(rp, device), dev(col, device), ncol, items_per_thread=ipt)


def test_transpose_matches_oracle(orc, device):
    rng = np.random.default_rng(2)
    for name, lens, ncol in edge_case_patterns(2048) + [
            ("pareto", (rng.pareto(1.0, 20000) * 5).astype(int), 15000)]:
        say(f"transpose {name}")
        rp, col = random_csr(rng, len(lens), ncol, lens=lens)
        op = make_op(rp, col, ncol, device)
        cp, ri, pm = op.transpose()
        ocp, ori, opm = orc.csr_transpose(rp, col, ncol)
        assert np.array_equal(cp, ocp) and np.array_equal(ri, ori) and np.array_equal(pm, opm)
        op.close()


@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_adjoint_dx_and_dval(orc, device, dtype):
    rng = np.random.default_rng(3)
    npdt = np.float64 if dtype == "f64" else np.float32
    tol = 1e-12 if dtype == "f64" else 1e-6
    for name, lens, ncol in edge_case_patterns(2048) + [
            ("pareto", (rng.pareto(1.0, 30000) * 5).astype(int), 25000)]:
        say(f"adjoint {dtype} {name}")
        rp, col = random_csr(rng, len(lens), ncol, lens=lens)
        nrow = len(lens)
        val = rng.uniform(0, 1, int(rp[-1])).astype(npdt)
        x = rng.standard_normal(ncol).astype(npdt)
        g = rng.standard_normal(nrow).astype(npdt)
        op = make_op(rp, col, ncol, device)
        dx = op.adjoint_x(dev(val, device), dev(g, device)).cpu().numpy()
        cp, ri, pm = orc.csr_transpose(rp, col, ncol)
        ya, absr = orc.spmv_accurate(cp, ri, np.ascontiguousarray(val[pm]), g)
        assert cond_err(dx.astype(np.float64), ya, absr) <= tol, name
        dval = op.adjoint_val(dev(g, device), dev(x, device)).cpu().numpy()
        rows = np.repeat(np.arange(nrow), np.diff(rp))
        want = (g[rows] * x[col]).astype(npdt)  # one rounding, as on the device
        assert np.array_equal(dval, want), name
        op.close()


def test_autograd_dense_reference(device):
    from merbit_amd.autograd import spmv
    rng = np.random.default_rng(5)
    n, m = 300, 250
    rp, col = random_csr(rng, n, m, 6.0)
    op = make_op(rp, col, m, device)
    val = dev(rng.uniform(0, 1, int(rp[-1])), device).requires_grad_(True)
    x = dev(rng.standard_normal(m), device).requires_grad_(True)
    w = dev(rng.standard_normal(n), device)
    say("autograd vs dense")
    y = spmv(op, val, x)
    (y * w).sum().backward()
    # dense FP64 reference on the CPU
    rows = torch.from_numpy(np.repeat(np.arange(n), np.diff(rp)))
    vc = val.detach().cpu().clone().requires_grad_(True)
    xc = x.detach().cpu().clone().requires_grad_(True)
    A = torch.zeros((n, m), dtype=torch.float64).index_put((rows, torch.from_numpy(col).long()),
                                                           vc, accumulate=True)
    ((A @ xc) * w.cpu()).sum().backward()
    assert torch.allclose(y.detach().cpu(), (A @ xc).detach(), rtol=1e-12, atol=1e-12)
    assert torch.allclose(val.grad.cpu(), vc.grad, rtol=1e-12, atol=1e-14)
    assert torch.allclose(x.grad.cpu(), xc.grad, rtol=1e-12, atol=1e-12)
    op.close()


def test_autograd_finite_differences(device):
    from merbit_amd.autograd import spmv
    rng = np.random.default_rng(6)
    n = 400
    rp, col = random_csr(rng, n, n, lens=(rng.pareto(1.2, n) * 4).astype(int))
    op = make_op(rp, col, n, device, ipt=4)
    val = dev(rng.uniform(0, 1, int(rp[-1])), device).requires_grad_(True)
    x = dev(rng.standard_normal(n), device).requires_grad_(True)
    w = dev(rng.standard_normal(n), device)
    say("finite differences")
    f = lambda v, xx: float((spmv(op, v, xx) * w).sum())
    (spmv(op, val, x) * w).sum().backward()
    for _ in range(3):
        dv = dev(rng.standard_normal(val.numel()), device)
        dxd = dev(rng.standard_normal(n), device)
        h = 1e-3
        with torch.no_grad():
            fd_v = (f(val + h * dv, x) - f(val - h * dv, x)) / (2 * h)
            fd_x = (f(val, x + h * dxd) - f(val, x - h * dxd)) / (2 * h)
        an_v = float((val.grad * dv).sum())
        an_x = float((x.grad * dxd).sum())
        assert abs(fd_v - an_v) <= 1e-8 * (1 + abs(an_v))
        assert abs(fd_x - an_x) <= 1e-8 * (1 + abs(an_x))
    # torch's own checker on a tiny instance
    rp2, col2 = random_csr(rng, 12, 10, 3.0)
    op2 = make_op(rp2, col2, 10, device, ipt=2)
    v2 = dev(rng.uniform(0, 1, int(rp2[-1])), device).requires_grad_(True)
    x2 = dev(rng.standard_normal(10), device).requires_grad_(True)
    assert torch.autograd.gradcheck(lambda a, b: spmv(op2, a, b), (v2, x2))
    op.close()
    op2.close()


@pytest.mark.slow
def test_adjoint_graph_c5(orc, device):
    """Random graph with 2.4M nodes, Poisson(25) out-degree, FP32 values."""
    from merbit_amd import workloads
    say("adjoint graph 2.4M nodes")
    A, x, g = workloads.graph()
    op = make_op(A.row_ptr, A.col, A.ncol, device)
    val = dev(A.val, device)
    y = op.forward(val, dev(x, device)).cpu().numpy().astype(np.float64)
    ya, absr = orc.spmv_accurate(A.row_ptr, A.col, A.val, x)
    assert cond_err(y, ya, absr) <= 1e-6
    dx = op.adjoint_x(val, dev(g, device)).cpu().numpy().astype(np.float64)
    cp, ri, pm = orc.csr_transpose(A.row_ptr, A.col, A.ncol)
    da, dabs = orc.spmv_accurate(cp, ri, np.ascontiguousarray(A.val[pm]), g)
    assert cond_err(dx, da, dabs) <= 1e-6
    dval = op.adjoint_val(dev(g, device), dev(x, device)).cpu().numpy()
    rows = np.repeat(np.arange(A.nrow), np.diff(A.row_ptr))
    assert np.array_equal(dval, g[rows] * x[A.col])
    op.close()


@pytest.mark.parametrize("dtype", ["f64", "f32", "bf16"])
def test_adjoint_permutation_boundaries(orc, device, dtype):
    """The CSR-to-CSC value permutation runs in chunks and buckets of 16384
    entries: at sizes around those boundaries dx must equal, bit for bit, the
    same tile SpMV on the transposed pattern fed with val[perm] gathered by
    torch."""
    from merbit_amd import gen
    from merbit_amd._lib import MBX_VALUE_BF16, MBX_VALUE_F32, MBX_VALUE_F64, check, lib
    from merbit_amd.spmv import MerbitPlan
    kinds = {"f64": MBX_VALUE_F64, "f32": MBX_VALUE_F32, "bf16": MBX_VALUE_BF16}
    rng = np.random.default_rng(11)
    for nnz in (1, 16383, 16384, 16385, 3 * 16384 + 7, 200_003):
        say(f"permutation {dtype} nnz={nnz}")
        nrow, ncol = max(1, nnz // 9), max(1, nnz // 7)
        lens = rng.multinomial(nnz, np.full(nrow, 1.0 / nrow))
        rp, col = random_csr(rng, nrow, ncol, lens=lens)
        assert int(rp[-1]) == nnz
        v32 = rng.uniform(-1, 1, nnz).astype(np.float32)
        if dtype == "f64":
            val, g = v32.astype(np.float64), rng.standard_normal(nrow)
        else:
            val = v32 if dtype == "f32" else gen.to_bf16(v32)
            g = rng.standard_normal(nrow).astype(np.float32)
        op = make_op(rp, col, ncol, device)
        d_val, d_g = dev(val, device), dev(g, device)
        dx = torch.empty(ncol, dtype=d_g.dtype, device=device)
        check(lib().mbx_adjoint_dx(op.adjoint_handle, kinds[dtype], d_val.data_ptr(), 1.0,
                                   d_g.data_ptr(), 0.0, dx.data_ptr(),
                                   torch.cuda.current_stream().cuda_stream), "mbx_adjoint_dx")
        cp, ri, pm = orc.csr_transpose(rp, col, ncol)
        d_cp = dev(cp, device)
        plan_t = MerbitPlan(d_cp, nrow, nnz, items_per_thread=8)
        idx = dev(pm.astype(np.int64), device)
        # gather the raw bits (torch indexes int16, not uint16)
        vt = (d_val.view(torch.int16)[idx].view(torch.uint16) if dtype == "bf16"
              else d_val[idx].contiguous())
        want = plan_t.spmv(d_cp, dev(ri, device), vt, d_g)
        assert torch.equal(dx.view(torch.uint8), want.view(torch.uint8)), nnz
        # values one element into a larger buffer: not 4-element aligned
        holder = torch.zeros(nnz + 1, dtype=d_val.dtype, device=device)
        holder[1:] = d_val
        dx_u = torch.empty_like(dx)
        check(lib().mbx_adjoint_dx(op.adjoint_handle, kinds[dtype], holder[1:].data_ptr(), 1.0,
                                   d_g.data_ptr(), 0.0, dx_u.data_ptr(),
                                   torch.cuda.current_stream().cuda_stream), "mbx_adjoint_dx")
        assert torch.equal(dx_u.view(torch.uint8), want.view(torch.uint8)), nnz
        plan_t.close()
        op.close()


@pytest.mark.slow
@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_operator_default_adjoint_rule_rmat(orc, device, dtype):
    """A default MerbitOperator runs the launch rule on the transposed
    pattern (mbx_adjoint_auto).  The R-MAT's rows are skewed like its
    columns, so A^T meets the hot-columns rule, which with 2^20 columns
    takes tile-sorted gathers at 16 items per thread; for FP64 dx is
    bit-identical to the plain single-product adjoint of the same shape, for
    FP32 the rule also picks FP32 products with FP64 sums, so dx matches the
    plain adjoint within the FP32 bar; both match the oracle's accurate A^T
    product."""
    from merbit_amd import workloads
    from merbit_amd.autograd import MerbitOperator
    A, _ = workloads.rmat()
    say(f"default operator adjoint rmat {dtype}")
    d_rp, d_col = dev(A.row_ptr, device), dev(A.col, device)
    rng = np.random.default_rng(31)
    g = rng.uniform(-1, 1, A.nrow)
    vals = A.val if dtype == torch.float64 else A.val.astype(np.float32)
    d_val, d_g = dev(vals, device), dev(g, device).to(dtype)
    op = MerbitOperator(d_rp, d_col, A.ncol, value_dtype=dtype)
    assert op.adjoint_blocks is None
    dx = op.adjoint_x(d_val, d_g)
    c = op.adjoint_choice
    assert c["rule"] == "hot_columns" and c["sort_cols"] == 1 and c["hot_columns"] == 0, c
    ref_op = MerbitOperator(d_rp, d_col, A.ncol, items_per_thread=c["items_per_thread"],
                            blocks=0, adjoint_blocks=0, value_dtype=dtype)
    dx_ref = ref_op.adjoint_x(d_val, d_g)
    cp, ri, pm = orc.csr_transpose(A.row_ptr, A.col, A.ncol)
    gh = d_g.double().cpu().numpy()
    da, dabs = orc.spmv_accurate(cp, ri, np.ascontiguousarray(vals[pm]).astype(np.float64), gh)
    tol = 1e-12 if dtype == torch.float64 else 1e-6
    if dtype == torch.float64:
        assert c["accumulation"] == 0
        assert torch.equal(dx.view(torch.uint8), dx_ref.view(torch.uint8))
    else:
        assert c["accumulation"] == 2  # FP32 products, FP64 sums
        assert cond_err(dx.double().cpu().numpy(), dx_ref.double().cpu().numpy(), dabs) <= 2 * tol
    assert cond_err(dx.double().cpu().numpy(), da, dabs) <= tol
    ref_op.close()
    op.close()


@pytest.mark.slow
def test_operator_default_adjoint_rule_banded(orc, device):
    """A banded pattern's transpose is banded: the rule on A^T picks the
    local form (16-bit columns, FP32 products with FP64 sums for FP32
    values); dx matches the oracle and the plain adjoint within the FP32
    bar."""
    from merbit_amd.autograd import MerbitOperator
    n, offs = 1_200_000, np.arange(-3, 4)
    say("default operator adjoint banded")
    rows = np.repeat(np.arange(n), len(offs))
    cols = rows + np.tile(offs, n)
    keep = (cols >= 0) & (cols < n)  # a clipped band: every tile spans few columns
    rows, cols = rows[keep], cols[keep].astype(np.int32)
    rp = np.zeros(n + 1, np.int32)
    rp[1:] = np.cumsum(np.bincount(rows, minlength=n))
    rng = np.random.default_rng(32)
    vals = rng.uniform(-1, 1, cols.size).astype(np.float32)
    g = rng.uniform(-1, 1, n).astype(np.float32)
    d_rp, d_col, d_val, d_g = dev(rp, device), dev(cols, device), dev(vals, device), \
        dev(g, device)
    op = MerbitOperator(d_rp, d_col, n, value_dtype=torch.float32)
    dx = op.adjoint_x(d_val, d_g)
    c = op.adjoint_choice
    assert c["rule"] == "local" and c["compress_cols"] == 1 and c["accumulation"] == 2, c
    cp, ri, pm = orc.csr_transpose(rp, cols, n)
    da, dabs = orc.spmv_accurate(cp, ri, np.ascontiguousarray(vals[pm]).astype(np.float64),
                                 g.astype(np.float64))
    assert cond_err(dx.double().cpu().numpy(), da, dabs) <= 1e-6
    ref_op = MerbitOperator(d_rp, d_col, n, items_per_thread=8, blocks=0, adjoint_blocks=0,
                            value_dtype=torch.float32)
    ref = ref_op.adjoint_x(d_val, d_g).double()
    assert float(torch.linalg.norm(dx.double() - ref) / torch.linalg.norm(ref)) <= 1e-6
    ref_op.close()
    op.close()
