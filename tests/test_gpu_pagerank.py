"""PageRank power iterations (mbx_pagerank_f64, the damping factor the
reference carries in PARAMETER::damp but leaves unused) against the oracle's
iterations, plus the graph variant, the C++ driver and the argument checks."""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def ref_graph(orc, scale, seed):
    """Reference-semantics operated matrix of an R-MAT edge list (CSR of P^T,
    values 1/out-degree); R-MAT leaves many nodes without out-edges."""
    from merbit_amd import gen
    A = gen.rmat(scale, 8 << scale, seed=seed)
    row = np.repeat(np.arange(A.nrow, dtype=np.int32), np.diff(A.row_ptr))
    rp, col, val = orc.ref_ingest(A.nrow, row, A.col.astype(np.int32))
    return A.nrow, rp, col, val, int((np.diff(A.row_ptr) == 0).sum())


@pytest.mark.parametrize("ipt", [4, 8, 16])
def test_pagerank_matches_oracle(orc, device, ipt):
    from merbit_amd.spmv import MerbitPlan
    n, rp, col, val, dangling = ref_graph(orc, 12, 5)
    assert dangling > 0
    print(f"[case] pagerank rmat scale 12 ipt {ipt}", flush=True)
    x0 = np.random.default_rng(1).uniform(0.5, 1.5, n)
    t = [torch.from_numpy(a).to(device) for a in (rp, col, val, x0)]
    plan = MerbitPlan(t[0], n, len(col), items_per_thread=ipt)
    for iters in (1, 2, 30):
        y = plan.pagerank(t[0], t[1], t[2], damp=0.85, iters=iters, x0=t[3])
        torch.cuda.synchronize()
        ref = orc.pagerank(rp, col, val, 0.85, iters, x0)
        got = y.cpu().numpy()
        assert np.max(np.abs(got - ref) / np.abs(ref)) <= 1e-12
        assert abs(got.sum() - x0.sum()) <= 1e-12 * x0.sum()
    assert np.array_equal(t[3].cpu().numpy(), x0)  # x0 is not modified
    plan.close()


def test_pagerank_graph_replays_match(orc, device):
    from merbit_amd._lib import check, lib
    from merbit_amd.spmv import MerbitPlan
    n, rp, col, val, _ = ref_graph(orc, 11, 8)
    print("[case] pagerank graph rmat scale 11", flush=True)
    x0 = torch.full((n,), 1.0 / n, dtype=torch.float64, device=device)
    rp_d, col_d, val_d = (torch.from_numpy(a).to(device) for a in (rp, col, val))
    plan = MerbitPlan(rp_d, n, len(col))
    eager = plan.pagerank(rp_d, col_d, val_d, damp=0.9, iters=12, x0=x0).cpu()
    nbytes = int(lib().mbx_pagerank_workspace_bytes(n))
    work = torch.empty(nbytes, dtype=torch.uint8, device=device)
    y = torch.empty(n, dtype=torch.float64, device=device)
    g = ctypes.c_void_p()
    check(lib().mbx_pagerank_graph_create(plan.handle, rp_d.data_ptr(), col_d.data_ptr(),
                                          val_d.data_ptr(), 0.9, 12, x0.data_ptr(),
                                          y.data_ptr(), work.data_ptr(), nbytes,
                                          ctypes.byref(g)), "mbx_pagerank_graph_create")
    s = torch.cuda.current_stream().cuda_stream
    for rep in range(3):
        y.fill_(-1.0)  # eager work between replays
        check(lib().mbx_graph_launch(g, s), "mbx_graph_launch")
        torch.cuda.synchronize()
        assert torch.equal(y.cpu(), eager), f"replay {rep}"
    lib().mbx_graph_destroy(g)
    plan.close()


def test_pagerank_edge_cases(orc, device):
    from merbit_amd._lib import MerbitError, lib
    from merbit_amd.spmv import MerbitPlan
    n, rp, col, val, _ = ref_graph(orc, 8, 2)
    print("[case] pagerank edge cases", flush=True)
    rp_d, col_d, val_d = (torch.from_numpy(a).to(device) for a in (rp, col, val))
    plan = MerbitPlan(rp_d, n, len(col))
    x0 = torch.from_numpy(np.random.default_rng(3).uniform(0, 1, n)).to(device)
    y = plan.pagerank(rp_d, col_d, val_d, damp=0.0, iters=1, x0=x0)
    assert torch.allclose(y, torch.full_like(y, float(x0.sum()) / n), rtol=1e-13, atol=0)
    with pytest.raises(MerbitError):
        plan.pagerank(rp_d, col_d, val_d, iters=0, x0=x0)
    with pytest.raises(MerbitError):
        plan.pagerank(rp_d, col_d, val_d, damp=1.5, x0=x0)
    nbytes = int(lib().mbx_pagerank_workspace_bytes(n))
    work = torch.empty(nbytes, dtype=torch.uint8, device=device)
    out = torch.empty(n, dtype=torch.float64, device=device)
    st = lib().mbx_pagerank_f64(plan.handle, rp_d.data_ptr(), col_d.data_ptr(), val_d.data_ptr(),
                                0.85, 3, x0.data_ptr(), out.data_ptr(), work.data_ptr(),
                                nbytes - 8, torch.cuda.current_stream().cuda_stream)
    assert st != 0  # workspace too small
    plan.close()
    rect_rp = torch.tensor([0, 1, 2], dtype=torch.int32, device=device)  # 2 x 3
    rect = MerbitPlan(rect_rp, 3, 2)
    with pytest.raises(ValueError):
        rect.pagerank(rect_rp, torch.tensor([0, 2], dtype=torch.int32, device=device),
                      torch.ones(2, dtype=torch.float64, device=device))
    rect.close()


def test_driver_pagerank(device, tmp_path):
    from test_gpu_driver import BIN, write_rmat_mtx
    write_rmat_mtx(tmp_path / "rmat.mtx", 13, 4)
    print("[case] driver --pagerank", flush=True)
    r = subprocess.run([BIN, "--mtx", str(tmp_path / "rmat.mtx"), "--pagerank", "--damp",
                        "0.85", "--maxiter", "25", "--sigma", "8", "--check"],
                       stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=120)
    assert r.returncode == 0, r.stdout
    assert "check ok" in r.stdout and "ms/iteration" in r.stdout
