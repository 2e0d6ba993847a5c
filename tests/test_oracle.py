"""Pins the CPU oracle (oracle/merbit_oracle.c) against independent checks:
brute force, closed forms, hand-derived bit images, scipy/numpy and exact
rational arithmetic.  CPU only."""
from __future__ import annotations

import json
import os
from fractions import Fraction

import numpy as np
import pytest
import scipy.sparse as sp

from helpers import cond_err, random_csr

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def brute_path(row_ptr):
    """All merge-path coordinates by the linear merge (reference tie rule:
    an nnz index is consumed before an equal row end)."""
    nrow = len(row_ptr) - 1
    nnz = int(row_ptr[-1])
    x = y = 0
    pts = [(0, 0)]
    steps = []
    while x + y < nrow + nnz:
        if y < nrow and x < row_ptr[y + 1]:
            x += 1
            steps.append(0)
        else:
            y += 1
            steps.append(1)
        pts.append((x, y))
    return pts, steps


def small_cases(rng):
    cases = [np.array([0], np.int32), np.array([0, 0], np.int32),
             np.array([0, 5], np.int32), np.array([0, 0, 0, 3, 3], np.int32),
             np.array([0, 2, 2, 5, 6], np.int32)]
    for n in [1, 2, 7, 33]:
        for dens in [0.0, 0.5, 3.0]:
            rp, _ = random_csr(rng, n, 10, dens)
            cases.append(rp)
    return cases


def test_merge_search_brute_force(orc):
    rng = np.random.default_rng(7)
    for rp in small_cases(rng):
        pts, _ = brute_path(rp)
        nnz = int(rp[-1])
        for d, (bx, by) in enumerate(pts):
            assert orc.merge_search(rp, nnz, d) == (bx, by), (rp.tolist(), d)


def test_merge_search_invariants(orc):
    rng = np.random.default_rng(11)
    rp, _ = random_csr(rng, 300, 50, lens=rng.poisson(4, 300) * (rng.random(300) < 0.6))
    nnz = int(rp[-1])
    nrow = len(rp) - 1
    prev = (0, 0)
    for d in range(nrow + nnz + 1):
        x, y = orc.merge_search(rp, nnz, d)
        assert x + y == d
        assert 0 <= y <= nrow and 0 <= x <= nnz
        if y < nrow:
            assert rp[y] <= x <= rp[y + 1] if y > 0 else x <= rp[1]
        assert x >= prev[0] and y >= prev[1]
        prev = (x, y)
    assert orc.merge_search(rp, nnz, 0) == (0, 0)
    assert orc.merge_search(rp, nnz, nrow + nnz) == (nnz, nrow)


def test_ref_plan_golden(orc):
    with open(os.path.join(GOLDEN, "ref_plan_cases.json")) as f:
        cases = json.load(f)["cases"]
    for c in cases:
        rp = np.array(c["row_ptr"], np.int32)
        tx, ty, desc = orc.ref_plan(rp, int(rp[-1]), c["sigma"])
        assert tx.tolist() == c["tile_x"], c["name"]
        assert ty.tolist() == c["tile_y"], c["name"]
        if "lane_desc" in c:
            assert desc.tolist() == c["lane_desc"], c["name"]
        for k, v in c.get("lane_desc_tail", {}).items():
            assert int(desc[int(k)]) == v, (c["name"], k)
        for k, v in c.get("lane_desc_sample", {}).items():
            assert int(desc[int(k)]) == v, (c["name"], k)


@pytest.mark.parametrize("sigma", [1, 3, 7, 14, 15, 22, 42])
def test_ref_plan_brute_force(orc, sigma):
    rng = np.random.default_rng(sigma)
    for rp in small_cases(rng) + [random_csr(rng, 200, 40, 6.0)[0]]:
        nrow, nnz = len(rp) - 1, int(rp[-1])
        tx, ty, desc = orc.ref_plan(rp, nnz, sigma)
        pts, steps = brute_path(rp)
        wide = sigma >= 15
        dsh, fsh = (11, 22) if wide else (9, 18)
        mask = (1 << dsh) - 1
        total_flags = 0
        for gid, d in enumerate(desc.tolist()):
            t = gid // 32
            lx, ly = pts[gid * sigma]
            wx, wy = pts[t * 32 * sigma]
            assert d & mask == lx - wx
            assert (d >> dsh) & mask == ly - wy
            flags = d >> fsh
            want = 0
            for i, s in enumerate(steps[gid * sigma: gid * sigma + sigma]):
                want |= s << i
            assert flags == want
            total_flags += bin(flags).count("1")
        assert total_flags == nrow
        ntiles = len(tx) - 1
        for t in range(ntiles):
            wx, wy = pts[t * 32 * sigma]
            end = min((t + 1) * 32 * sigma, nrow + nnz)
            fast = pts[end][1] == wy
            assert tx[t] == wx and (ty[t] & 0x7FFFFFFF) == wy
            assert bool(ty[t] >> 31) == fast
        assert tx[-1] == nnz and ty[-1] == nrow


def laplacian_csr(g):
    I = sp.identity(g)
    T = sp.diags([-1, 2, -1], [-1, 0, 1], shape=(g, g))
    A = (sp.kron(I, T) + sp.kron(T, I)).tocsr()
    A.sort_indices()
    return A


def test_spmv_csr_laplacian_closed_form(orc):
    g = 16
    A = laplacian_csr(g)
    y = orc.spmv_csr(A.indptr, A.indices, A.data, np.ones(g * g))
    # x = 1: 4 minus the number of neighbours: 0 inside, 1 on edges, 2 at corners
    r, c = np.divmod(np.arange(g * g), g)
    want = (r == 0).astype(float) + (r == g - 1) + (c == 0) + (c == g - 1)
    assert np.array_equal(y, want)


def test_spmv_csr_vs_scipy_and_dense(orc):
    rng = np.random.default_rng(3)
    for n, dens in [(1, 2.0), (50, 3.0), (700, 9.0)]:
        rp, col = random_csr(rng, n, n, dens)
        val = rng.standard_normal(int(rp[-1]))
        x = rng.uniform(-1, 1, n)
        A = sp.csr_matrix((val, col, rp), shape=(n, n))
        ya, absr = orc.spmv_accurate(rp, col, val, x)
        y = orc.spmv_csr(rp, col, val, x)
        assert cond_err(y, A @ x, absr) < 1e-14
        dense = A.toarray() @ x
        assert cond_err(ya, dense, absr) < 1e-14
        y2 = orc.spmv_csr(rp, col, val, x, alpha=2.0, beta=-0.5, y=np.ones(n))
        assert np.allclose(y2, 2 * (A @ x) - 0.5, rtol=1e-13, atol=1e-13)


def exact_row_sums(rp, col, val, x):
    out = []
    for r in range(len(rp) - 1):
        s = Fraction(0)
        for k in range(rp[r], rp[r + 1]):
            s += Fraction(float(val[k])) * Fraction(float(x[col[k]]))
        out.append(s)
    return out


def test_accurate_vs_exact_rationals(orc):
    rng = np.random.default_rng(5)
    n = 40
    rp, col = random_csr(rng, n, n, 12.0)
    val = rng.standard_normal(int(rp[-1])) * 10.0 ** rng.integers(-8, 9, int(rp[-1]))
    x = rng.uniform(-1, 1, n)
    # cancellation-heavy rows: exact sum far below the terms
    k = int(rp[3])
    val[k:k + 6] = [1e16, 1.0, -1e16, 3.0, 1e-3, -1e-3]
    assert rp[4] - rp[3] >= 6
    y, _ = orc.spmv_accurate(rp, col, val, x)
    ex = exact_row_sums(rp, col, val, x)
    for r in range(n):
        e = float(ex[r])
        assert abs(y[r] - e) <= 2.3e-16 * abs(e) + 1e-300, r
    # float32 and bf16 inputs are widened exactly
    v32 = val.astype(np.float32)
    x32 = x.astype(np.float32)
    y32, _ = orc.spmv_accurate(rp, col, v32, x32)
    ex32 = exact_row_sums(rp, col, v32.astype(np.float64), x32.astype(np.float64))
    for r in range(n):
        assert abs(y32[r] - float(ex32[r])) <= 2.3e-16 * abs(float(ex32[r])) + 1e-300
    h = (v32.view(np.uint32) >> 16).astype(np.uint16)
    vb = (h.astype(np.uint32) << 16).view(np.float32)
    yb, _ = orc.spmv_accurate(rp, col, h, x32)
    exb = exact_row_sums(rp, col, vb.astype(np.float64), x32.astype(np.float64))
    for r in range(n):
        assert abs(yb[r] - float(exb[r])) <= 2.3e-16 * abs(float(exb[r])) + 1e-300


def test_emulator_grouping_hand_derived(orc):
    # One row, 4 products, sigma = 2: lanes hold (p0+p1), (p2+p3); the third
    # lane ends the row; the segmented scan gives lane 1 (p2+p3)+(p0+p1),
    # which the row-ending lane adds to the zero the walk left in s_y.
    rp = np.array([0, 4], np.int32)
    col = np.arange(4, dtype=np.int32)
    val = np.array([1.0, 1e16, -1e16, 1.0])
    x = np.ones(4)
    assert orc.merbit_emulate(rp, col, val, x, sigma=2)[0] == 0.0
    assert orc.spmv_csr(rp, col, val, x)[0] == 1.0  # sequential order differs
    # 40 products in one row with sigma = 1: tile 0 is "fast" (xor-butterfly
    # tree of lanes 0..31), tile 1 is normal: lanes 32..39 scanned with
    # strides 1, 2, 4 and added by the row-ending lane 40.
    rng = np.random.default_rng(9)
    p = rng.standard_normal(40) * 10.0 ** rng.integers(-6, 7, 40)
    rp = np.array([0, 40], np.int32)
    col = np.arange(40, dtype=np.int32)
    v = [np.float64(t) for t in p]
    lanes = v[:32]
    for dl in (16, 8, 4, 2, 1):
        lanes = [lanes[i] + lanes[i ^ dl] for i in range(32)]
    tree = lanes[0]
    q = v[32:40]
    l7 = ((q[7] + q[6]) + (q[5] + q[4])) + ((q[3] + q[2]) + (q[1] + q[0]))
    want = (np.float64(0.0) + tree + np.float64(0.0)) + l7
    got = orc.merbit_emulate(rp, col, p, np.ones(40), sigma=1)[0]
    assert got == want


@pytest.mark.parametrize("sigma", [1, 2, 5, 14, 15, 32])
def test_emulator_matches_accurate(orc, sigma):
    rng = np.random.default_rng(100 + sigma)
    for n, lens in [(300, rng.poisson(3, 300)), (50, [0] * 20 + [900] + [1] * 29),
                    (2000, (rng.pareto(1.2, 2000) * 3).astype(int))]:
        rp, col = random_csr(rng, n, n, lens=lens)
        val = rng.standard_normal(int(rp[-1]))
        x = rng.uniform(-1, 1, n)
        ya, absr = orc.spmv_accurate(rp, col, val, x)
        ye = orc.merbit_emulate(rp, col, val, x, sigma=sigma)
        assert cond_err(ye, ya, absr) < 1e-13


def test_ref_ingest_hand_example(orc):
    # loaded COO (row -> col): 0->1, 0->2, 1->2, 2->0, 2->0 (duplicate kept)
    row = np.array([0, 0, 1, 2, 2], np.int32)
    col = np.array([1, 2, 2, 0, 0], np.int32)
    rp, c, v = orc.ref_ingest(3, row, col)
    # A^T rows are the loaded columns: col 0 <- rows 2,2 ; col 1 <- row 0 ;
    # col 2 <- rows 0,1.  values 1/outdeg(row): deg = [2, 1, 2]
    assert rp.tolist() == [0, 2, 3, 5]
    assert c.tolist() == [2, 2, 0, 0, 1]
    assert v.tolist() == [0.5, 0.5, 0.5, 0.5, 1.0]
    # P^T is column-stochastic: with x = 1, sum(y) = rows with outdeg > 0
    y = orc.spmv_csr(rp, c, v, np.ones(3))
    assert y.sum() == 3.0


def test_ref_ingest_vs_scipy(orc):
    rng = np.random.default_rng(21)
    n, m = 500, 4000
    row = rng.integers(0, n, m).astype(np.int32)
    col = rng.integers(0, n, m).astype(np.int32)
    rp, c, v = orc.ref_ingest(n, row, col)
    deg = np.bincount(row, minlength=n)
    A = sp.coo_matrix((1.0 / deg[row], (row, col)), shape=(n, n)).tocsr()
    At = A.T.tocsr()
    At.sum_duplicates()
    B = sp.csr_matrix((v, c, rp), shape=(n, n))
    B.sum_duplicates()
    assert np.allclose(B.toarray(), At.toarray(), rtol=1e-14, atol=0)
    for r in range(n):
        assert np.all(np.diff(c[rp[r]:rp[r + 1]]) >= 0)
    y = orc.spmv_csr(rp, c, v, np.ones(n))
    assert abs(y.sum() - np.count_nonzero(deg)) < 1e-9


def test_csr_transpose_vs_scipy(orc):
    rng = np.random.default_rng(4)
    rp, col = random_csr(rng, 300, 200, 5.0)
    cp, ri, perm = orc.csr_transpose(rp, col, 200)
    val = rng.standard_normal(int(rp[-1]))
    A = sp.csr_matrix((val, col, rp), shape=(300, 200))
    At = sp.csr_matrix((val[perm], ri, cp), shape=(200, 300))
    assert np.array_equal(At.toarray(), A.T.toarray())


def test_omp_baseline_matches(orc):
    rng = np.random.default_rng(8)
    rp, col = random_csr(rng, 5000, 5000, 7.0)
    val = rng.standard_normal(int(rp[-1]))
    x = rng.uniform(-1, 1, 5000)
    t, used, y = orc.spmv_csr_omp_time(rp, col, val, x, nthreads=2, warmup=1, reps=3)
    assert t > 0 and used >= 1
    assert np.array_equal(y, orc.spmv_csr(rp, col, val, x))


def _dangling_graph(seed=5, n=300, m=900):
    rng = np.random.default_rng(seed)
    row = rng.integers(0, n, m).astype(np.int32)
    col = rng.integers(0, n, m).astype(np.int32)
    return n, row, col


def test_pagerank_fixed_point_dense_solve(orc):
    """The oracle's iterations converge to the fixed point of
    x = d P^T x + (d/n) 1 a^T x + (1-d) s/n 1 (a: dangling indicator, s the
    mass), solved densely from the loaded edges without the oracle."""
    n, row, col = _dangling_graph()
    deg = np.bincount(row, minlength=n).astype(float)
    dangling = deg == 0
    assert dangling.sum() > 0
    P = np.zeros((n, n))
    np.add.at(P, (row, col), 1.0 / deg[row])
    d = 0.85
    x0 = np.random.default_rng(6).uniform(0.5, 1.5, n)
    s = x0.sum()
    M = np.eye(n) - d * P.T - (d / n) * np.outer(np.ones(n), dangling.astype(float))
    x_star = np.linalg.solve(M, np.full(n, (1 - d) * s / n))
    rp, c, v = orc.ref_ingest(n, row, col)
    x = orc.pagerank(rp, c, v, d, 300, x0)
    assert np.max(np.abs(x - x_star) / np.abs(x_star)) < 1e-12
    assert abs(x_star.sum() - s) < 1e-12 * s


def test_pagerank_mass_and_no_damping(orc):
    """Mass is kept every iteration; with d = 0 one iteration gives the uniform
    vector s/n; with no dangling nodes and d = 1 the iteration is P^T x."""
    n, row, col = _dangling_graph(seed=9)
    rp, c, v = orc.ref_ingest(n, row, col)
    x0 = np.random.default_rng(2).uniform(0, 1, n)
    for it in (1, 2, 7):
        assert abs(orc.pagerank(rp, c, v, 0.85, it, x0).sum() - x0.sum()) < 1e-12 * x0.sum()
    assert np.allclose(orc.pagerank(rp, c, v, 0.0, 1, x0), x0.sum() / n, rtol=1e-15, atol=0)
    ring_r = np.arange(n, dtype=np.int32)
    ring_c = ((ring_r + 1) % n).astype(np.int32)
    rp2, c2, v2 = orc.ref_ingest(n, ring_r, ring_c)
    assert np.array_equal(orc.pagerank(rp2, c2, v2, 1.0, 1, x0), np.roll(x0, 1))
