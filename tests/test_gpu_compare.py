"""The rocSPARSE comparator computes the same product as the merge path
(sanity of what the benchmark times beside it)."""
from __future__ import annotations

import numpy as np
import pytest
import torch

from helpers import cond_err, random_csr

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("alg", ["adaptive", "lrb", "coo", "coo_atomic"])
def test_rocsparse_matches_oracle(orc, device, alg):
    from merbit_amd.compare import RocsparseSpmv
    print(f"[case] rocsparse {alg}", flush=True)
    rng = np.random.default_rng(3)
    lens = (rng.pareto(1.0, 20000) * 4).astype(int)
    rp, col = random_csr(rng, len(lens), 20000, lens=lens)
    val = rng.standard_normal(int(rp[-1]))
    x = rng.uniform(-1, 1, 20000)
    t = [torch.from_numpy(a).to(device) for a in (rp, col, val, x)]
    y = torch.empty(len(lens), dtype=torch.float64, device=device)
    op = RocsparseSpmv(*t, y, alg=alg)
    op()
    torch.cuda.synchronize()
    ya, absr = orc.spmv_accurate(rp, col, val, x)
    assert cond_err(y.cpu().numpy(), ya, absr) <= 1e-12
    op.close()
