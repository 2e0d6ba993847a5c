"""Device ingest (mbx_ingest_ref) against the hand-derived MTX fixtures and
the oracle's restatement of the reference's input preparation
(src/class.cu:9-34,143-177; include/tile.cuh:571-574): bit-exact integers
and values; invalid indices are an error, not a fault."""
from __future__ import annotations

import json
import os

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "mtx")


@pytest.mark.parametrize("name", ["general", "symmetric", "basezero"])
def test_device_ingest_fixtures(device, name):
    from merbit_amd import ingest
    print(f"[case] device ingest {name}", flush=True)
    exp = json.load(open(os.path.join(GOLDEN, "expected.json")))[name]
    n, ri, ci, _ = ingest.read_mtx(os.path.join(GOLDEN, f"{name}.mtx"))
    rp, col, val = ingest.device_ingest(n, torch.from_numpy(ri).to(device),
                                        torch.from_numpy(ci).to(device))
    assert rp.cpu().tolist() == exp["row_ptr_t"]
    assert col.cpu().tolist() == exp["col_t"]
    assert val.cpu().tolist() == exp["val_t"]


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_device_ingest_matches_oracle(orc, device, dtype):
    from merbit_amd import ingest
    rng = np.random.default_rng(21)
    for n, m in [(1, 3), (50, 0), (300, 5000), (4000, 60000), (70000, 400000)]:
        print(f"[case] device ingest n={n} m={m} {dtype}", flush=True)
        ri = rng.integers(0, n, m).astype(np.int32)
        ci = rng.integers(0, n, m).astype(np.int32)
        if m:  # a hub row and a hub column, with duplicate entries
            ri[: m // 10] = n // 2
            ci[m // 10: m // 5] = 0
        rp, col, val = ingest.device_ingest(n, torch.from_numpy(ri).to(device),
                                            torch.from_numpy(ci).to(device), dtype)
        erp, ecol, evl = orc.ref_ingest(n, ri, ci)
        assert np.array_equal(rp.cpu().numpy(), erp)
        assert np.array_equal(col.cpu().numpy(), ecol)
        want = evl if dtype == torch.float64 else evl.astype(np.float32)
        assert np.array_equal(val.cpu().numpy(), want)


def test_device_ingest_rejects_bad_index(device):
    from merbit_amd import MerbitError, ingest
    print("[case] device ingest bad index", flush=True)
    ri = torch.tensor([0, 1, 5], dtype=torch.int32, device=device)
    ci = torch.tensor([1, 0, 0], dtype=torch.int32, device=device)
    with pytest.raises(MerbitError):
        ingest.device_ingest(3, ri, ci)
    ci = torch.tensor([1, -1, 0], dtype=torch.int32, device=device)
    ri = torch.tensor([0, 1, 2], dtype=torch.int32, device=device)
    with pytest.raises(MerbitError):
        ingest.device_ingest(3, ri, ci)
