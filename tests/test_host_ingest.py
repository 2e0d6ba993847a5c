"""Reference-semantics ingest (include/merbit_host.h) on the CPU: MTX
fixtures with hand-derived expectations, agreement with the oracle's
restatement, and error paths (missing config ID, malformed files)."""
from __future__ import annotations

import json
import os

import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "mtx")


@pytest.mark.parametrize("name", ["general", "symmetric", "basezero"])
def test_mtx_fixtures(name):
    from merbit_amd import ingest
    exp = json.load(open(os.path.join(GOLDEN, "expected.json")))[name]
    n, ri, ci, _ = ingest.read_mtx(os.path.join(GOLDEN, f"{name}.mtx"))
    assert n == exp["nrow"]
    assert ri.tolist() == exp["row_ind"] and ci.tolist() == exp["col_ind"]
    A = ingest.ref_ingest(n, ri, ci)
    assert A.row_ptr.tolist() == exp["row_ptr_t"]
    assert A.col.tolist() == exp["col_t"]
    assert A.val.tolist() == exp["val_t"]


def test_ingest_matches_oracle(orc):
    from merbit_amd import ingest
    rng = np.random.default_rng(12)
    for n, m in [(1, 3), (50, 0), (300, 5000), (4000, 60000)]:
        ri = rng.integers(0, n, m).astype(np.int32)
        ci = rng.integers(0, n, m).astype(np.int32)
        A = ingest.ref_ingest(n, ri, ci)
        rp, col, val = orc.ref_ingest(n, ri, ci)
        assert np.array_equal(A.row_ptr, rp) and np.array_equal(A.col, col)
        assert np.array_equal(A.val, val)


def test_config_lookup(tmp_path):
    from merbit_amd import MerbitError, ingest
    cfg = tmp_path / "config.txt"
    cfg.write_text("1 /data/a.mtx\t10 10 20\n7 " + os.path.join(GOLDEN, "general.mtx") +
                   " 4 4 6\n")
    assert ingest.config_lookup(str(cfg), 1) == "/data/a.mtx"
    A = ingest.load_graph(str(cfg), 7)
    assert A.nrow == 4 and A.nnz == 6
    with pytest.raises(MerbitError):  # the reference loops forever here
        ingest.config_lookup(str(cfg), 3)
    with pytest.raises(MerbitError):
        ingest.config_lookup(str(tmp_path / "missing.txt"), 1)


def test_malformed_mtx(tmp_path):
    from merbit_amd import MerbitError, ingest
    bad = {
        "extra.mtx": "%%MatrixMarket matrix coordinate real general\n2 2 1\n1 1 1\n2 2 1\n",
        "short.mtx": "%%MatrixMarket matrix coordinate real general\n2 2 3\n1 1 1\n",
        "rect.mtx": "%%MatrixMarket matrix coordinate real general\n2 3 1\n1 1 1\n",
        "range.mtx": "%%MatrixMarket matrix coordinate real general\n2 2 1\n3 1 1\n",
        "nosize.mtx": "%%MatrixMarket matrix coordinate real general\n% only comments\n",
    }
    for name, text in bad.items():
        p = tmp_path / name
        p.write_text(text)
        with pytest.raises(MerbitError):
            ingest.read_mtx(str(p))
