"""The C++ config-ID driver (merbit_main <ID>, the reference's ./main <ID>)
on a registry of small MTX graphs, checked against a host FP64 product."""
from __future__ import annotations

import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "merbit_amd", "bin", "merbit_main")
GOLDEN = os.path.join(ROOT, "tests", "golden", "mtx")


def write_rmat_mtx(path, scale, seed):
    from merbit_amd import gen
    A = gen.rmat(scale, 8 << scale, seed=seed)
    rows = []
    for r in range(A.nrow):
        for k in range(A.row_ptr[r], A.row_ptr[r + 1]):
            rows.append(f"{r + 1} {A.col[k] + 1} 1\n")
    with open(path, "w") as f:
        f.write("%%MatrixMarket matrix coordinate real general\n")
        f.write(f"{A.nrow} {A.nrow} {A.nnz}\n")
        f.writelines(rows)


@pytest.mark.parametrize("typ,sigma", [("f64", 8), ("f64", 4), ("f64", 14), ("f32", 16), ("f32", 10)])
def test_driver_config_ids(device, tmp_path, typ, sigma):
    write_rmat_mtx(tmp_path / "rmat.mtx", 14, 3)
    cfg = tmp_path / "config.txt"
    cfg.write_text(f"1 {GOLDEN}/general.mtx 4 4 6\n2 {GOLDEN}/symmetric.mtx 3 3 3\n"
                   f"3 {tmp_path}/rmat.mtx\n")
    log = tmp_path / "log.txt"
    for gid in (1, 2, 3):
        print(f"[case] driver gid={gid} type={typ} sigma={sigma}", flush=True)
        r = subprocess.run([BIN, str(gid), "--config", str(cfg), "--log", str(log),
                            "--type", typ, "--sigma", str(sigma), "--perrnd", "5",
                            "--maxiter", "2", "--check"],
                           stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                           timeout=120)
        assert r.returncode == 0, r.stdout
        assert "check ok" in r.stdout
    lines = log.read_text().splitlines()
    assert len(lines) == 3 and all("&MERBIT&" in l for l in lines)
    assert lines[0].split("&")[0].strip() == "1"


def test_driver_missing_id_errors(device, tmp_path):
    cfg = tmp_path / "config.txt"
    cfg.write_text(f"1 {GOLDEN}/general.mtx\n")
    r = subprocess.run([BIN, "9", "--config", str(cfg)], stdout=subprocess.PIPE,
                       stderr=subprocess.STDOUT, text=True, timeout=60)
    assert r.returncode != 0 and "config ID 9" in r.stdout


def test_driver_coo_sort_round_trip(device, tmp_path):
    """--sort-rows puts the graph's COO entries in (row, col) order (the
    reference's GRAPH::sort(true)); Fun_Merbit sorts them back before it
    takes the CSR view, so the product still checks."""
    write_rmat_mtx(tmp_path / "rmat.mtx", 12, 5)
    cfg = tmp_path / "config.txt"
    cfg.write_text(f"7 {tmp_path}/rmat.mtx\n")
    for typ in ("f64", "f32"):
        print(f"[case] driver sort-rows type={typ}", flush=True)
        r = subprocess.run([BIN, "7", "--config", str(cfg), "--type", typ, "--sigma", "14",
                            "--perrnd", "2", "--maxiter", "2", "--sort-rows", "--check"],
                           stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                           timeout=120)
        assert r.returncode == 0 and "check ok" in r.stdout, r.stdout


@pytest.mark.parametrize("typ", ["f64", "f32"])
def test_driver_sigma_auto(device, tmp_path, typ):
    """--sigma auto: Fun_Merbit_Auto builds the plan with the launch rule
    (a power-law R-MAT gets the hot-column split, a banded-like stencil
    row-aligned tiles) and the product still checks against the host."""
    write_rmat_mtx(tmp_path / "rmat.mtx", 16, 7)
    n = 300
    with open(tmp_path / "band.mtx", "w") as f:
        ent = [(r, c) for r in range(n) for c in range(max(0, r - 2), min(n, r + 3))]
        f.write("%%MatrixMarket matrix coordinate real general\n")
        f.write(f"{n} {n} {len(ent)}\n")
        f.writelines(f"{r + 1} {c + 1} 1\n" for r, c in ent)
    cfg = tmp_path / "config.txt"
    cfg.write_text(f"1 {tmp_path}/rmat.mtx\n2 {tmp_path}/band.mtx\n")
    log = tmp_path / "log.txt"
    for gid in (1, 2):
        print(f"[case] driver auto gid={gid} type={typ}", flush=True)
        r = subprocess.run([BIN, str(gid), "--config", str(cfg), "--log", str(log), "--type", typ,
                            "--sigma", "auto", "--perrnd", "3", "--maxiter", "2", "--check"],
                           stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                           timeout=120)
        assert r.returncode == 0 and "check ok" in r.stdout, r.stdout
        assert "launch rule" in r.stdout and "sigma = auto:" in r.stdout, r.stdout
    lines = log.read_text().splitlines()
    assert len(lines) == 2 and all("&MERBIT&auto:" in l for l in lines)
    r = subprocess.run([BIN, "1", "--config", str(cfg), "--sigma", "auto", "--pagerank"],
                       stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=60)
    assert r.returncode == 2 and "fixed --sigma" in r.stdout
