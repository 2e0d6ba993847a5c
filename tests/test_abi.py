"""C-ABI library checks that need no GPU: it loads, exports every symbol the
headers declare, carries gfx950 code, and rejects bad arguments on the host
before touching the device."""
from __future__ import annotations

import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in ("merbit_amd.h", "merbit_gen.h", "merbit_host.h"):
        text = open(os.path.join(ROOT, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names |= set(re.findall(r"\b(mb[xgh]_[a-z0-9_]+)\s*\(", text))
    return names


def test_library_exports_every_declared_symbol():
    from merbit_amd import _lib
    L = _lib.lib()
    decl = declared_functions()
    assert decl == set(_lib.ABI_SYMBOLS) | set(_lib.GEN_SYMBOLS) | set(_lib.HOST_SYMBOLS)
    for name in sorted(decl):
        assert hasattr(L, name), name


def test_library_carries_gfx950():
    from merbit_amd.build import LIB, _fatbin_targets
    assert "hipv4-amdgcn-amd-amdhsa--gfx950" in _fatbin_targets(LIB)


def test_host_side_argument_checks():
    from merbit_amd import _lib
    L = _lib.lib()
    h = ctypes.c_void_p()
    assert L.mbx_plan_create(None, 0, 0, 0, 8, 256, None, None) == -1
    assert L.mbx_plan_create(None, -1, 0, 0, 8, 256, None, ctypes.byref(h)) == -1
    assert L.mbx_plan_create(None, 5, 5, 0, 8, 256, None, ctypes.byref(h)) == -1
    assert L.mbx_plan_create(None, 0, 0, 0, 7, 256, None, ctypes.byref(h)) == -3
    assert L.mbx_plan_create(None, 0, 0, 0, 8, 100, None, ctypes.byref(h)) == -3
    big = ctypes.c_void_p(16)  # never dereferenced: the size check comes first
    assert L.mbx_plan_create(big, 2**30, 2**30, 2**30, 8, 256, None, ctypes.byref(h)) == -2
    assert L.mbx_spmv_f64(None, None, None, None, 1.0, None, 0.0, None, None) == -1
    assert L.mbx_spmv_f32(None, None, None, None, 1.0, None, 0.0, None, None) == -1
    assert L.mbx_spmv_bf16(None, None, None, None, 1.0, None, 0.0, None, None) == -1
    assert L.mbx_plan_info(None, None, None, None) == -1
    assert L.mbx_pagerank_f64(None, None, None, None, 0.85, 20, None, None, None, 0, None) == -1
    assert L.mbx_pagerank_workspace_bytes(-1) == 0
    assert L.mbx_pagerank_workspace_bytes(1000) >= 8 * 1000
    assert L.mbx_ingest_ref(-1, 0, None, None, 0, None, None, None, None) == -1
    h = ctypes.c_void_p()
    assert L.mbx_colblock_create(None, None, 4, 4, 0, 0, 8, None, ctypes.byref(h)) == -1
    assert L.mbx_colblock_create(1, 1, 4, 4, 8, 17, 8, None, ctypes.byref(h)) == -1
    assert L.mbx_colblock_create(1, 1, 4, 0, 8, 0, 8, None, ctypes.byref(h)) == -1
    assert L.mbx_colblock_create(1, 1, 4, 4, 8, 0, 8, None, None) == -1
    assert L.mbx_colblock_info(None, None, None, None) == -1
    assert L.mbx_colblock_values(None, 0, None, None, None) == -1
    assert L.mbx_colblock_spmv(None, 0, None, 1.0, None, 0.0, None, None) == -1
    assert L.mbx_colblock_sddmm(None, 0, None, None, None, None) == -1
    assert L.mbx_colblock_values_to_csr(None, 0, None, None, None) == -1
    assert L.mbx_plan_compress_cols(None, None, None) == -1
    assert L.mbx_plan_set_accumulation(None, 0) == -1
    assert L.mbx_adjoint_set_blocks(None, 0, None) == -1
    assert L.mbx_adjoint_auto(None, 0, None, None) == -1
    assert L.mbx_colblock_set_stream(None, 0) == -1
    assert L.mbx_colblock_set_residency(None, 0) == -1
    assert L.mbx_colblock_adjoint_prepare(None, None, None, 0, 8, None, None) == -1
    assert L.mbx_colblock_dx(None, 0, None, 1.0, None, 0.0, None, None) == -1
    rep = _lib.CsrReport()
    assert L.mbx_csr_validate(None, None, -1, 0, 0, None, ctypes.byref(rep)) == -1
    assert L.mbx_csr_validate(None, None, 4, 4, 0, None, ctypes.byref(rep)) == -1
    assert L.mbx_csr_validate(None, None, 0, 0, 0, None, ctypes.byref(rep)) == 0
    assert rep.nnz_mismatch == 0 and rep.first_bad_col == -1
    # no rows but entries: refused on the host with the mismatch reported
    assert L.mbx_csr_validate(None, ctypes.c_void_p(16), 0, 4, 3, None, ctypes.byref(rep)) == -1
    assert rep.nnz_mismatch == 1
    assert L.mbx_plan_create(None, 0, 0, 3, 8, 256, None, ctypes.byref(h)) == -1
    # the no-allocation product: bad sizes and the scratch limit, on the host
    assert L.mbx_spmv_f64_noalloc(None, None, None, -1, 0, 0, 1.0, None, 0.0, None, None) == -1
    assert L.mbx_spmv_f64_noalloc(None, None, None, 100, 100, 268433408, 1.0, None, 0.0,
                                  None, None) == -6
    assert L.mbx_spmv_f64_noalloc(None, None, None, 5, 5, 0, 1.0, None, 0.0, None, None) == -1
    assert "scratch" in _lib.status_str(-6)
    assert _lib.status_str(0) == "ok"
    assert "overflow" not in _lib.status_str(-1)
    assert "int32" in _lib.status_str(-2)


def test_generators_laplacian_matches_scipy():
    import scipy.sparse as sp
    from merbit_amd import gen
    g = 9
    A = gen.laplacian5(g)
    I = sp.identity(g)
    T = sp.diags([-1, 2, -1], [-1, 0, 1], shape=(g, g))
    B = (sp.kron(I, T) + sp.kron(T, I)).tocsr()
    B.sort_indices()
    assert np.array_equal(A.row_ptr, B.indptr) and np.array_equal(A.col, B.indices)
    assert np.array_equal(A.val, B.data)
    assert A.nnz == 5 * g * g - 4 * g


def test_generators_rmat_properties():
    from merbit_amd import gen
    A = gen.rmat(12, 16 << 12, 0.57, 0.19, 0.19, seed=1)
    lens = np.diff(A.row_ptr)
    assert A.nnz == 16 << 12 and A.nrow == 1 << 12
    assert all(np.all(np.diff(A.col[A.row_ptr[r]:A.row_ptr[r + 1]]) >= 0)
               for r in range(A.nrow))
    # row 0 takes the top-left quadrant (a+b = 0.76) at every level
    assert abs(lens[0] / A.nnz - 0.76 ** 12) < 0.3 * 0.76 ** 12
    assert 0.2 < np.mean(lens == 0) < 0.6
    B = gen.rmat(12, 16 << 12, 0.57, 0.19, 0.19, seed=1)
    assert np.array_equal(A.col, B.col)


def test_generators_banded_and_poisson():
    from merbit_amd import gen
    A = gen.banded(1000, 13)
    assert A.nnz == 27 * 1000 - 13 * 14
    r = 500
    assert A.col[A.row_ptr[r]:A.row_ptr[r + 1]].tolist() == list(range(487, 514))
    P = gen.poisson_rows(20000, 20000, 25.0, seed=3)
    lens = np.diff(P.row_ptr)
    assert abs(lens.mean() - 25.0) < 0.2 and abs(lens.var() - 25.0) < 1.5
    for rr in range(0, 20000, 997):
        c = P.col[P.row_ptr[rr]:P.row_ptr[rr + 1]]
        assert np.all(np.diff(c) > 0)  # distinct and sorted


def test_bf16_round_to_nearest_even():
    from merbit_amd import gen
    rng = np.random.default_rng(0)
    f = np.concatenate([rng.standard_normal(20000).astype(np.float32),
                        np.array([1.0, 1.00390625, 1.01171875, -0.0, 3.0e38], np.float32)])
    h = gen.to_bf16(f)
    back = gen.bf16_to_f32(h).astype(np.float64)
    # the nearest bf16 neighbours of each value, found independently by ulp steps
    ulp = np.float64(2.0) ** (np.floor(np.log2(np.abs(f.astype(np.float64)) + 1e-300)) - 7)
    err = np.abs(back - f.astype(np.float64))
    assert np.all(err <= ulp / 2 + 1e-300)
    ties = err == ulp / 2
    assert np.all((h[ties] & 1) == 0)
    # above the largest bf16 plus half an ulp the value rounds to infinity
    assert gen.to_bf16(np.array([3.4e38, -3.4e38], np.float32)).tolist() == [0x7F80, 0xFF80]


def test_c_program_links_the_abi(tmp_path):
    """A C99 program includes the public headers, links libmerbit_amd.so and
    calls host-checked entry points (no device work): the boundary is usable
    from plain C."""
    import shutil
    import subprocess
    if shutil.which("gcc") is None:
        pytest.skip("no C compiler")
    src = tmp_path / "abi_check.c"
    src.write_text(r"""
#include <stdio.h>
#include <string.h>
#include "merbit_amd.h"
#include "merbit_gen.h"
#include "merbit_host.h"
int main(void) {
  mbx_plan* p = 0;
  if (mbx_plan_create(0, -1, 0, 0, 8, 256, 0, &p) != MBX_EINVAL) return 1;
  if (mbx_plan_create(0, 0, 0, 0, 7, 256, 0, &p) != MBX_EUNSUPPORTED) return 2;
  if (mbx_spmv_f64(0, 0, 0, 0, 1.0, 0, 0.0, 0, 0) != MBX_EINVAL) return 3;
  if (mbx_plan_set_stream(0, MBX_STREAM_CACHED) != MBX_EINVAL) return 4;
  if (mbx_pagerank_workspace_bytes(-1) != 0) return 5;
  if (strcmp(mbx_status_str(MBX_OK), "ok") != 0) return 6;
  printf("abi ok\n");
  return 0;
}
""")
    lib_dir = os.path.join(ROOT, "merbit_amd")
    exe = tmp_path / "abi_check"
    r = subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", f"-I{ROOT}/include", str(src),
                        f"-L{lib_dir}", "-lmerbit_amd", f"-Wl,-rpath,{lib_dir}", "-o", str(exe)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and "abi ok" in r.stdout, (r.returncode, r.stdout, r.stderr)


def test_values_to_csr_checks_out_before_any_launch():
    """ColumnBlockedCsr.values_to_csr refuses a short or mistyped ``out``
    before calling the library (host tensors: nothing reaches a device)."""
    import torch
    from merbit_amd.spmv import ColumnBlockedCsr
    cb = ColumnBlockedCsr.__new__(ColumnBlockedCsr)
    cb.nnz, cb.value_slots, cb._h = 10, 16, None
    vb = torch.zeros(16, dtype=torch.float64)
    with pytest.raises(ValueError):
        cb.values_to_csr(vb, out=torch.zeros(9, dtype=torch.float64))
    with pytest.raises(ValueError):
        cb.values_to_csr(vb, out=torch.zeros(10, dtype=torch.float32))


def test_reference_style_cpp_compiles(tmp_path):
    """Host code written against the reference's API compiles against the
    C++ mirror: GRAPH's COO fields and sort(bool) (include/class.cuh:62-81),
    Fun_Merbit<V,BLOCK,SIGMA>(env, para) returning void (include/tile.cuh:
    563-564), Fun_Check, FSTR.  Compile only (no device work)."""
    import shutil
    import subprocess
    if shutil.which("g++") is None:
        pytest.skip("no C++ compiler")
    src = tmp_path / "ref_style.cpp"
    src.write_text(r"""
#include "merbit.hpp"
int run(const std::string& config, int gid) {
  ENV env(0, config, "log.txt");
  GRAPH<double> graph(config, gid);
  graph.sort(true);
  const int* r = graph.row_ind;
  const int* c = graph.col_ind;
  double* v = graph.values;
  (void)r; (void)c; (void)v;
  PARAMETER<double> para(&graph, 0.85, 20, 20);
  para.graph->sort(false);
  Fun_Merbit<double, 256, 14>(env, para);
  static_assert(std::is_same<decltype(Fun_Merbit<float, 256, 14>(env, *(PARAMETER<float>*)0)),
                             void>::value, "Fun_Merbit returns void");
  std::vector<double> y(graph.nrow), x(graph.nrow);
  return Fun_Check<double>(x.data(), y.data(), graph.nrow);
}
int main() { return 0; }
""")
    rocm = "/opt/rocm"
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-Wall", "-D__HIP_PLATFORM_AMD__",
                        f"-I{ROOT}/include", f"-I{ROOT}/merbit_amd/host", f"-I{rocm}/include",
                        str(src)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
