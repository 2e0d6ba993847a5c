"""Build recipe for the MI355X (gfx950) library and the host driver.

``python -m merbit_amd.build`` (or ``__graft_entry__.build()``) compiles, in
tree and every time it is called:

* ``merbit_amd/libmerbit_amd.so`` -- HIP kernels + C ABI (include/merbit_amd.h)
  and the host-side generators (include/merbit_gen.h);
* ``merbit_amd/bin/merbit_main`` -- the C++ config-ID driver (``./main <ID>``
  of the reference, src/main.cu:3-84) over that library.
"""
from __future__ import annotations

import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "merbit_amd")
CSRC = os.path.join(PKG, "csrc")
HOST = os.path.join(PKG, "host")
LIB = os.path.join(PKG, "libmerbit_amd.so")
BIN = os.path.join(PKG, "bin", "merbit_main")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

# -packed-fp32-ops: no v_pk_{add,mul,fma}_f32 (build flags);
# -ffp-contract=off: the compensated (TwoSum/TwoProduct) arithmetic must be
# evaluated exactly as written.
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
]

LIB_SOURCES = [
    os.path.join(CSRC, "merbit_spmv.hip"),
    os.path.join(CSRC, "merbit_gen.cpp"),
]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    # hipcc's host pass does not know the packed-fp32 feature: harmless noise
    out = "\n".join(l for l in r.stdout.splitlines()
                    if "packed-fp32-ops" not in l and l.strip())
    if r.returncode != 0:
        raise RuntimeError(f"build failed: {' '.join(cmd)}\n{out}")
    if verbose and out:
        print(out)


def build_library(verbose: bool = False) -> str:
    inc = os.path.join(ROOT, "include")
    cmd = [HIPCC, *HIP_FLAGS, "-shared", f"-I{inc}", "-x", "hip",
           LIB_SOURCES[0], "-x", "c++", LIB_SOURCES[1], "-o", LIB + ".tmp"]
    _run(cmd, verbose)
    os.replace(LIB + ".tmp", LIB)
    return LIB


def build_driver(verbose: bool = False) -> str | None:
    main = os.path.join(HOST, "main.cpp")
    if not os.path.exists(main):
        return None
    os.makedirs(os.path.dirname(BIN), exist_ok=True)
    inc = os.path.join(ROOT, "include")
    cmd = [HIPCC, "-O2", "-std=c++17", f"-I{inc}", f"-I{HOST}",
           main, os.path.join(HOST, "graph.cpp"),
           f"-L{PKG}", "-lmerbit_amd", f"-Wl,-rpath,{PKG}", "-Wl,-rpath,$ORIGIN/..",
           "-o", BIN]
    _run(cmd, verbose)
    return BIN


def build(verbose: bool = False) -> None:
    build_library(verbose)
    build_driver(verbose)


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print(LIB)
