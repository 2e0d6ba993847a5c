"""Benchmark of the merge-path tile SpMV on one MI355X.

Workload (default): FP64 merge SpMV y = A*x on the R-MAT power-law matrix
(2^20 rows, 2^24 entries with duplicates kept, a=.57 b=c=.19, values N(0,1),
x ~ U[-1,1)).  A step is one SpMV (tile kernel + carry fixup) with the matrix
and x already resident in HBM.  Three copies of the matrix are rotated so
that consecutive steps cannot be served from the 256 MB MALL; the warm
(single copy) rate is reported beside it.

value = algorithmic GB/s = (4(nrow+1) + 12 nnz + 8 ncol + 8 nrow) / time per
step.  The roofline object is for the dominant (tile) kernel, timed with HIP
events recorded around it on its own stream.  The CPU baseline is the
row-parallel FP64 CSR product of the oracle, timed on the host cores.

Prints ONE JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def time_steps(calls, steps, warmup, plan_events=None):
    """Run warmup + steps calls; return (ms per step, mean tile-kernel ms)."""
    import torch
    for i in range(warmup):
        calls[i % len(calls)]()
    torch.cuda.synchronize()
    evs = None
    if plan_events is not None:
        evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
               for _ in range(steps)]
        for b, e in evs:  # materialise the underlying hipEvents
            b.record()
            e.record()
        torch.cuda.synchronize()
    start = torch.cuda.Event(enable_timing=True)
    stop = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    start.record()
    for i in range(steps):
        if evs is not None:
            plan_events(*evs[i])
        calls[i % len(calls)]()
    stop.record()
    torch.cuda.synchronize()
    ms = start.elapsed_time(stop) / steps
    kern = None
    if evs is not None:
        kern = float(np.mean([b.elapsed_time(e) for b, e in evs]))
    return ms, kern


def cpu_baseline(A, x, b_alg):
    from oracle import oracle
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
    threads = max(1, min(threads, os.cpu_count() or 1))
    t0 = time.time()
    t, used, _ = oracle.spmv_csr_omp_time(A.row_ptr, A.col, A.val, x, threads,
                                          warmup=2, reps=10)
    log(f"cpu baseline {t * 1e3:.2f} ms/SpMV on {used} threads ({time.time() - t0:.1f}s)")
    return {"value": round(b_alg / t / 1e9, 3), "unit": "GB/s", "cores": used,
            "kind": "port",
            "sample": "full R-MAT FP64 product, OpenMP row-parallel CSR, median of 10 "
                      "after 2 warm-ups", "ms_per_spmv": round(t * 1e3, 3)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--ipt", type=int, default=8, help="items per thread (SIGMA)")
    ap.add_argument("--copies", type=int, default=3)
    ap.add_argument("--scale", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-mixed", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not "
              "supported (the merge-path SpMV here runs on a single MI355X)",
              file=sys.stderr)
        sys.exit(2)

    import torch

    from merbit_amd import gen, workloads
    from merbit_amd.spmv import MerbitPlan

    dev = torch.device("cuda:0")
    t0 = time.time()
    A, x = workloads.rmat(args.scale)
    log(f"rmat: nrow={A.nrow} nnz={A.nnz} ({time.time() - t0:.1f}s)")
    nrow, ncol, nnz = A.nrow, A.ncol, A.nnz
    b_alg = workloads.algorithmic_bytes(nrow, ncol, nnz, 8, 8, 8)
    flops = 2 * nnz

    copies = []
    for _ in range(max(1, args.copies)):
        copies.append(dict(
            rp=torch.from_numpy(A.row_ptr).to(dev), col=torch.from_numpy(A.col).to(dev),
            val=torch.from_numpy(A.val).to(dev), x=torch.from_numpy(x).to(dev),
            y=torch.empty(nrow, dtype=torch.float64, device=dev)))
    plan = MerbitPlan(copies[0]["rp"], ncol, nnz, items_per_thread=args.ipt)
    torch.cuda.synchronize()

    def call_for(c):
        return lambda: plan.spmv(c["rp"], c["col"], c["val"], c["x"], c["y"])

    calls = [call_for(c) for c in copies]
    ms, kern_ms = time_steps(calls, args.steps, args.warmup, plan.set_events)
    warm_ms, warm_kern = time_steps(calls[:1], args.steps, args.warmup, plan.set_events)
    value = b_alg / (ms * 1e-3) / 1e9
    achieved = b_alg / (kern_ms * 1e-3) / 1e9
    log(f"fp64 rotated {ms * 1e3:.1f} us/step ({value:.0f} GB/s), tile kernel "
        f"{kern_ms * 1e3:.1f} us ({achieved:.0f} GB/s); warm {warm_ms * 1e3:.1f} us")
    y64 = copies[0]["y"].clone()

    extras = {
        "gflops": round(flops / (ms * 1e-3) / 1e9, 2),
        "pct_hbm_peak": round(100.0 * value / HBM_PEAK_GBS, 2),
        "algorithmic_bytes": b_alg,
        "tile_kernel_us": round(kern_ms * 1e3, 2),
        "warm_single_copy": {"gbps": round(b_alg / (warm_ms * 1e-3) / 1e9, 1),
                             "us_per_step": round(warm_ms * 1e3, 2),
                             "tile_kernel_us": round(warm_kern * 1e3, 2)},
        "ntiles": plan.ntiles, "tile_items": plan.tile_items,
    }

    if not args.no_mixed:
        mixed = {}
        v32 = A.val.astype(np.float32)
        x32 = x.astype(np.float32)
        for kind in ("f32", "bf16"):
            stored = v32 if kind == "f32" else gen.to_bf16(v32)
            s_val = 4 if kind == "f32" else 2
            wide = v32.astype(np.float64) if kind == "f32" else \
                gen.bf16_to_f32(stored).astype(np.float64)
            mc = []
            for _ in range(max(1, args.copies)):
                mc.append(dict(rp=copies[0]["rp"], col=torch.from_numpy(A.col).to(dev),
                               val=torch.from_numpy(stored).to(dev),
                               x=torch.from_numpy(x32).to(dev),
                               y=torch.empty(nrow, dtype=torch.float32, device=dev)))
            mcalls = [call_for(c) for c in mc]
            mms, mkern = time_steps(mcalls, args.steps, args.warmup, plan.set_events)
            mb = workloads.algorithmic_bytes(nrow, ncol, nnz, s_val, 4, 4)
            # error against the FP64 path (oracle-verified) on the same stored values
            ref = plan.spmv(copies[0]["rp"], copies[0]["col"],
                            torch.from_numpy(wide).to(dev),
                            torch.from_numpy(x32.astype(np.float64)).to(dev))
            ym = mc[0]["y"].double()
            rel = float(torch.linalg.norm(ym - ref) / torch.linalg.norm(ref))
            mixed[kind] = {"gbps": round(mb / (mms * 1e-3) / 1e9, 1),
                           "gflops": round(flops / (mms * 1e-3) / 1e9, 2),
                           "us_per_step": round(mms * 1e3, 2),
                           "tile_kernel_us": round(mkern * 1e3, 2),
                           "speedup_vs_f64": round(ms / mms, 3),
                           "rel_err_vs_f64_same_values": rel}
            log(f"{kind}: {mms * 1e3:.1f} us/step, x{ms / mms:.2f} vs fp64, err {rel:.2e}")
            del mc, mcalls
        extras["mixed_rmat"] = mixed

    out = {
        "metric": "SpMV GB/s as % HBM peak & GFLOP/s: FP64 plain vs mixed/batched at stated err",
        "value": round(value, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "fp64 merge SpMV, R-MAT power law 2^20 rows 2^24 nnz "
                               "(a=.57,b=c=.19), 3 rotated copies",
                   "nrow": nrow, "nnz": nnz, "items_per_thread": args.ipt,
                   "wg_size": 256, "copies_rotated": len(copies)},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": None},
        "cpu_baseline": None,
        **extras,
    }
    del y64
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(A, x, b_alg)
    plan.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
