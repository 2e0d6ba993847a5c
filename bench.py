"""Benchmark of the merge-path tile SpMV on one MI355X.

Headline workload: FP64 merge SpMV y = A*x on the R-MAT power-law matrix
(2^20 rows, 2^24 entries with duplicates kept, a=.57 b=c=.19, values N(0,1),
x ~ U[-1,1)).  A step is one SpMV (hot gather, tile kernel and carry fixup)
with the matrix and x already resident in HBM; the step time is taken over
the whole loop of steps, with no per-step events, the steps captured as one
HIP graph and replayed once (--no-graph: eager launches).  Three copies of
the matrix, each with its own plan, are rotated so consecutive steps cannot
be served from the 256 MB MALL; the warm (single copy) rate is reported
beside it.  The launch shape (items per thread, stream cache policy, tiles
per CU, hot columns) is picked by timing the candidates on the rotated
workload, as a user would tune for their access pattern (--ipt / --stream /
--residency / --hot pin them); the launch rule's plans (the library's
default) are timed beside it.

value = algorithmic GB/s = (4(nrow+1) + 12 nnz + 8 ncol + 8 nrow) / time per
step.  The roofline object is for the dominant (tile) kernel, timed with HIP
events attached to its own dispatch (hipExtLaunchKernel) on the stream it is
launched on, the interval a kernel trace reports.  `traffic` (measured HBM bytes
per launch) is read from the committed counter summary for this workload.
`extras.ceilings` times the tile kernel's staging loads alone and its x
gathers alone (probes at the same shape).

Beside the headline, `extras` reports the other measured forms: the 256^2
Laplacian, mixed precision (FP32 / BF16 values) on the R-MAT and on a
banded 4M-row matrix (tuned, with FP64 sums and 16-bit columns, and the
launch rule's default), multi-RHS SpMM (k = 8, 32), 64 uneven matrices in
one persistent graph launch (autotuned), the adjoint on a 2.4M-node graph
(CSR, column-blocked and the operator's defaults) and on the R-MAT (the
operator's defaults against a pinned shape), PageRank iterations on the
R-MAT edge list after device ingest, rocSPARSE's CSR and COO SpMV on the
headline matrix (measured comparators), and a plain device copy as the
measured bandwidth ceiling.  Every rate uses the same algorithmic-bytes
definition.  The CPU leg times the oracle's OpenMP FP64 CSR product (the
CPU baseline) on the R-MAT, the Laplacian and the banded matrix, and checks
every GPU result gathered by the other legs against the oracle's accurate
product (`extras.oracle_errors`).

Prints ONE JSON line on stdout; progress goes to stderr.
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
METRIC = "SpMV GB/s as % HBM peak & GFLOP/s: FP64 plain vs mixed/batched at stated err"


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def time_steps(calls, steps, warmup, plan_events=None, graph=False):
    """Run warmup + steps calls round-robin; return (ms per step, mean
    tile-kernel ms or None).  The step time comes from events around the
    whole loop only.  graph: the `steps` calls are captured once as one HIP
    graph (as the reference's timed loop of repeated launches) and the timed
    region is its single replay.  With plan_events the loop is run once more,
    eagerly, with events attached to every tile kernel's dispatch (on the stream it
    is launched on), for the kernel's own duration."""
    import torch
    for i in range(warmup):
        calls[i % len(calls)]()
    torch.cuda.synchronize()
    g = None
    if graph:
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            for i in range(steps):
                calls[i % len(calls)]()
        torch.cuda.synchronize()
    start = torch.cuda.Event(enable_timing=True)
    stop = torch.cuda.Event(enable_timing=True)
    start.record()
    if g is not None:
        g.replay()
    else:
        for i in range(steps):
            calls[i % len(calls)]()
    stop.record()
    torch.cuda.synchronize()
    ms = start.elapsed_time(stop) / steps
    del g
    kern = None
    if plan_events is not None:
        evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
               for _ in range(steps)]
        for b, e in evs:  # materialise the underlying hipEvents
            b.record()
            e.record()
        torch.cuda.synchronize()
        for i in range(steps):
            plan_events(*evs[i])
            calls[i % len(calls)]()
        torch.cuda.synchronize()
        kern = float(np.mean([b.elapsed_time(e) for b, e in evs]))
    return ms, kern


def rel_err(a, b):
    import torch
    a, b = a.double(), b.double()
    return float(torch.linalg.norm(a - b) / torch.linalg.norm(b))


def gbs(nbytes, ms):
    return round(nbytes / (ms * 1e-3) / 1e9, 1)


def load_traffic(name):
    p = os.path.join(ROOT, "profiles", f"traffic_{name}.json")
    if not os.path.exists(p):
        return None
    with open(p) as f:
        return json.load(f).get("hbm_bytes_per_launch")


# ---------------------------------------------------------------------------


HOT_COLUMNS = {"f64": 4096, "f32": 5800, "bf16": 5800}
# FP64 at 12 items per thread runs 2 workgroups per CU either way: their LDS
# holds 6900 hot values (94.2 us against 96.1 with 4096 on the R-MAT)
HOT_COLUMNS_IPT12 = {"f64": 6900}


def make_plans(copies, ncol, nnz, ipt, hot=0, kind="f64", srt=False):
    """One plan per rotated copy (each keeps its own hot-column or sorted
    copy of the columns when hot > 0 or srt, so no array is shared across
    the rotation)."""
    from merbit_amd.spmv import MerbitPlan
    plans = []
    for c in copies:
        p = MerbitPlan(c["rp"], ncol, nnz, items_per_thread=ipt)
        if hot:
            p.split_hot(c["col"], hot)
        if srt and not p.sort_cols(c["col"]):
            raise RuntimeError("tile-sorted columns refused for this plan")
        plans.append(p)
    return plans


def rule_plans(copies, ncol, nnz, value_dtype):
    """The launch rule's plans (MerbitPlan's default), one per copy."""
    from merbit_amd.spmv import MerbitPlan
    return [MerbitPlan(c["rp"], ncol, nnz, col=c["col"], value_dtype=value_dtype)
            for c in copies]


def event_setter(plans):
    """For time_steps' event pass: the i-th call's plan attaches the events
    to its tile kernel (the calls cycle over the plans in order)."""
    state = {"i": 0}

    def setter(b, e):
        plans[state["i"] % len(plans)].set_events(b, e)
        state["i"] += 1
    return setter


def spmv_calls(plans, copies):
    return [(lambda p=p, c=c: p.spmv(c["rp"], c["col"], c["val"], c["x"], c["y"]))
            for p, c in zip(plans, copies)]


def tune_shape(args, copies, ncol, nnz, kind, hot_ok, sort_ok=True):
    """Launch shape: --ipt / --stream / --residency when given, else the
    (items per thread, stream policy, tiles per CU, hot columns, tile-sorted
    columns) whose plans run the timed workload (the rotated copies) fastest
    over a short timing of each (the first, in candidate order, within 1 %
    of the fastest).
    Tuning on the rotated copies matters: on one copy resident in the 256 MB
    cache the shapes time alike on the R-MAT matrix, while rotated copies
    separate them; a matrix read once per product gains from non-temporal
    stream loads that cost a cache-resident one 6 %; fewer tiles per CU cut
    the L1 misses of gather-bound products while HBM-bound ones want every
    tile resident; the hot-column split (items per thread 8 and 12) pays on
    power-law patterns, tile-sorted columns (8, 12 and 16, without the
    split) where a tile's gathers share x lines."""
    from merbit_amd.spmv import ITEMS_PER_THREAD
    ipts = [args.ipt] if args.ipt else [i for i in ITEMS_PER_THREAD if i in (4, 8, 12, 16)]
    modes = [args.stream] if args.stream else ["cached", "nontemporal"]
    wgs = [args.residency] if args.residency is not None else [0, 3, 4, 5]
    hots = ([args.hot] if args.hot is not None else
            [0, HOT_COLUMNS[kind]] + ([HOT_COLUMNS_IPT12[kind]] if kind in HOT_COLUMNS_IPT12
                                      else [])) if hot_ok else [0]

    def hot_fits(i, h):  # hot kernels exist for 8 and 12 items per thread
        return not h or (i in (8, 12) and (h != HOT_COLUMNS_IPT12.get(kind) or i == 12
                                           or args.hot is not None))
    sorts = [False] + ([True] if sort_ok and not args.no_sort else [])
    forms = [(h, False) for h in hots] + [(0, True) for s_ in sorts if s_]
    if args.sort and sort_ok:  # only the tile-sorted columns
        forms = [(0, True)]

    def form_fits(i, h, srt):
        return (i in (8, 12, 16) and not h) if srt else hot_fits(i, h)
    cands = [(i, m, w, h, srt) for i in ipts for h, srt in forms if form_fits(i, h, srt)
             for m in modes for w in (wgs if not h else [0])]
    if len(cands) == 1:
        return cands[0]
    times = {}
    for ipt in ipts:
        for hot, srt in forms:
            if not form_fits(ipt, hot, srt):
                continue
            plans = make_plans(copies, ncol, nnz, ipt, hot, kind, srt)
            calls = spmv_calls(plans, copies)
            for mode in modes:
                for p in plans:
                    p.set_stream(mode)
                for wg in (wgs if not hot else [0]):
                    for p in plans:
                        p.set_residency(wg)
                    times[(ipt, mode, wg, hot, srt)] = min(
                        time_steps(calls, 10 * len(calls), 2 * len(calls))[0] for _ in range(3))
            for p in plans:
                p.close()
    fastest = min(times.values())
    return next(k for k in cands if k in times and times[k] <= 1.01 * fastest)


def bench_rmat_f64(A, x, dev, args, checks):
    import torch
    from merbit_amd import workloads
    from merbit_amd.spmv import MerbitPlan
    nrow, ncol, nnz = A.nrow, A.ncol, A.nnz
    b_alg = workloads.algorithmic_bytes(nrow, ncol, nnz, 8, 8, 8)
    copies = [dict(rp=torch.from_numpy(A.row_ptr).to(dev), col=torch.from_numpy(A.col).to(dev),
                   val=torch.from_numpy(A.val).to(dev), x=torch.from_numpy(x).to(dev),
                   y=torch.empty(nrow, dtype=torch.float64, device=dev))
              for _ in range(max(1, args.copies))]
    c0 = copies[0]
    ipt, stream, resid, hot, srt = tune_shape(args, copies, ncol, nnz, "f64", not args.no_hot)
    plan_ms = []
    for _ in range(5):  # plan construction (the reference's tim1), host clock
        t0 = time.perf_counter()
        plan = MerbitPlan(c0["rp"], ncol, nnz, items_per_thread=ipt)
        plan_ms.append((time.perf_counter() - t0) * 1e3)
        plan.close()
    plans = make_plans(copies, ncol, nnz, ipt, hot, srt=srt)
    for p in plans:
        p.set_stream(stream)
        p.set_residency(resid)
    calls = spmv_calls(plans, copies)
    # the tile kernel's own duration: events the library attaches to its
    # dispatch on the call's stream, for whichever copy's plan runs
    ms, kern = time_steps(calls, args.steps, args.warmup, event_setter(plans),
                          graph=not args.no_graph)
    checks.append(("headline_rmat_f64", A, A.val, x, c0["y"].cpu().numpy(), 1e-12))
    warm_ms = warm_kern = None
    if not args.no_warm:  # one copy over and over: cache-resident, plain loads
        plans[0].set_stream("cached")
        warm_ms, warm_kern = time_steps(calls[:1], args.steps, args.warmup, plans[0].set_events)
    ceilings = None
    if not args.no_probes and ipt in (8, 12, 16):
        ceilings = bench_ceilings(A, copies, ncol, nnz, ipt, stream, resid, args, srt)
    # the launch rule's plans (the library default), on the same rotation,
    # and on one copy over and over (the rule's own use: a matrix reused)
    auto_ms = []
    for _ in range(3):  # rule + hot split + plan, host clock
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        MerbitPlan(c0["rp"], ncol, nnz, col=c0["col"], value_dtype=torch.float64).close()
        auto_ms.append((time.perf_counter() - t0) * 1e3)
    rp_rule = rule_plans(copies, ncol, nnz, torch.float64)
    rule_ms, _ = time_steps(spmv_calls(rp_rule, copies), args.steps, args.warmup,
                            graph=not args.no_graph)
    rule_warm_ms = None
    if not args.no_warm:
        rule_warm_ms, _ = time_steps(spmv_calls(rp_rule, copies)[:1], args.steps, args.warmup,
                                     graph=not args.no_graph)
    rule = dict(rp_rule[0].choice)
    for p in rp_rule:
        p.close()
    log(f"rmat fp64 (ipt {ipt}, {stream} stream, {resid or 'all'} tiles/CU, hot {hot}"
        f"{', sorted columns' if srt else ''}): "
        f"{ms * 1e3:.1f} us/step rotated ({gbs(b_alg, ms)} GB/s), tile kernel "
        f"{kern * 1e3:.1f} us; rule ({rule['rule']}) {rule_ms * 1e3:.1f} us"
        + (f"; warm {warm_ms * 1e3:.1f} us, rule's warm {rule_warm_ms * 1e3:.1f} us"
           if warm_ms else "") + f"; rule's plan creation {np.median(auto_ms):.2f} ms")
    res = dict(b_alg=b_alg, ms=ms, kern=kern, warm_ms=warm_ms, warm_kern=warm_kern,
               ntiles=plans[0].ntiles, tile_items=plans[0].tile_items, ipt=ipt, stream=stream,
               resid=resid, hot=hot, srt=srt, plan_ms=float(np.median(plan_ms)),
               ceilings=ceilings,
               rule={"us_per_step": round(rule_ms * 1e3, 2), "gbs": gbs(b_alg, rule_ms),
                     "warm_us_per_step": (round(rule_warm_ms * 1e3, 2) if rule_warm_ms
                                          else None),
                     "plan_create_ms": round(float(np.median(auto_ms)), 3),
                     "choice": {k: (round(v, 4) if isinstance(v, float) else v)
                                for k, v in rule.items()}})
    for p in plans:
        p.close()
    del copies
    return res


def bench_ceilings(A, copies, ncol, nnz, ipt, stream, resid, args, srt=False):
    """The headline tile kernel's two halves on their own (probes.hip, at
    the same tile shape, residency and stream policy, rotated copies): the
    stream probe reads row_ptr, col and the values as the tile kernel stages
    them and writes y; the gather probe reads col (the plan's tile-sorted
    columns when srt) and gathers x.  Beside them the product of the same
    plans (no hot split)."""
    import torch
    from merbit_amd.compare import probe
    plans = make_plans(copies, ncol, nnz, ipt, 0, srt=srt)
    for p in plans:
        p.set_stream(stream)
        p.set_residency(resid)
    outs = [torch.empty(max(A.nrow, plans[0].ntiles), dtype=torch.float64,
                        device=copies[0]["x"].device) for _ in copies]
    res = {}
    for which in ("stream", "gather"):
        calls = [(lambda p=p, c=c, o=o: probe(p, which, c["rp"], c["col"], c["val"], c["x"], o))
                 for p, c, o in zip(plans, copies, outs)]
        res[which] = time_steps(calls, args.steps, args.warmup, graph=not args.no_graph)[0]
    res["plain"] = time_steps(spmv_calls(plans, copies), args.steps, args.warmup,
                              graph=not args.no_graph)[0]
    for p in plans:
        p.close()
    b_stream = 4 * (A.nrow + 1) + 12 * nnz + 8 * A.nrow
    b_gather = 4 * nnz + 8 * ncol
    out = {"stream_only_us": round(res["stream"] * 1e3, 2),
           "stream_only_gbs": gbs(b_stream, res["stream"]),
           "gather_only_us": round(res["gather"] * 1e3, 2),
           "gather_only_gathers_per_s": round(nnz / (res["gather"] * 1e-3) / 1e9, 1),
           "plain_product_us": round(res["plain"] * 1e3, 2),
           "stream_plus_gather_us": round((res["stream"] + res["gather"]) * 1e3, 2),
           "shape": {"items_per_thread": ipt, "stream": stream, "tiles_per_cu": resid,
                     "sorted_columns": srt},
           "what": "probes.hip: the tile kernel's staging loads alone (stream) and its column "
                   "loads plus x gathers alone (gather), G gathers/s"}
    log(f"ceilings: stream only {out['stream_only_us']} us ({out['stream_only_gbs']} GB/s), "
        f"gather only {out['gather_only_us']} us, plain product {out['plain_product_us']} us")
    return out


def bench_mixed(name, A, x, dev, args, checks, f64_ms=None, copies=1):
    """FP64 vs FP32 vs BF16 values on one pattern: the tuned launch (with
    the hot split as a candidate on the power-law pattern), the same plans
    with FP64 sums and (banded) 16-bit columns, and the launch rule's plans
    (the library default).  Outputs go to `checks` for the oracle leg."""
    import torch
    from merbit_amd import gen, workloads
    from merbit_amd.spmv import ACCUMULATION_NAMES
    nrow, ncol, nnz = A.nrow, A.ncol, A.nnz
    rp = torch.from_numpy(A.row_ptr).to(dev)
    v32 = np.ascontiguousarray(A.val, dtype=np.float32)
    x32 = np.ascontiguousarray(x, dtype=np.float32)
    out = {}

    def make_copies(stored, xs, ydt):
        xx = torch.from_numpy(xs).to(dev)
        return [dict(rp=rp, col=torch.from_numpy(A.col).to(dev),
                     val=torch.from_numpy(stored).to(dev), x=xx,
                     y=torch.empty(nrow, dtype=ydt, device=dev)) for _ in range(copies)]

    def timed(plans, cps):
        return time_steps(spmv_calls(plans, cps), args.steps, args.warmup, event_setter(plans))

    def compressed(plans, cps, ref_ms):
        """The same products with 16-bit column offsets (banded only: its
        tiles span a few hundred columns); None when a plan refuses."""
        if name != "banded" or not all(p.compress_cols(c["col"]) for p, c in zip(plans, cps)):
            return None
        ms_c, kern_c = timed(plans, cps)
        for p in plans:
            p.compress_cols(None)
        return {"us_per_step": round(ms_c * 1e3, 2),
                "tile_kernel_us": None if kern_c is None else round(kern_c * 1e3, 2),
                "speedup_vs_f64": round(ref_ms / ms_c, 3), "column_bytes_per_nnz": 2}

    kinds = (["f64"] if f64_ms is None else []) + ["f32", "bf16"]
    for kind in kinds:
        if kind == "f64":  # FP64 values = the FP32 values widened
            stored, xs, s_val, s_x, ydt = v32.astype(np.float64), x32.astype(np.float64), 8, 8, \
                torch.float64
        else:
            stored = v32 if kind == "f32" else gen.to_bf16(v32)
            xs, s_val, s_x, ydt = x32, (4 if kind == "f32" else 2), 4, torch.float32
        cps = make_copies(stored, xs, ydt)
        b = workloads.algorithmic_bytes(nrow, ncol, nnz, s_val, s_x, s_x)
        ipt, stream, resid, hot, srt = tune_shape(args, cps, ncol, nnz, kind,
                                                  name == "rmat" and not args.no_hot,
                                                  name == "rmat")
        plans = make_plans(cps, ncol, nnz, ipt, hot, kind, srt)
        for p in plans:
            p.set_stream(stream)
            p.set_residency(resid)
        ms, kern = timed(plans, cps)
        if kind == "f64":
            f64_ms = ms
        tol = 1e-12 if kind == "f64" else 1e-6
        checks.append((f"{name}_{kind}", A, stored, xs, cps[0]["y"].cpu().numpy(), tol))
        rec = {"gbs": gbs(b, ms), "gflops": round(2 * nnz / (ms * 1e-3) / 1e9, 1),
               "us_per_step": round(ms * 1e3, 2),
               "tile_kernel_us": None if kern is None else round(kern * 1e3, 2),
               "items_per_thread": ipt, "stream": stream, "tiles_per_cu": resid,
               "hot_columns": hot, "sorted_columns": srt,
               "pct_hbm_peak": round(100 * b / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 1)}
        if kind != "f64":
            rec["speedup_vs_f64"] = round(f64_ms / ms, 3)
        c = compressed(plans, cps, f64_ms)
        if c:
            rec["compressed_cols"] = c
        if kind != "f64":  # the same plans with FP64 sums: exact or FP32-rounded products
            from merbit_amd.spmv import launch_rule
            vdt = torch.float32 if kind == "f32" else torch.bfloat16
            for acc, key in (("f64", "acc_f64"), ("f32prod", "acc_f32prod")):
                aplans, ahot = plans, hot
                if acc == "f32prod" and hot:
                    # 4-byte product slots leave LDS for a larger hot set at the
                    # same tiles per CU: the launch rule's count for this form
                    ahot = launch_rule(cps[0]["rp"], cps[0]["col"], ncol, vdt)["hot_columns"]
                    if ahot and ahot != hot:
                        aplans = make_plans(cps, ncol, nnz, ipt, ahot, kind)
                        for p in aplans:
                            p.set_stream(stream)
                            p.set_residency(resid)
                    else:
                        ahot = hot
                for p in aplans:
                    p.set_accumulation(acc)
                ms_a, kern_a = timed(aplans, cps)
                checks.append((f"{name}_{kind}_{'fp64_sums' if acc == 'f64' else 'f32_products'}",
                               A, stored, xs, cps[0]["y"].cpu().numpy(), tol))
                rec[key] = {"us_per_step": round(ms_a * 1e3, 2),
                            "tile_kernel_us": None if kern_a is None else round(kern_a * 1e3, 2),
                            "gbs": gbs(b, ms_a), "speedup_vs_f64": round(f64_ms / ms_a, 3),
                            "hot_columns": ahot}
                c = compressed(aplans, cps, f64_ms)
                if c:
                    rec[key]["compressed_cols"] = c
                if aplans is not plans:
                    for p in aplans:
                        p.close()
        for p in plans:
            p.close()
        # the launch rule's plans: what MerbitPlan does by default
        rps = rule_plans(cps, ncol, nnz, {"f64": torch.float64, "f32": torch.float32,
                                          "bf16": torch.bfloat16}[kind])
        ms_r, _ = timed(rps, cps)
        checks.append((f"{name}_{kind}_rule", A, stored, xs, cps[0]["y"].cpu().numpy(), tol))
        ch = rps[0].choice
        rec["rule"] = {"us_per_step": round(ms_r * 1e3, 2), "gbs": gbs(b, ms_r),
                       "speedup_vs_f64": round(f64_ms / ms_r, 3), "rule": ch["rule"],
                       "items_per_thread": ch["items_per_thread"],
                       "tiles_per_cu": ch["wg_per_cu"], "compress_cols": ch["compress_cols"],
                       "accumulation": ACCUMULATION_NAMES[ch["accumulation"]],
                       "hot_columns": ch["hot_columns"], "row_tiles": ch["row_tiles"]}
        for p in rps:
            p.close()
        out[kind] = rec
        log(f"{name} {kind}: tuned {ms * 1e3:.1f} us/step (ipt {ipt}, {stream}, "
            f"{resid or 'all'}/CU, hot {hot}{', sorted' if srt else ''})"
            + ("" if kind == "f64" else f" x{f64_ms / ms:.2f} vs fp64")
            + ("" if kind == "f64" else
               f"; FP32 products {rec['acc_f32prod']['us_per_step']:.1f} us "
               f"(hot {rec['acc_f32prod']['hot_columns']})")
            + f"; rule {ch['rule']} {ms_r * 1e3:.1f} us (x{f64_ms / ms_r:.2f})")
        del cps
    return out


def bench_spmm(A, dev, args, checks):
    import torch
    from merbit_amd import gen, workloads
    from merbit_amd.spmv import MerbitPlan
    nrow, ncol, nnz = A.nrow, A.ncol, A.nnz
    rp = torch.from_numpy(A.row_ptr).to(dev)
    col = torch.from_numpy(A.col).to(dev)
    v32 = A.val.astype(np.float32)
    val = torch.from_numpy(v32).to(dev)
    # 32 right-hand sides, column-major; k = 8 takes the first 8
    X32 = gen.uniform(ncol * 32, -1, 1, workloads.SEED["spmm_x"], np.float32)
    X32_dev = torch.from_numpy(X32).to(dev).view(32, ncol)
    out = {}
    for k in (8, 32):
        X = X32_dev[:k].t()  # col-major view, leading dimension ncol
        Y = torch.empty((k, nrow), dtype=torch.float32, device=dev).t()
        # items per thread: the faster of the two SpMM tile shapes on this k
        tuned = {}
        for ipt in (2, 4):
            plan = MerbitPlan(rp, ncol, nnz, items_per_thread=ipt)
            ws = torch.empty(plan.spmm_workspace_bytes(X), dtype=torch.uint8, device=dev)
            tuned[ipt] = min(time_steps([lambda: plan.spmm(rp, col, val, X, Y, workspace=ws)],
                                        10, 2)[0] for _ in range(2))
            plan.close()
        ipt = min(tuned, key=tuned.get)
        plan = MerbitPlan(rp, ncol, nnz, items_per_thread=ipt)
        ws = torch.empty(plan.spmm_workspace_bytes(X), dtype=torch.uint8, device=dev)
        ms, kern = time_steps([lambda: plan.spmm(rp, col, val, X, Y, workspace=ws)],
                              max(20, args.steps // 4), args.warmup, plan.set_events)
        b = workloads.algorithmic_bytes(nrow, ncol, nnz, 4, 4, 4, k=k)
        checks.append((f"spmm_k{k}", A, v32, X32.reshape(32, ncol)[:k],
                       Y.t().contiguous().cpu().numpy(), 1e-5))
        out[f"k{k}"] = {"gbs": gbs(b, ms), "gflops": round(2 * nnz * k / (ms * 1e-3) / 1e9, 1),
                        "us_per_step": round(ms * 1e3, 2),
                        "tile_kernel_us": round(kern * 1e3, 2),
                        "us_per_rhs": round(ms * 1e3 / k, 2),
                        "items_per_thread": ipt,
                        "layout": "column-major X and Y, repacked per call"}
        log(f"spmm k={k} (ipt {ipt}): {ms * 1e3:.1f} us/step ({out[f'k{k}']['gbs']} GB/s)")
        plan.close()
        del Y, ws
    return out


def bench_laplacian(dev, args, checks):
    """The 256 x 256 5-point Laplacian (65,536 rows): FP64 products on the
    launch rule's plan, the steps captured as one HIP graph (the product is
    launch-bound: a few microseconds)."""
    import torch
    from merbit_amd import workloads
    from merbit_amd.spmv import MerbitPlan
    A, x = workloads.laplacian(256)
    c = dict(rp=torch.from_numpy(A.row_ptr).to(dev), col=torch.from_numpy(A.col).to(dev),
             val=torch.from_numpy(A.val).to(dev), x=torch.from_numpy(x).to(dev),
             y=torch.empty(A.nrow, dtype=torch.float64, device=dev))
    plan = MerbitPlan(c["rp"], A.ncol, A.nnz, col=c["col"], value_dtype=torch.float64)
    calls = spmv_calls([plan], [c])
    ms, _ = time_steps(calls, 10 * args.steps, args.warmup, graph=not args.no_graph)
    checks.append(("laplacian_f64", A, A.val, x, c["y"].cpu().numpy(), 1e-12))
    b = workloads.algorithmic_bytes(A.nrow, A.ncol, A.nnz, 8, 8, 8)
    out = {"nrow": A.nrow, "nnz": A.nnz, "us_per_step": round(ms * 1e3, 3), "gbs": gbs(b, ms),
           "gflops": round(2 * A.nnz / (ms * 1e-3) / 1e9, 1), "rule": plan.choice["rule"],
           "items_per_thread": plan.items_per_thread, "row_tiles": plan.choice["row_tiles"],
           "timed_as": "eager launches" if args.no_graph else "one HIP graph of the steps"}
    log(f"laplacian 256^2 fp64: {ms * 1e3:.2f} us/step ({out['gbs']} GB/s)")
    plan.close()
    return out


def bench_batch(dev, args):
    import torch
    from merbit_amd import workloads
    from merbit_amd.batch import MerbitBatch, autotune_batch
    t0 = time.time()
    mats, xs = workloads.uneven_batch(args.batch_count)
    log(f"batch: {len(mats)} matrices, {sum(m.nrow for m in mats)} rows, "
        f"{sum(m.nnz for m in mats)} nnz ({time.time() - t0:.1f}s)")
    items = []
    for A, x in zip(mats, xs):
        items.append((torch.from_numpy(A.row_ptr).to(dev), torch.from_numpy(A.col).to(dev),
                      torch.from_numpy(A.val).to(dev), torch.from_numpy(x).to(dev),
                      torch.empty(A.nrow, dtype=torch.float32, device=dev), A.ncol))
    ipt, tuned_us = autotune_batch(items, reps=3)
    # the autotuned shape, and the members' tile-sorted columns at each shape
    # that has them; then tiles per CU, timed like the single-matrix shapes
    # (first within 1 %)
    forms = [(ipt, False)] + ([] if args.no_sort else [(i, True) for i in (8, 12, 16, 32)])
    wgs = [args.residency] if args.residency is not None else [0, 3, 4, 5]
    t_form = {}
    for fi, fs in forms:
        bb = MerbitBatch(items, items_per_thread=fi)
        if fs and not bb.sort_cols(True):
            bb.close()
            continue
        for wg in wgs:
            bb.set_residency(wg)
            t_form[(fi, fs, wg)] = min(time_steps([bb.spmv], 5, 2)[0] for _ in range(3))
        bb.close()
    best = min(t_form.values())
    ipt, srt, resid = next(k for k in t_form if t_form[k] <= 1.01 * best)
    b = MerbitBatch(items, items_per_thread=ipt)
    if srt:
        b.sort_cols(True)
    b.set_residency(resid)
    reps = 10
    g = b.graph(repeats=reps)
    ms, _ = time_steps([g.launch], max(5, args.steps // 20), 2)
    ms /= reps
    nbytes = sum(workloads.algorithmic_bytes(A.nrow, A.ncol, A.nnz, 4, 4, 4) for A in mats)
    nnz = sum(A.nnz for A in mats)
    # per-matrix launches for comparison (same arithmetic, one launch pair each)
    from merbit_amd.spmv import MerbitPlan
    # (single-matrix plans stop at 16 items per thread: the batch's 32 is its
    # 512-thread tile)
    plans = [MerbitPlan(it[0], it[5], it[1].numel(), items_per_thread=min(ipt, 16))
             for it in items]
    for p, it in zip(plans, items):
        p.set_residency(resid)
        if srt:
            p.sort_cols(it[1])
    calls = [lambda: [p.spmv(it[0], it[1], it[2], it[3], it[4]) for p, it in zip(plans, items)]]
    ms_loop, _ = time_steps(calls, max(5, args.steps // 20), 2)
    out = {"matrices": len(mats), "rows": int(sum(A.nrow for A in mats)), "nnz": int(nnz),
           "items_per_thread": ipt, "autotuned_items_per_thread": forms[0][0],
           "sorted_columns": srt, "tiles_per_cu": resid or "all",
           "tile_items": b.tile_items,
           "persistent_grid": b.grid, "gbs": gbs(nbytes, ms),
           "gflops": round(2 * nnz / (ms * 1e-3) / 1e9, 1), "us_per_batch": round(ms * 1e3, 1),
           "us_per_batch_separate_launches": round(ms_loop * 1e3, 1),
           "speedup_vs_separate_launches": round(ms_loop / ms, 2),
           "launch": f"one hipGraph of {reps} persistent batch products, per-product time"}
    log(f"batch: {ms * 1e3:.1f} us per batch product ({out['gbs']} GB/s; ipt {ipt}"
        f"{', sorted columns' if srt else ''}, {resid or 'all'}/CU), separate launches "
        f"{ms_loop * 1e3:.1f} us")
    for p in plans:
        p.close()
    g.close()
    b.close()
    return out


def bench_rocsparse(A, x, dev, args, f64_ms):
    """rocSPARSE SpMV (FP64; CSR adaptive and LRB, COO segmented and atomic)
    on the same R-MAT copies: a measured comparator, not part of the product
    path."""
    import torch
    from merbit_amd import workloads
    from merbit_amd.compare import RocsparseSpmv
    nrow, ncol, nnz = A.nrow, A.ncol, A.nnz
    b_alg = workloads.algorithmic_bytes(nrow, ncol, nnz, 8, 8, 8)
    copies = [dict(rp=torch.from_numpy(A.row_ptr).to(dev), col=torch.from_numpy(A.col).to(dev),
                   val=torch.from_numpy(A.val).to(dev), x=torch.from_numpy(x).to(dev),
                   y=torch.empty(nrow, dtype=torch.float64, device=dev))
              for _ in range(max(1, args.copies))]
    out = {}
    for alg in ("adaptive", "lrb", "coo", "coo_atomic"):
        try:
            t0 = time.time()
            ops = [RocsparseSpmv(c["rp"], c["col"], c["val"], c["x"], c["y"], alg)
                   for c in copies]
            prep_s = (time.time() - t0) / len(ops)
        except Exception as e:  # an algorithm this rocSPARSE build refuses
            out[alg] = {"error": str(e)}
            continue
        ms, _ = time_steps(ops, args.steps, args.warmup)
        out[alg] = {"gbs": gbs(b_alg, ms), "us_per_step": round(ms * 1e3, 2),
                    "merge_path_speedup": round(ms / f64_ms, 3),
                    "preprocess_ms": round(prep_s * 1e3, 1),
                    "buffer_bytes": ops[0].buffer_bytes}
        log(f"rocsparse {alg}: {ms * 1e3:.1f} us/step ({out[alg]['gbs']} GB/s), "
            f"merge path x{ms / f64_ms:.2f}")
        for o in ops:
            o.close()
    del copies
    return out


def bench_adjoint(dev, args, checks):
    """The differentiable product on a random graph (2.4M nodes, Poisson(25)
    out-degree, FP32 values): forward y = A x, dx = A^T g through the
    transposed merge path, dval = g[row] x[col] (SDDMM)."""
    import torch
    from merbit_amd import workloads
    from merbit_amd.autograd import MerbitOperator
    t0 = time.time()
    A, x, g = workloads.graph()
    log(f"graph: nrow={A.nrow} nnz={A.nnz} ({time.time() - t0:.1f}s)")
    rp = torch.from_numpy(A.row_ptr).to(dev)
    col = torch.from_numpy(A.col).to(dev)
    val = torch.from_numpy(A.val).to(dev)
    xx = torch.from_numpy(x).to(dev)
    gg = torch.from_numpy(g).to(dev)
    op = MerbitOperator(rp, col, A.ncol, items_per_thread=args.ipt or 8, blocks=0,
                        adjoint_blocks=0)
    t0 = time.time()
    op.adjoint_handle  # builds the transposed pattern once
    torch.cuda.synchronize()
    build_ms = (time.time() - t0) * 1e3
    n, nnz = A.nrow, A.nnz
    steps = max(20, args.steps // 4)
    out = {"nrow": n, "nnz": nnz, "transpose_build_ms": round(build_ms, 1)}
    fwd_ms, _ = time_steps([lambda: op.forward(val, xx)], steps, args.warmup)
    dx_ms, _ = time_steps([lambda: op.adjoint_x(val, gg)], steps, args.warmup)
    dv_ms, _ = time_steps([lambda: op.adjoint_val(gg, xx)], steps, args.warmup)
    b_fwd = workloads.algorithmic_bytes(n, A.ncol, nnz, 4, 4, 4)
    b_dx = workloads.algorithmic_bytes(A.ncol, n, nnz, 4, 4, 4)
    b_dv = 4 * (n + 1) + 4 * nnz + 4 * nnz + 4 * n + 4 * A.ncol
    for name, ms, b in (("forward", fwd_ms, b_fwd), ("dx", dx_ms, b_dx), ("dval", dv_ms, b_dv)):
        out[name] = {"us": round(ms * 1e3, 2), "gbs": gbs(b, ms),
                     "gflops": round(2 * nnz / (ms * 1e-3) / 1e9, 1)}
    out["backward_us"] = round((dx_ms + dv_ms) * 1e3, 2)
    checks.append(("graph_forward_f32", A, A.val, x, op.forward(val, xx).cpu().numpy(), 1e-6))
    log(f"adjoint: fwd {fwd_ms * 1e3:.1f} us, dx {dx_ms * 1e3:.1f} us, "
        f"dval {dv_ms * 1e3:.1f} us")
    # the same forward product on the column-blocked form of the pattern
    # (values moved into block-major order once; x slices stay in L2)
    from merbit_amd.spmv import ColumnBlockedCsr
    t0 = time.time()
    cb = ColumnBlockedCsr(rp, col, A.ncol)
    torch.cuda.synchronize()
    cb_build_ms = (time.time() - t0) * 1e3
    vb = cb.values(val)
    yb = torch.empty(n, dtype=torch.float32, device=dev)
    tuned = {}
    for srt in ((False,) if args.no_sort else (False, True)):
        if srt and not cb.sort_cols(True):
            break
        for mode in ("cached", "nontemporal"):
            cb.set_stream(mode)
            for wg in (0, 3, 4):
                cb.set_residency(wg)
                tuned[(mode, wg, srt)] = min(time_steps([lambda: cb.spmv(vb, xx, yb)], 10, 2)[0]
                                             for _ in range(2))
    cb_mode, cb_wg, cb_srt = min(tuned, key=tuned.get)
    cb.sort_cols(cb_srt)
    cb.set_stream(cb_mode)
    cb.set_residency(cb_wg)
    vals_ms, _ = time_steps([lambda: cb.values(val, out=vb)], steps, args.warmup)
    fwdb_ms, _ = time_steps([lambda: cb.spmv(vb, xx, yb)], steps, args.warmup)
    db = torch.zeros(max(cb.value_slots, 1), dtype=torch.float32, device=dev)
    dvb_ms, _ = time_steps([lambda: cb.sddmm(gg, xx, out=db)], steps, args.warmup)
    t0 = time.time()
    # column-blocked A^T (blocks of A's rows, g slices of at most 3.5 MiB)
    nbt = cb.prepare_adjoint(rp, col)
    torch.cuda.synchronize()
    adj_build_ms = (time.time() - t0) * 1e3
    cb.set_stream(cb_mode)
    cb.set_residency(cb_wg)
    dxb = torch.empty(A.ncol, dtype=torch.float32, device=dev)
    dxb_ms, _ = time_steps([lambda: cb.adjoint_x(vb, gg, dxb)], steps, args.warmup)
    dx_ref = op.adjoint_x(val, gg)
    # the CSR-order adjoint with a blocked A^T: values stay in CSR order
    opb = MerbitOperator(rp, col, A.ncol, items_per_thread=args.ipt or 8, blocks=0,
                         adjoint_blocks=-1)
    opb.adjoint_handle
    torch.cuda.synchronize()
    dxc_ms, _ = time_steps([lambda: opb.adjoint_x(val, gg)], steps, args.warmup)
    out["dx_blocked_transpose"] = {"us": round(dxc_ms * 1e3, 2), "gbs": gbs(b_dx, dxc_ms),
                                   "speedup": round(dx_ms / dxc_ms, 2),
                                   "rel_diff": rel_err(opb.adjoint_x(val, gg), dx_ref)}
    log(f"csr adjoint dx through a blocked A^T: {dxc_ms * 1e3:.1f} us (x{dx_ms / dxc_ms:.2f})")
    opb.close()
    # the CSR-order operator with every product on blocked forms
    # the operator's defaults: the launch rule picks its plan and, for this
    # spread-column graph, both blocked forms
    opc = MerbitOperator(rp, col, A.ncol)
    opc.adjoint_handle
    torch.cuda.synchronize()
    f_ms, _ = time_steps([lambda: opc.forward(val, xx)], steps, args.warmup)
    v_ms, _ = time_steps([lambda: opc.adjoint_val(gg, xx)], steps, args.warmup)
    x_ms, _ = time_steps([lambda: opc.adjoint_x(val, gg)], steps, args.warmup)
    out["csr_operator_blocked"] = {
        "forward_us": round(f_ms * 1e3, 2), "dval_us": round(v_ms * 1e3, 2),
        "dx_us": round(x_ms * 1e3, 2), "step_us": round((f_ms + v_ms + x_ms) * 1e3, 2),
        "step_speedup": round((fwd_ms + dv_ms + dx_ms) / (f_ms + v_ms + x_ms), 2),
        "values": "CSR order (moved into block order per forward call); dval by the CSR SDDMM",
        "operator": "MerbitOperator defaults (launch rule: spread columns -> blocked forms)",
        "blocked": opc._cb is not None}
    log(f"csr operator on blocked forms: fwd {f_ms * 1e3:.1f}, dval {v_ms * 1e3:.1f}, "
        f"dx {x_ms * 1e3:.1f} us (step x{(fwd_ms + dv_ms + dx_ms) / (f_ms + v_ms + x_ms):.2f})")
    opc.close()
    out["forward_column_blocked"] = {
        "us": round(fwdb_ms * 1e3, 2), "gbs": gbs(b_fwd, fwdb_ms),
        "gflops": round(2 * nnz / (fwdb_ms * 1e-3) / 1e9, 1), "blocks": cb.nblocks,
        "stream_loads": cb_mode, "tiles_per_cu": cb_wg, "sorted_columns": cb_srt,
        "speedup_vs_csr": round(fwd_ms / fwdb_ms, 2),
        "dval_blocked_us": round(dvb_ms * 1e3, 2), "dval_speedup": round(dv_ms / dvb_ms, 2),
        "dx_blocked_us": round(dxb_ms * 1e3, 2), "dx_speedup": round(dx_ms / dxb_ms, 2),
        "dx_blocks": nbt, "dx_build_ms": round(adj_build_ms, 1),
        "dx_rel_diff_vs_csr": rel_err(dxb, dx_ref),
        "step_us": round((fwdb_ms + dvb_ms + dxb_ms) * 1e3, 2),
        "step_speedup": round((fwd_ms + dv_ms + dx_ms) / (fwdb_ms + dvb_ms + dxb_ms), 2),
        "values_to_blocked_us": round(vals_ms * 1e3, 2), "build_ms": round(cb_build_ms, 1),
        "rel_diff_vs_csr": rel_err(yb, op.forward(val, xx))}
    log(f"graph forward, {cb.nblocks} column blocks: {fwdb_ms * 1e3:.1f} us "
        f"(x{fwd_ms / fwdb_ms:.2f} vs CSR), dval {dvb_ms * 1e3:.1f} us (x{dv_ms / dvb_ms:.2f}), "
        f"dx {dxb_ms * 1e3:.1f} us (x{dx_ms / dxb_ms:.2f}, {nbt} row blocks); "
        f"values to block order {vals_ms * 1e3:.1f} us, "
        f"build {cb_build_ms:.1f} ms)")
    cb.close()
    op.close()
    return out


def bench_adjoint_rmat(A, x, dev, args, checks):
    """The differentiable product on the headline R-MAT matrix (one copy, as
    a training step reuses its matrix): forward y = A x, dx = A^T g and
    dval = g[row] x[col] for the operator's defaults (the launch rule on A
    and, through mbx_adjoint_auto, on A^T: hot-column splits of both) and
    for a pinned operator (8 items per thread, single products, no split),
    FP64 and FP32 values."""
    import torch
    from merbit_amd.autograd import MerbitOperator
    from merbit_amd.gen import Csr
    rp = torch.from_numpy(A.row_ptr).to(dev)
    col = torch.from_numpy(A.col).to(dev)
    g = np.random.default_rng(41).uniform(-1.0, 1.0, A.nrow)
    steps = max(20, args.steps // 4)
    out = {}
    for name, dt in (("f64", torch.float64), ("f32", torch.float32)):
        vals = A.val if name == "f64" else A.val.astype(np.float32)
        gh = g if name == "f64" else g.astype(np.float32)
        xh = x if name == "f64" else x.astype(np.float32)
        val = torch.from_numpy(vals).to(dev)
        xx, gg = torch.from_numpy(xh).to(dev), torch.from_numpy(gh).to(dev)
        ops = {"defaults": MerbitOperator(rp, col, A.ncol, value_dtype=dt),
               "pinned": MerbitOperator(rp, col, A.ncol, items_per_thread=8, blocks=0,
                                        adjoint_blocks=0, value_dtype=dt)}
        res = {}
        for key, op in ops.items():
            op.adjoint_handle  # the transposed pattern, built once
            torch.cuda.synchronize()
            f_ms, _ = time_steps([lambda: op.forward(val, xx)], steps, args.warmup)
            x_ms, _ = time_steps([lambda: op.adjoint_x(val, gg)], steps, args.warmup)
            v_ms, _ = time_steps([lambda: op.adjoint_val(gg, xx)], steps, args.warmup)
            res[key] = {"forward_us": round(f_ms * 1e3, 2), "dx_us": round(x_ms * 1e3, 2),
                        "dval_us": round(v_ms * 1e3, 2),
                        "step_us": round((f_ms + x_ms + v_ms) * 1e3, 2)}
        d = ops["defaults"]
        keys = ("rule", "items_per_thread", "sort_cols", "hot_columns")
        res["defaults"]["forward_rule"] = {k: d.plan.choice[k] for k in keys}
        res["defaults"]["adjoint_rule"] = {k: d.adjoint_choice[k] for k in keys}
        res["step_speedup"] = round(res["pinned"]["step_us"] / res["defaults"]["step_us"], 3)
        res["dx_speedup"] = round(res["pinned"]["dx_us"] / res["defaults"]["dx_us"], 3)
        cp, ri, pm = d.transpose()  # A^T from the device's own transpose
        At = Csr(cp, ri, np.ascontiguousarray(vals[pm]), A.nrow)
        checks.append((f"rmat_dx_{name}", At, At.val, gh, d.adjoint_x(val, gg).cpu().numpy(),
                       1e-12 if name == "f64" else 1e-6))
        for op in ops.values():
            op.close()
        out[name] = res
        log(f"rmat adjoint {name}: defaults fwd {res['defaults']['forward_us']}, dx "
            f"{res['defaults']['dx_us']}, dval {res['defaults']['dval_us']} us; pinned "
            f"dx {res['pinned']['dx_us']} us (dx x{res['dx_speedup']}, step "
            f"x{res['step_speedup']})")
    return out


def bench_pagerank(A, dev, args):
    """The reference's pipeline on the R-MAT edge list: device ingest (CSR of
    P^T, values 1/out-degree), then 20 damped PageRank iterations captured
    as one HIP graph, against 20 plain products on the same matrix."""
    import ctypes

    import torch
    from merbit_amd._lib import check, lib
    from merbit_amd.ingest import device_ingest
    from merbit_amd.spmv import MerbitPlan
    n = A.nrow
    row = torch.repeat_interleave(torch.arange(n, dtype=torch.int32, device=dev),
                                  torch.from_numpy(np.diff(A.row_ptr)).to(dev))
    col = torch.from_numpy(A.col).to(dev)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    rp, ct, vt = device_ingest(n, row, col)
    torch.cuda.synchronize()
    ingest_ms = (time.perf_counter() - t0) * 1e3
    del row, col
    iters = 20
    # the launch rule's plan (the library default) unless a shape is pinned
    plan = MerbitPlan(rp, n, vt.numel(), items_per_thread=args.ipt) if args.ipt else \
        MerbitPlan(rp, n, vt.numel(), col=ct, value_dtype=torch.float64)
    x0 = torch.full((n,), 1.0 / n, dtype=torch.float64, device=dev)
    y = torch.empty(n, dtype=torch.float64, device=dev)
    nbytes = int(lib().mbx_pagerank_workspace_bytes(n))
    work = torch.empty(nbytes, dtype=torch.uint8, device=dev)
    g = ctypes.c_void_p()
    check(lib().mbx_pagerank_graph_create(plan.handle, rp.data_ptr(), ct.data_ptr(),
                                          vt.data_ptr(), 0.85, iters, x0.data_ptr(),
                                          y.data_ptr(), work.data_ptr(), nbytes,
                                          ctypes.byref(g)), "mbx_pagerank_graph_create")
    s = torch.cuda.current_stream().cuda_stream
    pr_ms, _ = time_steps([lambda: check(lib().mbx_graph_launch(g, s), "mbx_graph_launch")],
                          max(5, args.steps // 10), 2)
    lib().mbx_graph_destroy(g)
    spmv_ms, _ = time_steps([lambda: plan.spmv(rp, ct, vt, x0, y)], args.steps, args.warmup)
    mass = float(plan.pagerank(rp, ct, vt, damp=0.85, iters=iters, x0=x0).sum())
    ch = getattr(plan, "choice", None)
    shape = {"items_per_thread": plan.items_per_thread}
    if ch:
        shape.update(rule=ch["rule"], sorted_columns=bool(ch["sort_cols"]),
                     hot_columns=ch["hot_columns"])
    plan.close()
    out = {"nrow": n, "nnz": int(vt.numel()), "iterations": iters, "damp": 0.85, "plan": shape,
           "ingest_ms": round(ingest_ms, 2), "us_per_iteration": round(pr_ms * 1e3 / iters, 2),
           "us_per_plain_spmv": round(spmv_ms * 1e3, 2), "mass_after": mass}
    log(f"pagerank: ingest {ingest_ms:.1f} ms, {out['us_per_iteration']} us/iteration "
        f"(plain SpMV {out['us_per_plain_spmv']} us), mass {mass:.15f}")
    return out


def bench_copy(dev):
    import torch
    n = 256 << 20  # 256 Mi doubles = 2 GiB
    a = torch.empty(n, dtype=torch.float64, device=dev).fill_(1.0)
    b = torch.empty_like(a)
    ms, _ = time_steps([lambda: b.copy_(a)], 20, 3)
    return {"gbs": gbs(2 * 8 * n, ms), "bytes": 2 * 8 * n, "what": "torch device copy of 2 GiB"}


def _errors(y, ya, absr):
    """(max condition-scaled error |y - ya| / (|A||x|), normwise relative
    error); rows with (|A||x|) = 0 must be exactly 0 (else inf)."""
    y = np.asarray(y, np.float64)
    d = np.abs(y - ya)
    zero = absr == 0
    cond = float("inf") if np.any(d[zero] != 0) else \
        (float(np.max(d[~zero] / absr[~zero])) if np.any(~zero) else 0.0)
    den = float(np.linalg.norm(ya))
    return cond, float(np.linalg.norm(y - ya) / (den if den > 0 else 1.0))


def cpu_leg(A, x, b_alg, checks, laplacian, banded):
    """The host leg: the oracle's OpenMP FP64 CSR product timed on the host
    cores (the CPU baseline; the reference has no CPU SpMV) for the R-MAT,
    the Laplacian and the banded matrix, at every thread and at one, and the
    oracle as the checker of every GPU result gathered in `checks`: its
    accurate (Dot2) product on the same stored values."""
    from oracle import oracle
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
    threads = max(1, min(threads, os.cpu_count() or 1))
    t0 = time.time()
    t, used, _ = oracle.spmv_csr_omp_time(A.row_ptr, A.col, A.val, x, threads,
                                          warmup=3, reps=20)
    t1, _, _ = oracle.spmv_csr_omp_time(A.row_ptr, A.col, A.val, x, 1, warmup=3, reps=20)
    log(f"cpu baseline {t * 1e3:.2f} ms/SpMV on {used} threads, {t1 * 1e3:.2f} ms on 1 "
        f"({time.time() - t0:.1f}s)")
    base = {"value": round(b_alg / t / 1e9, 3), "unit": "GB/s", "cores": used,
            "kind": "port",
            "sample": "full R-MAT FP64 product (2^20 rows, 2^24 nnz), OpenMP row-parallel "
                      "CSR, median of 20 after 3 warm-ups", "ms_per_spmv": round(t * 1e3, 3),
            "one_thread": {"value": round(b_alg / t1 / 1e9, 3), "ms_per_spmv": round(t1 * 1e3, 3)}}
    from merbit_amd import workloads
    per = {"rmat_f64": {"threads": used, "ms": round(t * 1e3, 3), "gbs": base["value"],
                        "one_thread_ms": round(t1 * 1e3, 3),
                        "one_thread_gbs": base["one_thread"]["value"]}}
    for name, (M, v, xm, reps) in (("laplacian_f64", laplacian), ("banded_f64", banded)):
        if M is None:
            continue
        v64 = np.ascontiguousarray(v, dtype=np.float64)
        x64 = np.ascontiguousarray(xm, dtype=np.float64)
        b = workloads.algorithmic_bytes(M.nrow, M.ncol, M.nnz, 8, 8, 8)
        tn, un, _ = oracle.spmv_csr_omp_time(M.row_ptr, M.col, v64, x64, threads, warmup=2,
                                             reps=reps)
        to, _, _ = oracle.spmv_csr_omp_time(M.row_ptr, M.col, v64, x64, 1, warmup=1,
                                            reps=max(3, reps // 4))
        per[name] = {"threads": un, "ms": round(tn * 1e3, 3), "gbs": round(b / tn / 1e9, 3),
                     "one_thread_ms": round(to * 1e3, 3), "one_thread_gbs": round(b / to / 1e9, 3)}
        log(f"cpu baseline {name}: {tn * 1e3:.2f} ms on {un} threads, {to * 1e3:.2f} ms on 1")
    errs = {}
    t0 = time.time()
    for name, M, stored, xs, y, tol in checks:
        if xs.ndim == 2:  # right-hand sides by rows (k x ncol), Y by rows (k x nrow)
            worst_c = worst_n = 0.0
            for j in range(xs.shape[0]):
                ya, absr = oracle.spmv_accurate(M.row_ptr, M.col, stored,
                                                np.ascontiguousarray(xs[j]))
                c, n = _errors(y[j], ya, absr)
                worst_c, worst_n = max(worst_c, c), max(worst_n, n)
            errs[name] = {"cond_max": worst_c, "normwise_max": worst_n, "columns": xs.shape[0],
                          "bar": tol, "within_bar": worst_c <= tol}
        else:
            ya, absr = oracle.spmv_accurate(M.row_ptr, M.col, stored, xs)
            c, n = _errors(y, ya, absr)
            errs[name] = {"cond_max": c, "normwise": n, "bar": tol, "within_bar": c <= tol}
    log(f"oracle checks: {len(errs)} results, all within their bars: "
        f"{all(e['within_bar'] for e in errs.values())} ({time.time() - t0:.1f}s)")
    return base, per, errs


def north_star_summary(extras):
    """The ratios the targets are stated in, gathered from the records above:
    mixed and batched forms against the plain FP64 product of the same
    matrix (its tuned launch with 32-bit columns), with the library's
    default (the launch rule's plan) and the tuned launch, the largest
    oracle error of the mixed results, and the differentiable step on the
    column-blocked forms."""
    ns = {}
    errs = extras.get("oracle_errors", {})
    for name in ("banded", "rmat"):
        mb = extras.get(f"mixed_{name}")
        if not mb:
            continue
        m = {}
        for kind in ("f32", "bf16"):
            m[kind] = mb[kind]["rule"]["speedup_vs_f64"]
            m[f"{kind}_tuned_compensated"] = mb[kind]["speedup_vs_f64"]
            acc = mb[kind].get("acc_f64", {})
            if "compressed_cols" in acc:
                m[f"{kind}_tuned_fp64_sums_16bit_cols"] = acc["compressed_cols"]["speedup_vs_f64"]
            acc = mb[kind].get("acc_f32prod", {})
            if "compressed_cols" in acc:
                m[f"{kind}_tuned_f32_products_16bit_cols"] = \
                    acc["compressed_cols"]["speedup_vs_f64"]
            elif acc:
                m[f"{kind}_tuned_f32_products"] = acc["speedup_vs_f64"]
        e = [v["cond_max"] for k, v in errs.items()
             if k.startswith(f"{name}_f32") or k.startswith(f"{name}_bf16")]
        if e:
            m["max_cond_err_vs_fp64_oracle"] = max(e)
        ns[f"mixed_vs_fp64_{name}"] = m
    sp = extras.get("spmm_rmat")
    f64_us = extras.get("tile_kernel_us")
    if sp and f64_us:
        ns["spmm_per_rhs_vs_fp64_spmv"] = {k: round(f64_us / v["us_per_rhs"], 2)
                                           for k, v in sp.items()}
    bt = extras.get("batch_uneven")
    if bt:
        ns["batch_vs_separate_launches"] = bt["speedup_vs_separate_launches"]
    ag = extras.get("adjoint_graph", {})
    ad = ag.get("forward_column_blocked")
    if ad:
        ns["graph_forward_backward_colblocked_vs_csr"] = ad["step_speedup"]
    if "dx_blocked_transpose" in ag:
        ns["graph_dx_csr_values_blocked_transpose_vs_csr"] = ag["dx_blocked_transpose"]["speedup"]
    if "csr_operator_blocked" in ag:
        ns["graph_forward_backward_csr_values_blocked_vs_csr"] = \
            ag["csr_operator_blocked"]["step_speedup"]
    return ns


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--ipt", type=int, default=0,
                    help="items per thread (SIGMA); 0 = picked by the timing autotuner")
    ap.add_argument("--stream", choices=["cached", "nontemporal"], default=None,
                    help="stream cache policy; default: picked with the tile shape")
    ap.add_argument("--residency", type=int, default=None,
                    help="tiles per CU (0 = as many as fit); default: picked with the shape")
    ap.add_argument("--copies", type=int, default=3)
    ap.add_argument("--scale", type=int, default=20)
    ap.add_argument("--batch-count", type=int, default=64)
    ap.add_argument("--only", choices=["headline", "laplacian", "mixed", "banded", "spmm", "batch",
                                       "adjoint", "pagerank", "all"],
                    default="all",
                    help="headline: only the FP64 R-MAT line; laplacian / mixed / banded / spmm / "
                         "batch / adjoint / pagerank: the headline plus that part (for profiling)")
    ap.add_argument("--copy-probe", action="store_true",
                    help="also run the plain device copy (counter calibration)")
    ap.add_argument("--compare", action="store_true",
                    help="also time rocSPARSE on the headline matrix (implied by --only all)")
    ap.add_argument("--no-warm", action="store_true",
                    help="skip the warm single-copy line (so a kernel trace of the run holds "
                         "only rotated-copy products)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-graph", action="store_true",
                    help="time the headline steps as eager launches instead of one graph replay")
    ap.add_argument("--no-probes", action="store_true",
                    help="skip the stream-only and gather-only probes of the headline kernel")
    ap.add_argument("--no-hot", action="store_true",
                    help="leave the hot-column split out of the tuned candidates")
    ap.add_argument("--no-sort", action="store_true",
                    help="leave tile-sorted columns out of the tuned candidates")
    ap.add_argument("--sort", action="store_true",
                    help="only tile-sorted columns for the power-law products (no hot split)")
    ap.add_argument("--hot", type=int, default=None,
                    help="hot columns of the power-law products (0 = none); default: tuned")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported "
              "(the merge-path SpMV library runs on a single MI355X)",
              file=sys.stderr)
        sys.exit(2)

    import torch

    from merbit_amd import workloads

    dev = torch.device("cuda:0")
    t0 = time.time()
    A, x = workloads.rmat(args.scale)
    log(f"rmat: nrow={A.nrow} nnz={A.nnz} ({time.time() - t0:.1f}s)")
    checks = []  # GPU results for the oracle leg: (name, A, values, x, y, bar)
    h = bench_rmat_f64(A, x, dev, args, checks)
    b_alg, ms, kern = h["b_alg"], h["ms"], h["kern"]
    value = b_alg / (ms * 1e-3) / 1e9
    achieved = b_alg / (kern * 1e-3) / 1e9
    extras = {
        "gflops": round(2 * A.nnz / (ms * 1e-3) / 1e9, 2),
        "pct_hbm_peak": round(100.0 * value / HBM_PEAK_GBS, 2),
        "algorithmic_bytes": b_alg,
        "tile_kernel_us": round(kern * 1e3, 2),
        "warm_single_copy": None if h["warm_ms"] is None else {
            "gbs": gbs(b_alg, h["warm_ms"]), "us_per_step": round(h["warm_ms"] * 1e3, 2),
            "tile_kernel_us": round(h["warm_kern"] * 1e3, 2)},
        "ntiles": h["ntiles"], "tile_items": h["tile_items"],
        "plan_create_ms": round(h["plan_ms"], 3),
        "headline_rule": h["rule"],
        "hot_columns": h["hot"],
        "sorted_columns": h["srt"],
        "ceilings": h["ceilings"],
    }
    lap = (None, None, None, 0)
    band = (None, None, None, 0)
    if args.only in ("all", "laplacian"):
        extras["laplacian"] = bench_laplacian(dev, args, checks)
        La, lx = workloads.laplacian(256)
        lap = (La, La.val, lx, 200)
    if args.only in ("all", "mixed"):
        extras["mixed_rmat"] = bench_mixed("rmat", A, x, dev, args, checks, f64_ms=ms, copies=3)
    if args.only in ("all", "spmm"):
        extras["spmm_rmat"] = bench_spmm(A, dev, args, checks)
    if args.only in ("all", "banded"):
        t0 = time.time()
        B, xb = workloads.banded()
        log(f"banded: nrow={B.nrow} nnz={B.nnz} ({time.time() - t0:.1f}s)")
        extras["mixed_banded"] = bench_mixed("banded", B, xb, dev, args, checks)
        band = (B, B.val, xb, 10)
    if args.only in ("all", "batch"):
        extras["batch_uneven"] = bench_batch(dev, args)
    if args.only in ("all", "adjoint"):
        extras["adjoint_graph"] = bench_adjoint(dev, args, checks)
        extras["adjoint_rmat"] = bench_adjoint_rmat(A, x, dev, args, checks)
    if args.only in ("all", "pagerank"):
        extras["pagerank_rmat"] = bench_pagerank(A, dev, args)
    if args.only == "all" or args.compare:
        extras["rocsparse_rmat_f64"] = bench_rocsparse(A, x, dev, args, ms)
    if args.only == "all" or args.copy_probe:
        extras["copy_ceiling"] = bench_copy(dev)

    out = {
        "metric": METRIC,
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
                   "nrow": A.nrow, "nnz": A.nnz, "items_per_thread": h["ipt"],
                   "stream_loads": h["stream"], "tiles_per_cu": h["resid"] or "all",
                   "hot_columns": h["hot"], "sorted_columns": h["srt"],
                   "wg_size": 256, "copies_rotated": max(1, args.copies),
                   "timed_as": "eager launches" if args.no_graph else "one HIP graph of the steps"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": load_traffic("rmat_f64")},
        "cpu_baseline": None,
        "extras": extras,
    }
    if not args.no_cpu:
        base, per, errs = cpu_leg(A, x, b_alg, checks, lap, band)
        out["cpu_baseline"] = base
        extras["cpu_baselines"] = per
        extras["oracle_errors"] = errs
    extras["north_star"] = north_star_summary(extras)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
