#!/usr/bin/env python3
"""Benchmark of the SBP hot path on one MI355X.

A step is one full partition() (GSAP/partition.hpp:852-954, timed like GSAP.cu:24-30) of
the 200,000-vertex directed DC-SBM graph (71 planted blocks, Dirichlet(1) sizes,
intra/inter ratio 1.5, ~4.5M edges, seed 3) with the graph already resident in HBM.

value = nodal-move edges/s = E * sweeps / sum of sweep wall time, where a sweep is the body
of partition.hpp:912-935 (4 nodal batches, each with its block-matrix rebuild, plus the
description length).  The timed steps run as a caller gets them by default, without event
timing.  The roofline object is for the largest
single kernel, the fused nodal-move kernel (one launch per batch): its average launch time
comes from HIP events on its stream in as many profiled partitions of the same workload.
The CPU baseline times the oracle on the host's cores (and on one thread) on a bounded
sample of the same workload: a few seconds of nodal batches on one labelling per
block-count regime, combined over the GPU partition's own batches per regime.
"""
import argparse
import dataclasses
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import gsap_amd  # noqa: E402
from gsap_amd import workloads  # noqa: E402

CONFIGS = {k: v for k, v in workloads.CONFIGS.items() if k != "c5"}
HBM_PEAK_GBS = 8000.0
KCLASS = ["nodal", "rebuild", "block_terms", "merge", "mdl", "accept_sum", "nodal_propose"]
# device kernels of each class (the library times a class with HIP events around all of them;
# the first matching class names a kernel)
KERNELS_OF = {
    "nodal_propose": ["k_nodal_propose"],
    "nodal": ["k_nodal"],
    "rebuild": ["k_label_", "k_tile_hist", "k_edge_hist", "k_hist_reduce", "k_dense_",
                "k_small_bm", "k_transpose", "k_logs", "k_seg_", "k_scan_", "k_lab16",
                "k_colsum"],
    "accept_sum": ["k_apply_tiles", "k_final_sum"],
    "block_terms": ["k_block_terms"],
    "merge": ["k_merge", "k_argmin", "k_order_keys", "k_radix_", "k_big_blocks", "k_jump",
              "k_relabel_roots"],
}
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "c3_fp64_traffic.json")


def kernel_class(name):
    short = name.split("(")[0].replace("void ", "").replace("gsap::", "")
    for k, pats in KERNELS_OF.items():
        if any(short.startswith(pt) for pt in pats):
            return k
    return "other"


def summarize_profile(stats_csv, fetch_csv, write_csv, cal_fetch_csv, cal_write_csv, bench_log,
                      out_md, out_json, command,
                      title="full partition of the 200k-vertex graph (FP64 form)"):
    """rocprofv3 CSVs -> a short markdown summary + per-class HBM bytes per launch."""
    import csv
    from collections import defaultdict
    stats = list(csv.DictReader(open(stats_csv)))

    def per_dispatch(path):
        rows = list(csv.DictReader(open(path)))
        agg = defaultdict(float)
        for r in rows:
            agg[r["Kernel_Name"]] += float(r["Counter_Value"])
        return agg, rows

    fetch, frows = per_dispatch(fetch_csv)
    write, _ = per_dispatch(write_csv)
    # partitions in the profiled process (one k_iota dispatch each)
    n_part = max(1, sum(1 for r in frows if "k_iota" in r["Kernel_Name"]))
    cf, crow = per_dispatch(cal_fetch_csv)
    cw, _ = per_dispatch(cal_write_csv)
    copy = [k for k in cf if "copyBuffer" in k][0]
    ncopy = sum(1 for r in crow if r["Kernel_Name"] == copy)
    copy_bytes = 512 * 1024 * 1024 * ncopy  # 3 copies of a 512 MiB buffer (bench --calibrate-copy)
    f_fac = copy_bytes / (cf[copy] * 1024)
    w_fac = copy_bytes / (cw[copy] * 1024)
    tot = sum(float(r["TotalDurationNs"]) for r in stats)
    lines = [f"# Kernel profile: {title}", "",
             f"Command: `{command}`", "",
             f"PMC calibration on {ncopy} plain 512 MiB device copies (bench.py --calibrate-copy): "
             f"bytes read = {f_fac:.3f} x FETCH_SIZE KiB x 1024, "
             f"bytes written = {w_fac:.3f} x WRITE_SIZE KiB x 1024.", "",
             "| kernel | class | calls | total ms | avg us | % | HBM read MB/call | HBM write MB/call |",
             "|---|---|---|---|---|---|---|---|"]
    cls_bytes = defaultdict(lambda: [0.0, 0.0])
    per_kernel = {}
    for r in sorted(stats, key=lambda r: -float(r["TotalDurationNs"])):
        n = r["Name"]
        c = kernel_class(n)
        calls = int(r["Calls"])
        rb = fetch.get(n, 0.0) * 1024 * f_fac
        wb = write.get(n, 0.0) * 1024 * w_fac
        cls_bytes[c][0] += rb
        cls_bytes[c][1] += wb
        per_kernel[n.split("(")[0].replace("void ", "").replace("gsap::", "")] = {
            "calls": calls, "avg_us": float(r["AverageNs"]) / 1e3,
            "read_bytes_per_call": rb / calls, "write_bytes_per_call": wb / calls}
        if float(r["TotalDurationNs"]) / tot >= 0.003:
            lines.append(f"| {n.split('(')[0][:48]} | {c} | {calls} | "
                         f"{float(r['TotalDurationNs']) / 1e6:.2f} | {float(r['AverageNs']) / 1e3:.1f} | "
                         f"{100 * float(r['TotalDurationNs']) / tot:.1f} | {rb / calls / 1e6:.2f} | "
                         f"{wb / calls / 1e6:.2f} |")
    lines += ["", f"Total kernel time {tot / 1e6:.1f} ms.", ""]
    open(out_md, "w").write("\n".join(lines) + "\n")
    bench = json.loads([l for l in open(bench_log) if l.startswith("{")][-1])
    json.dump({"command": command, "fetch_factor": f_fac, "write_factor": w_fac,
               "partitions_profiled": n_part,
               "class_launches": {k: n_part * v
                                  for k, v in bench.get("kernel_calls_per_step", {}).items()},
               "class_bytes_total": {k: v[0] + v[1] for k, v in cls_bytes.items()},
               "class_read_total": {k: v[0] for k, v in cls_bytes.items()},
               "class_write_total": {k: v[1] for k, v in cls_bytes.items()},
               "kernels": per_kernel},
              open(out_json, "w"), indent=1)


def nodal_algorithmic_bytes(evals, adj, launches):
    """Algorithmic HBM bytes of one launch of the nodal moves kernel: per vertex it
    evaluates (a proposal of another block) its list entry, row pointers, label, proposal
    and delta-S (4 + 16 + 4 + 4 + 8 = 36 bytes), per out/in adjacency entry of those
    vertices the neighbour id, the weight and the neighbour's block as a 16-bit label
    (4 + 4 + 2 = 10 bytes); evals and adj are totals over `launches` launches."""
    return (36 * evals + 10 * adj) / max(launches, 1)


def cpu_threads():
    """Host threads of the all-cores oracle: the box's share (OMP_NUM_THREADS), at most 16."""
    n = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
    return max(1, min(16, n))


# block-count regimes of a partition's nodal batches (the rebuild and lookup forms differ)
REGIMES = [(16385, 1 << 30), (8193, 16384), (513, 8192), (129, 512), (1, 128)]


def regime_of(B):
    return next(k for k, (lo, hi) in enumerate(REGIMES) if lo <= B <= hi)


def batch_mix(ctx, g, thr):
    """Nodal batches per block-count regime in one traced partition of the workload (every
    batch record of the trace carries its B)."""
    ctx.partition(g, gsap_amd.Options(seed=1, trace=1, **thr))
    n = [0] * len(REGIMES)
    for rec in ctx.trace():
        if rec[0] == 2:
            n[regime_of(rec[1])] += 1
    return n


def regime_labellings(ctx, g, final_lab, final_B):
    """A labelling per regime: above 128 blocks the first block count of the regime that the
    merge phases alone reach from one block per vertex (on the GPU, the partition's merge
    kernels and reduction rate, no sweeps in between); at most 128 blocks the partition's
    final labelling."""
    o = gsap_amd.Options(seed=1)
    lab, B, m = np.arange(g.N, dtype=np.int32), g.N, 0
    out = {regime_of(final_B): (final_lab, final_B)}
    while B > 128:
        r = ctx.merge_phase(g, o, lab, B, int(np.float32(B) * np.float32(o.reduction_rate)), m)
        lab, B, m = r["labels"], r["B"], m + 1
        out.setdefault(regime_of(B), (lab.copy(), B))
    return out


def cpu_baseline(g, labs, mix, seconds, threads):
    """The oracle's nodal batches (each with its block-matrix rebuild), seconds / 5 on the
    labelling of each block-count regime, on `threads` host threads; the regime rates are
    combined over the GPU partition's own batch mix (the time the oracle would take for those
    batches), so CPU and GPU values cover the same work."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    opts = gsap_amd.Options(seed=1)
    per_batch, parts = {}, []
    for k, (lab, B) in sorted(labs.items()):
        t0 = time.perf_counter()
        nbat = 0
        while True:
            if threads == 1:
                O.nodal_batch(g, opts, lab, B, nbat % opts.num_batches, 100000 + nbat)
            else:
                O.nodal_batch_mt(g, opts, lab, B, nbat % opts.num_batches, 100000 + nbat, threads)
            nbat += 1
            el = time.perf_counter() - t0
            if el >= seconds / len(labs) or nbat >= 256:
                break
        per_batch[k] = el / nbat
        parts.append(f"B={B}: {nbat} batches in {el:.1f} s")
    edges_per_batch = g.E / opts.num_batches
    t_cpu = sum(mix[k] * per_batch[k] for k in per_batch)
    n_all = sum(mix[k] for k in per_batch)
    return dict(value=n_all * edges_per_batch / t_cpu, unit="edges/s", cores=threads,
                kind="port",
                sample=f"nodal batches (each with its block-matrix rebuild) of the CPU oracle on "
                       f"{threads} thread(s), same 200k-vertex graph, one labelling per "
                       f"block-count regime ({'; '.join(parts)}), combined over the GPU "
                       f"partition's batches per regime {dict(zip(map(str, REGIMES), mix))}",
                regime_edges_per_s={str(REGIMES[k]): edges_per_batch / per_batch[k]
                                    for k in per_batch})


def compare_forms(ctx, g, lab, prec, thr):
    """Plain FP64 form against the mixed form (FP32 delta-S with FP64 for accepted and
    near-threshold proposals): one untimed and one timed partition of the other form, no
    event profiling in either timed run."""
    other = gsap_amd.PRECISION_MIXED if prec == gsap_amd.PRECISION_FP64 else gsap_amd.PRECISION_FP64
    walls, res = {}, {}
    for form in (prec, other):
        if form == other:
            ctx.partition(g, gsap_amd.Options(seed=1, precision=form, **thr))
        t = time.perf_counter()
        res[form] = ctx.partition(g, gsap_amd.Options(seed=1, precision=form, **thr))
        walls[form] = time.perf_counter() - t
        print(f"form {'fp64' if form == 0 else 'mixed'}: {walls[form]:.3f} s", file=sys.stderr,
              flush=True)
    F, M = gsap_amd.PRECISION_FP64, gsap_amd.PRECISION_MIXED
    (l64, r64), (l32, r32) = res[F], res[M]
    return {"fp64_s": walls[F], "mixed_s": walls[M], "mixed_over_fp64": walls[M] / walls[F],
            "identical_partition": bool(np.array_equal(l64, l32)),
            "identical_S": bool(np.float64(r64.S).tobytes() == np.float64(r32.S).tobytes()),
            "mixed_max_abs_err_fp32": r32.max_abs_err_fp32, "mixed_rechecks": r32.rechecks,
            "mixed_fp64_evals": r32.fp64_evals,
            "nodal_proposals": int(r32.batches * (g.N // 4))}


def bench_multi(a, prec):
    """C4 (batched graphs) and C5 (parallel chains): aggregate nodal-move edges/s over the
    wall time of the whole call."""
    t0 = time.perf_counter()
    if a.config == "c4":
        graphs = workloads.c4_graphs()
        desc = (f"batched: {len(graphs)} DC-SBM graphs, N log-uniform [500, 50000] "
                f"(sum N={sum(g.N for g in graphs)}, sum E={sum(g.E for g in graphs)}), "
                f"K=round(N^(1/3)), rho 4/1.5, seeds 4+g")
    else:
        g5 = workloads.graph("c5")
        graphs = [g5]
        desc = (f"chains: {a.chains} chains on N={g5.N}, E={g5.E}, K=150, alpha=0.5, rho=2, "
                f"seed 5; lowest S kept")
    t_gen = time.perf_counter() - t0
    ctx = gsap_amd.Context(0)
    opts = gsap_amd.Options(seed=1, precision=prec, workers=a.workers, graph=int(a.graph))
    thr, t_tune = tune(a, ctx, max(graphs, key=lambda x: x.E), opts)
    opts = dataclasses.replace(opts, **thr)
    walls, edges = [], []
    labs = None
    for k in range(a.warmup + a.steps):
        t1 = time.perf_counter()
        if a.config == "c4":
            labs, reps = ctx.partition_batch(graphs, opts)
            e = sum(r.sweep_edges for r in reps)
            extra = {}
        else:
            _, rep, S, best = ctx.partition_chains(graphs[0], a.chains, opts)
            e = None
            extra = {"best_chain": best, "S_best": rep.S, "S_chains": [float(x) for x in S]}
        w = time.perf_counter() - t1
        print(f"{'warmup' if k < a.warmup else 'step'} {k}: {w:.2f} s", file=sys.stderr,
              flush=True)
        if k >= a.warmup:
            walls.append(w)
            edges.append(e)
    if a.config == "c4":
        # the nodal-move kernel over every graph of the batch: HIP events on each worker's
        # stream around each of its launches (kernels of other graphs run beside it), in one
        # profiled batch call of the same graphs
        plabs, preps = ctx.partition_batch(graphs, dataclasses.replace(opts, profile=1,
                                                                        profile_classes=1))
        assert all(np.array_equal(x, y) for x, y in zip(plabs, labs)), "profiled batch differs"
        n_launch = sum(r.kernel_calls[0] for r in preps)
        nodal_ms = sum(r.kernel_ms[0] for r in preps)
        evals = sum(r.nodal_evals for r in preps)
        adj = sum(r.nodal_eval_adj for r in preps)
        abytes = nodal_algorithmic_bytes(evals, adj, n_launch)
        avg_ms = nodal_ms / max(n_launch, 1)
        achieved = abytes / (avg_ms * 1e-3) / 1e9
        extra["roofline"] = {"bound": "hbm", "kernel": "k_nodal_fused5 (all graphs of the batch)",
                             "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                             "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                             "avg_launch_ms": avg_ms, "launches": n_launch,
                             "algorithmic_bytes_per_launch": abytes,
                             "evaluated_vertices_per_launch": evals / max(n_launch, 1)}
        extra["nodal_kernel_ms_total"] = nodal_ms
    ctx.close()
    out = {"metric": "aggregate nodal-move edges/s over the whole call" if a.config == "c4"
                     else "partition seconds of the parallel-chain run",
           "value": (sum(edges) / sum(walls)) if a.config == "c4" else float(np.mean(walls)),
           "unit": "edges/s" if a.config == "c4" else "s", "n_gpus": 1, "steps": a.steps,
           "warmup": a.warmup, "ms_per_step": 1e3 * float(np.mean(walls)),
           "higher_is_better": a.config == "c4", "scaling": "weak", "vs_baseline": None,
           "dtype": "f64", "data": "synthetic (seeded DC-SBM generator)",
           "config": {"workload": desc, "workers": a.workers or 8,
                      "sweep_launch": "hip graph replay" if a.graph else "kernel by kernel",
                      "thresholds": thr or "defaults"},
           "graph_generation_s": t_gen, "autotune_s": t_tune}
    out.update(extra)
    print(json.dumps(out))


def mixed_report():
    """The mixed form against the plain FP64 form on every configuration: identical labels
    and S, max |dS32 - dS64| over the proposals evaluated in FP64, the share of proposals
    whose FP32 delta-S lay within its error bound (rechecks) and the wall times; for the
    200k-vertex graph also per level (cumulative counters of partitions cut after k levels,
    differenced)."""
    F, M = gsap_amd.PRECISION_FP64, gsap_amd.PRECISION_MIXED
    out = {"configs": {}}
    ctx = gsap_amd.Context(0)

    def one(name, graphs, chains=None):
        res = {}
        for prec in (F, M):
            o = gsap_amd.Options(seed=1, precision=prec)
            if name == "c4":
                ctx.partition_batch(graphs, o)  # warm-up
                t = time.perf_counter()
                labs, reps = ctx.partition_batch(graphs, o)
            else:
                ctx.partition(graphs[0], o)
                t = time.perf_counter()
                lab, rep = ctx.partition(graphs[0], o)
                labs, reps = [lab], [rep]
            res[prec] = (time.perf_counter() - t, labs, reps)
        (tf, lf, rf), (tm, lm, rm) = res[F], res[M]
        props = sum(r.batches * (g.N // 4) for r, g in zip(rm, graphs))
        d = {"fp64_s": tf, "mixed_s": tm, "mixed_over_fp64": tm / tf,
             "identical": all(np.array_equal(a, b) for a, b in zip(lf, lm)) and
             all(x.S == y.S for x, y in zip(rf, rm)),
             "max_abs_err_fp32": max(r.max_abs_err_fp32 for r in rm),
             "S": [r.S for r in rf][:4],
             "rechecks": sum(r.rechecks for r in rm), "fp64_evals": sum(r.fp64_evals for r in rm),
             "nodal_proposals": props}
        d["recheck_share"] = d["rechecks"] / max(props, 1)
        print(f"{name}: {json.dumps(d)}", file=sys.stderr, flush=True)
        return d

    for name in ("c1", "c2", "c3"):
        out["configs"][name] = one(name, [workloads.graph(name)])
    out["configs"]["c4"] = one("c4", workloads.c4_graphs())
    out["configs"]["c5"] = one("c5", [workloads.graph("c5")])
    # per level of the 200k-vertex graph
    g = workloads.graph("c3")
    prev = (0, 0, 0)
    levels = []
    k = 1
    while True:
        _, rep = ctx.partition(g, gsap_amd.Options(seed=1, precision=M, max_levels=k))
        cur = (rep.batches, rep.rechecks, rep.fp64_evals)
        nb = cur[0] - prev[0]
        if nb > 0:
            props = nb * (g.N // 4)
            levels.append({"level": k - 1, "batches": nb,
                           "recheck_share": (cur[1] - prev[1]) / props,
                           "fp64_share": (cur[2] - prev[2]) / props})
        prev = cur
        if rep.levels < k:
            break
        k += 1
    # the block count of each level from the FP64 trace
    _, _ = ctx.partition(g, gsap_amd.Options(seed=1, trace=1))
    gold = [t for t in ctx.trace() if t[0] == 1]
    for lv in levels:
        if lv["level"] < len(gold):
            lv["B"] = gold[lv["level"]][2]
    out["c3_levels"] = levels
    ctx.close()
    print(json.dumps(out))


def tune(a, ctx, g, opts):
    """Thresholds measured on g (untimed) when --autotune is given, else none."""
    if not a.autotune:
        return {}, None
    t0 = time.perf_counter()
    o = ctx.autotune(g, opts)
    t = time.perf_counter() - t0
    thr = {f: getattr(o, f) for f in gsap_amd.TUNED}
    print(f"autotune: {thr} in {t:.2f} s", file=sys.stderr, flush=True)
    return thr, t


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="c3", choices=list(CONFIGS) + ["c4", "c5"])
    ap.add_argument("--chains", type=int, default=16)
    ap.add_argument("--no-compare", action="store_true",
                    help="skip the FP64-vs-mixed comparison run")
    ap.add_argument("--workers", type=int, default=0)
    ap.add_argument("--graph", action="store_true",
                    help="replay each level's captured sweep as a HIP graph")
    ap.add_argument("--precision", default="fp64", choices=["fp64", "mixed"])
    ap.add_argument("--autotune", action="store_true",
                    help="measure the row-tile and dense-lookup thresholds on the workload "
                         "graph first (untimed) and run with them")
    ap.add_argument("--cpu-seconds", type=float, default=12.0,
                    help="bounded CPU-oracle sample per thread count (two samples)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a plain 512 MiB device copy (PMC byte-counter calibration)")
    ap.add_argument("--summarize-profile", nargs=8,
                    metavar=("STATS", "FETCH", "WRITE", "CALF", "CALW", "BENCH_LOG", "OUT_MD",
                             "OUT_JSON"),
                    help="turn rocprofv3 CSVs into the profiles/ summary")
    ap.add_argument("--profile-command", default="")
    ap.add_argument("--profile-title", default="full partition of the 200k-vertex graph (FP64 form)")
    ap.add_argument("--mixed-report", action="store_true",
                    help="mixed against FP64 form on C1-C5: identity, FP32 error, rechecks")
    a = ap.parse_args()
    if a.mixed_report:
        mixed_report()
        return
    if a.summarize_profile:
        summarize_profile(*a.summarize_profile, command=a.profile_command, title=a.profile_title)
        return
    if a.calibrate_copy:
        import torch
        x = torch.ones(128 << 20, dtype=torch.int32, device="cuda")
        y = torch.empty_like(x)
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": x.numel() * 4, "copy_bytes_written": y.numel() * 4,
                          "launches": 3}))
        return
    if a.gpus != 1:
        print(f"bench.py runs on one GPU: this build partitions one graph per GPU and has no "
              f"multi-GPU path (asked for --gpus {a.gpus})", file=sys.stderr)
        sys.exit(2)
    prec = gsap_amd.PRECISION_FP64 if a.precision == "fp64" else gsap_amd.PRECISION_MIXED
    if a.config in ("c4", "c5"):
        bench_multi(a, prec)
        return
    cfg = CONFIGS[a.config]
    t0 = time.perf_counter()
    g = workloads.graph(a.config)
    t_gen = time.perf_counter() - t0
    ctx = gsap_amd.Context(0)
    thr, t_tune = tune(a, ctx, g, gsap_amd.Options(seed=1, precision=prec))

    def opt(**kw):
        return gsap_amd.Options(seed=1, precision=prec, **thr, **kw)

    # untimed: warmups, the last one (or one extra run) with HIP events around every kernel
    # class, to pick the dominant class and report the per-class split
    prep = None
    for w in range(max(a.warmup, 1)):
        last = w == max(a.warmup, 1) - 1
        _, r0 = ctx.partition(g, opt(profile=int(last)))
        if last:
            prep = r0
        print(f"warmup {w} done", file=sys.stderr, flush=True)
    kms = np.asarray(prep.kernel_ms)
    kcalls = np.asarray(prep.kernel_calls)
    # untimed: one partition with the graph handed over in host memory (its upload over PCIe
    # inside the call), reported beside the timed value
    t1 = time.perf_counter()
    _, rup = ctx.partition(g, opt())
    wall_upload = time.perf_counter() - t1
    # timed steps on the graph the calls above left resident in device memory: no event
    # timing (what a caller gets by default), no host-to-device copy of the graph
    walls, reps = [], []
    for k in range(a.steps):
        t1 = time.perf_counter()
        lab, rep = ctx.partition_resident(g.N, opt())
        walls.append(time.perf_counter() - t1)
        reps.append(rep)
        print(f"step {k}: {walls[-1]:.3f} s, B={rep.B}, sweeps={rep.sweeps}", file=sys.stderr,
              flush=True)
    sweep_edges = sum(r.sweep_edges for r in reps)
    t_sweep = sum(r.t_sweep for r in reps)
    value = sweep_edges / t_sweep
    # roofline of the largest single kernel, the fused nodal-move kernel (one launch per
    # batch): HIP events on the work stream around each of its launches, in as many profiled
    # partitions of the same workload
    dom = 0
    pwalls, preps = [], []
    for k in range(a.steps):
        t1 = time.perf_counter()
        plab, prp = ctx.partition_resident(g.N, opt(profile=1, profile_classes=1 << dom))
        pwalls.append(time.perf_counter() - t1)
        preps.append(prp)
        assert np.array_equal(plab, lab), "profiled partition differs"
    avg_ms = sum(r.kernel_ms[dom] for r in preps) / max(sum(r.kernel_calls[dom] for r in preps), 1)
    n_launch = sum(r.kernel_calls[dom] for r in preps)
    evals = sum(r.nodal_evals for r in preps)
    adj = sum(r.nodal_eval_adj for r in preps)
    assert evals >= 0 and adj >= 0, "evaluation counts missing (unit-weight graph expected)"
    abytes = nodal_algorithmic_bytes(evals, adj, n_launch)
    achieved = abytes / (avg_ms * 1e-3) / 1e9
    traffic = None
    if os.path.exists(TRAFFIC_FILE):
        # the default-variant instance of the nodal kernel (reference delta, greedy accept)
        want = ("k_nodal_fused<true, 0>" if prec == gsap_amd.PRECISION_MIXED
                else "k_nodal_fused5<false, 0>")
        kn = json.load(open(TRAFFIC_FILE)).get("kernels", {}).get(want)
        if kn:
            traffic = kn["read_bytes_per_call"] + kn["write_bytes_per_call"]
    p, r = gsap_amd.pairwise(g.truth, lab)
    qual = gsap_amd.quality(g.truth, lab)
    out = {
        "metric": "nodal-move edges/s (full partition of the 200k-vertex graph); full-partition "
                  "seconds and % HBM peak reported alongside",
        "value": value,
        "unit": "edges/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": 1e3 * float(np.mean(walls)),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if prec == gsap_amd.PRECISION_FP64 else "f32+f64",
        "data": "synthetic (seeded DC-SBM generator)",
        "config": {"workload": f"partition {a.config}: N={g.N}, E={g.E}, K={cfg['K']}, "
                               f"alpha={cfg['alpha']}, rho={cfg['rho']}, seed={cfg['seed']}",
                   "precision": a.precision, "parallelism": "single GPU",
                   "thresholds": thr or "defaults"},
        "autotune_s": t_tune,
        "full_partition_s": float(np.mean(walls)),
        "profiled_partition_s": float(np.mean(pwalls)),
        "final_B": rep.B, "final_S": rep.S, "sweeps": rep.sweeps, "levels": rep.levels,
        "pairwise_precision": p, "pairwise_recall": r,
        "quality": {k: qual[k] for k in ("accuracy", "nmi", "adjusted_rand_index", "B_truth",
                                         "B_given")},
        "t_merge_s": float(np.mean([x.t_merge for x in reps])),
        "t_sweep_s": float(np.mean([x.t_sweep for x in reps])),
        "t_prepare_s": float(np.mean([x.t_upload for x in reps])),
        "full_partition_with_upload_s": wall_upload,
        "t_upload_s": rup.t_upload,
        "graph_generation_s": t_gen,
        "kernel_ms_per_step": {KCLASS[i]: float(kms[i]) for i in range(len(KCLASS))},
        "kernel_calls_per_step": {KCLASS[i]: int(kcalls[i]) for i in range(len(KCLASS))},
        "roofline": {"bound": "hbm", "kernel": "k_nodal_fused5" if prec != gsap_amd.PRECISION_MIXED else "k_nodal_fused", "achieved": achieved,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": traffic, "avg_launch_ms": avg_ms,
                     "algorithmic_bytes_per_launch": abytes,
                     "evaluated_vertices_per_launch": evals / max(n_launch, 1),
                     "evaluated_adjacency_per_launch": adj / max(n_launch, 1)},
    }
    if not a.no_compare:
        out["precision_forms"] = compare_forms(ctx, g, lab, prec, thr)
    if not a.no_cpu_baseline:
        nt = cpu_threads()
        mix = batch_mix(ctx, g, thr)
        labs = regime_labellings(ctx, g, lab, rep.B)
        out["cpu_baseline"] = cpu_baseline(g, labs, mix, a.cpu_seconds, nt)
        out["cpu_baseline_1thread"] = cpu_baseline(g, labs, mix, a.cpu_seconds, 1)
    ctx.close()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
